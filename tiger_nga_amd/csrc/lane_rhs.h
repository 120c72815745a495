// Lane-per-state chemistry right-hand side: one lane evaluates
// finitechem_compute_rhs (reference combustion/finitechem.f90:93-175) for a
// whole state.  The specialisations LaneRhs<NEQ> are generated per mechanism
// (lane_rhs_<mech>.h, tiger_nga_amd/mech/lanegen.py) as straight-line code
// with compile-time species indices and immediate coefficients.
#pragma once
#include "chem_device.h"

namespace tgr {

template <int NEQ>
struct LaneRhs;

}  // namespace tgr

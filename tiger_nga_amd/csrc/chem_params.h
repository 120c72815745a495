// Parameters of one chemistry step and per-cell status codes, shared by the
// HIP kernels, the C-ABI layer and the CPU wave emulator.
#pragma once
#include <stdint.h>

#include "chem_device.h"

// per-cell integrator status (same codes as the CPU oracle)
#define TGR_ST_OK 0
#define TGR_ST_TOO_MUCH_WORK (-1)
#define TGR_ST_TOO_MUCH_ACC (-2)
#define TGR_ST_ERR_FAIL (-3)
#define TGR_ST_CONV_FAIL (-4)
#define TGR_ST_SETUP_FAIL (-5)
#define TGR_ST_BAD_INPUT (-6)

// counters per cell: nst nfe nfe_dq nje nsetups nni netf ncfn nlu status
#define TGR_NCOUNT 10

namespace tgr {

// One chemistry step over the interior cells of a Fortran-layout block.
// Cell c in [0, nx*ny*nz) maps to (i, j, k) with i fastest (the reference's
// k-j-i loop order, finitechem.f90:993-995); its element offset in RHO_/SC_
// planes is base + i*sx + j*sy + k*sz.
struct StepParams {
  DevMech m;
  double dt, p_gp, rtol, atol;
  int max_steps, qmax;
  int nx, ny, nz;
  int64_t sx, sy, sz, base, plane;
  int isc_sc0, isc_T0;  // 0-based scalar planes of the first species and of T
  const int* mask;      // mask(imino:imaxo, jmino:jmaxo) or null
  int64_t mask_ld, mask_base;
  double* rho;
  double* sc;
  double* hr;       // HR_gp or null
  double* chemsrc;  // chemsrc_gp(..., 3) or null
  int idx_h2o, idx_co, idx_co2;
  int64_t ncell;
  unsigned int* queue;           // device work counter
  double* ws;                    // per-block workspace: saved Jacobian, neq * 64 doubles per block
  int* cell_stats;               // [ncell][TGR_NCOUNT] or null
  unsigned long long* totals;    // [TGR_NCOUNT] (last: failed cells) or null
};

// Transport / equation-of-state properties of a whole Fortran-layout block.
// ptab: mucoeff[64], kOverEps[64], W[64], then A[nsp][64] and B[nsp][64]
// (pair factors, row j, lane k).
struct PropsParams {
  const double* ptab;
  const char* mtab;  // mechanism table block (thermo rows)
  double tmid, p_thermo;
  int nx, ny, nz;    // whole block, ghosts included
  int64_t plane;
  int isc_sc0, isc_T0, isc_z0;  // 0-based scalar planes; isc_z0 < 0: none
  const int* mask;
  int64_t mask_ld, mask_base;
  const double* lewis;  // [nsp] or null (all 1)
  const double* sc;
  double* rho;   // RHO_ or null
  double* visc;  // VISC_ or null
  double* diff;  // DIFF_(..., nscalar) or null
};

}  // namespace tgr

// Launch interface between the C-ABI layer (chem_capi.cpp) and the HIP
// kernels (chem_bdf.hip).  Internal to the library.
#pragma once
#include <hip/hip_runtime.h>
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

}  // namespace tgr

extern "C" {
int tgr_step_occupancy(int neq);
int tgr_launch_step(const tgr::StepParams* p, int neq, int grid, hipStream_t s);
int tgr_launch_prodrates(const tgr::DevMech* m, int64_t n, const double* C, const double* T, const double* P,
                         double* cdot, hipStream_t s);
int tgr_launch_rhs(const tgr::DevMech* m, int64_t n, double p_gp, const double* y, double* f, hipStream_t s);
}

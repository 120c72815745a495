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
  int jac;              // 0: difference-quotient Jacobian (CVODE's, as the reference); 1: analytic
  int waves;            // wave-per-cell step kernel: resident waves per SIMD (2; 4 for 10 equations)
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
  const int* cell_list;          // cells to integrate (indices into [0, ncell)) or null: all
  const unsigned int* nlist;     // with cell_list: device count of listed cells
  unsigned int* queue;           // device work counter
  double* ws;                    // per-block workspace: saved Jacobian, neq * 64 doubles per block
  int* cell_stats;               // [ncell][TGR_NCOUNT] or null
  unsigned long long* totals;    // [TGR_NCOUNT] (last: failed cells) or null
};

// Surrogate-guided step: every interior cell gets a one-step SDIRK solve
// whose Newton iterations start from the network's predictions; cells that
// fail its checks are listed for the plain integrator.
constexpr int kSurrMaxSub = 8;   // sub-steps of the guided solve
constexpr int kSurrStages = 3;   // stages per sub-step
struct SurrParams {
  StepParams s;                  // field addressing, tolerances, queue, workspace
  const float* guess;            // network outputs, stage j of sub-step i at guess + (i*3 + j) * guess_stride
  int64_t guess_stride;          // rows * guess_ld
  int64_t guess_ld;
  const double* out_scale;       // [neq]: increment of Y_k / T per unit network output
  const double* elem_w;          // [TGR_NELEM][64]: atoms of element e per molecule of species k / W_k
  int nsub;                      // sub-steps (1..kSurrMaxSub)
  double hsub[kSurrMaxSub];      // their sizes (sum dt)
  int use_guess;                 // 0: Newton starts from the previous stage value
  int adaptive;                  // 1: step-size control (sdirk_adaptive) instead of the fixed schedule
  int max_att;                   // adaptive: attempted steps before the cell is rejected
  double h0_fac;                 // adaptive: first step c / max_k |f0_k| wacc_k in [dt/1024, dt] (0: hsub[0])
  int ntau;                      // number of network output times
  int split_jac;                 // 1: Jacobian with the rate constants evaluated once (Bdf::jac_split)
  double grow;                   // adaptive: largest step growth factor
  double keep;                   // adaptive: growth factors in [1, keep) keep h (and the LU factors)
  double jac_theta;              // adaptive: Newton contraction rate above which the Jacobian is re-evaluated
  int max_jac;                   // adaptive: Jacobians per cell (1: the initial one only)
  double tau[kSurrStages * kSurrMaxSub];  // times of the network outputs, increasing
  int newton_max;
  double newton_tol;             // Newton stop: |correction| <= newton_tol x acceptance tolerance
  double tol_y, tol_t;           // acceptance: |error of Y_k| <= tol_y, |error of T| <= tol_t K
  unsigned int* rej_count;       // cells handed to the plain integrator
  int* rej_list;
  unsigned long long* stotals;   // [kSurrNCount]
};
// surrogate counters: accepted, rejected, Newton iterations, RHS evaluations,
// rejected for Newton failure, for the error estimate, for the enthalpy check
constexpr int kSurrNCount = 10;

// Geometry of a ghosted block for the diffusion stencils (Fortran indices as
// in tgr_grid; plane = nxo * nyo * nzo; 0-based scalar planes).
struct DiffGeom {
  int imin_, imax_, jmin_, jmax_, kmin_, kmax_;
  int imino_, jmino_, kmino_, nxo, nyo, nzo;
  int imino, jmino, mx;  // mask(imino:, jmino:), leading dimension mx
  int64_t plane;
  int isc_sc0, isc_T0;
};

// Metric operator arrays (device pointers) of tgr_metric.
struct tgr_metric_dev {
  int st1, st2;
  const double *interp_sc_x, *interp_sc_y, *interp_sc_z;
  const double *grad_x, *grad_y, *grad_z;
  const double *div_u, *div_v, *div_w;
  const double *interp_u_xm, *interp_v_ym, *interp_w_zm;
  const double *grad_xm, *grad_ym, *grad_zm;
};

// One launch of the three kernels (filled by the C-ABI layer).
struct DiffParams {
  DiffGeom gm;
  tgr_metric_dev m;
  DevMech mech;
  double dt;
  int molar, gasrad, cv;
  const double* rho;
  const double* sc;
  const double* diff;
  const int* mask;
  double* wg;     // [plane] W_gp
  double* cpm;    // [plane] cp_mix (cv)
  double* cvel;   // [3][plane] corr_vel_x, _y, _z
  double* corr;   // diff_corr(..., nsp + 1) or null
  double* dgp;    // diff_gp(..., 3) or null
  double* src;    // the transport source (..., nscalar) or null
  int i_h2o, i_co, i_co2, i_ch4;  // species indices (-1: absent)
  double w_h2o, w_co, w_co2, w_ch4;
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

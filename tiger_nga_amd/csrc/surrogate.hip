// Surrogate-guided chemistry step (gfx950): network features, the guided
// one-step solve per cell, and its launchers.
//
// The step does what finitechem_source_chemistry does (reference
// combustion/finitechem.f90:996-1058: pack, integrate to dt, unpack, heat
// release and H2O/CO/CO2 sources) but integrates a cell by nsub steps of an
// L-stable SDIRK method whose Newton iterations start from the perceptron's
// prediction of the state at each stage time.  A cell is accepted when every
// Newton solve converges, the embedded error estimate is within the
// acceptance tolerance, the species are projected back onto exact
// conservation of every element's mass, and the temperature implied by
// enthalpy conservation
// (the reference's energy equation, finitechem.f90:158-166, conserves the
// mixture enthalpy exactly) agrees with the solved temperature; the written
// temperature is the enthalpy-consistent one.  Every other cell is appended to
// a device list that the plain integrator (step_kernel) then processes, so
// the plain FP64 solver is both the fallback and the reference the
// acceptance thresholds are measured against.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bdf_wave.h"
#include "chem_kernels.h"
#include "mech_tables.h"

namespace tgr {

constexpr int kSurrWaves = 8;
constexpr int kSurrBlock = 64 * kSurrWaves;

__device__ __forceinline__ uint16_t f2bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ int64_t cell_offset(const StepParams& p, int64_t cell, int& i, int& j) {
  i = (int)(cell % p.nx);
  j = (int)((cell / p.nx) % p.ny);
  const int k = (int)(cell / ((int64_t)p.nx * p.ny));
  return p.base + i * p.sx + j * p.sy + k * p.sz;
}

// Network input of each cell, one lane per cell: (log10 max(Y_k, 1e-30) -
// shift_k) * scale_k for the species, (T - shift) * scale, (log10 dt - shift)
// * scale, as bfloat16, row c of X[rows][in_pad] (padded columns and rows
// zero).
template <int NEQ>
__global__ __launch_bounds__(256) void surr_features_kernel(StepParams p, int64_t rows, int in_pad,
                                                            const float* shift, const float* scale, float log_dt,
                                                            uint16_t* X) {
  constexpr int nsp = NEQ - 1;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= rows) return;
  uint16_t* x = X + c * in_pad;
  if (c >= p.ncell) {
    for (int q = 0; q < in_pad; ++q) x[q] = 0;
    return;
  }
  int i, j;
  const int64_t off = cell_offset(p, c, i, j);
#pragma unroll
  for (int k = 0; k < nsp; ++k) {
    const float y = (float)p.sc[(int64_t)(p.isc_sc0 + k) * p.plane + off];
    x[k] = f2bf16((log10f(fmaxf(y, 1.0e-30f)) - shift[k]) * scale[k]);
  }
  x[nsp] = f2bf16(((float)p.sc[(int64_t)p.isc_T0 * p.plane + off] - shift[nsp]) * scale[nsp]);
  x[nsp + 1] = f2bf16((log_dt - shift[nsp + 1]) * scale[nsp + 1]);
  for (int q = nsp + 2; q < in_pad; ++q) x[q] = 0;
}

// the time-step column of every row (network inputs for another stage time)
__global__ __launch_bounds__(256) void surr_set_column_kernel(uint16_t* X, int64_t rows, int in_pad, int col,
                                                              float value) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c < rows) X[c * in_pad + col] = f2bf16(value);
}

__global__ void surr_reset_kernel(unsigned int* queue, unsigned int* rej_count, unsigned long long* stotals) {
  if (threadIdx.x == 0) {
    *queue = 0u;
    *rej_count = 0u;
  }
  if (threadIdx.x < kSurrNCount) stotals[threadIdx.x] = 0ull;
}

// One cell: -1 wall cell (skipped), 0 accepted (fields written), 1..3
// rejected (Newton, error estimate, enthalpy check; fields untouched).
// AD: the error-controlled schedule; NET: the stages start from the
// network's guesses (one kernel instance per combination keeps the other
// paths' code out of the register allocation)
template <int NEQ, bool AD, bool NET, bool AJ>
__device__ __forceinline__ int surr_cell(const SurrParams& P, const DevMech& m, tchar* tab, TGR_LDS RhsLds<NEQ>* rl,
                                         TGR_LDS RateCache<NEQ>* rc, int64_t cell, double* ws,
                                         int& nit, int& nfe, int& nlu, int& nsub, int& nje) {
  const StepParams& p = P.s;
  const int lane = lane_id();
  const int nsp = NEQ - 1;
  int i, j;
  const int64_t off = cell_offset(p, cell, i, j);
  if (p.mask && p.mask[p.mask_base + (int64_t)j * p.mask_ld + i] == 1) return -1;
  const double rho = p.rho[off];
  double sol = 0.0;
  if (lane < nsp) sol = fmax(rho * p.sc[(int64_t)(p.isc_sc0 + lane) * p.plane + off], 0.0);
  else if (lane == nsp) sol = rho * p.sc[(int64_t)p.isc_T0 * p.plane + off];
  const double sol0 = sol;
  if (__ballot(!isfinite(sol) || !(rho > 0.0)) != 0ull) return 1;
  // stage guesses in the packed variables: rho*Y + rho*dY, rho*T + rho*dT,
  // loaded when a stage starts
  const double osc = lane < NEQ ? P.out_scale[lane] : 0.0;
  const float* gcell = P.guess + cell * P.guess_ld + (lane < NEQ ? lane : 0);
  auto guess = [&](int i, int j) -> double {
    if (lane >= NEQ) return 0.0;
    return sol0 + rho * osc * (double)gcell[(int64_t)(i * kSurrStages + j) * P.guess_stride];
  };
  const double wacc = lane < nsp ? 1.0 / (P.tol_y * rho) : (lane == nsp ? 1.0 / (P.tol_t * rho) : 0.0);
  Bdf<NEQ, AJ> b(m, tab, rl, ws, p.p_gp, p.rtol, p.atol);
  if (P.split_jac && !AJ) b.rc = rc;
  int st;
  if constexpr (AD) {
    // the network's state at time t: linear in t between its stage times
    // (increment 0 at t = 0, the last output beyond the last time)
    const int nt = P.ntau;
    auto gat = [&](double t) -> double {
      if (lane >= NEQ || !(t > 0.0)) return sol0;
      int k = 0;
      while (k < nt && P.tau[k] < t) ++k;
      const double t1 = k < nt ? P.tau[k] : P.tau[nt - 1];
      const double v1 = (double)gcell[(int64_t)(k < nt ? k : nt - 1) * P.guess_stride];
      const double t0 = k > 0 ? P.tau[k - 1] : 0.0;
      const double v0 = k > 0 ? (double)gcell[(int64_t)(k - 1) * P.guess_stride] : 0.0;
      const double w = (k < nt && t1 > t0) ? (t - t0) / (t1 - t0) : 1.0;
      return sol0 + rho * osc * (v0 + w * (v1 - v0));
    };
    st = b.sdirk_adaptive(sol, p.dt, P.hsub[0], P.h0_fac, gat, NET, wacc, P.newton_max, P.newton_tol, P.max_att,
                          P.grow, P.keep, P.jac_theta, P.max_jac, nit);
  } else {
    st = b.sdirk_guided(sol, P.nsub, P.hsub, guess, NET, wacc, P.newton_max, P.newton_tol, nit);
  }
  nfe = b.c_nfe + b.c_nfe_dq;
  nlu = b.c_nlu;
  nsub = b.nst;
  nje = b.c_nje;
  if (st) return st;
  // element-conserving projection: the least change of the species (weights
  // |rho Y_k|) that restores every element's mass to its initial value
  // exactly; Newton with a difference-quotient Jacobian conserves elements
  // only to its convergence tolerance
  {
    double a[TGR_NELEM], r[TGR_NELEM], G[TGR_NELEM][TGR_NELEM];
    const double wk = lane < nsp ? fabs(sol) + 1.0e-20 * rho : 0.0;
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e) a[e] = lane < nsp ? P.elem_w[e * kWave + lane] : 0.0;
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e) {
      r[e] = wave_sum(a[e] * (sol - sol0));
#pragma unroll
      for (int f2 = 0; f2 <= e; ++f2) G[e][f2] = G[f2][e] = wave_sum(a[e] * wk * a[f2]);
    }
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e)
      if (!(G[e][e] > 0.0)) {  // element absent from the mechanism
        G[e][e] = 1.0;
        r[e] = 0.0;
      }
    // Cholesky solve G lambda = r (G symmetric positive definite)
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e) {
      double d = G[e][e];
#pragma unroll
      for (int q = 0; q < e; ++q) d -= G[e][q] * G[e][q];
      d = sqrt(fmax(d, 1.0e-300));
      G[e][e] = d;
#pragma unroll
      for (int f2 = e + 1; f2 < TGR_NELEM; ++f2) {
        double v = G[f2][e];
#pragma unroll
        for (int q = 0; q < e; ++q) v -= G[f2][q] * G[e][q];
        G[f2][e] = v / d;
      }
    }
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e) {
      double v = r[e];
#pragma unroll
      for (int q = 0; q < e; ++q) v -= G[e][q] * r[q];
      r[e] = v / G[e][e];
    }
#pragma unroll
    for (int e = TGR_NELEM - 1; e >= 0; --e) {
      double v = r[e];
#pragma unroll
      for (int q = e + 1; q < TGR_NELEM; ++q) v -= G[q][e] * r[q];
      r[e] = v / G[e][e];
    }
    double dl = 0.0;
#pragma unroll
    for (int e = 0; e < TGR_NELEM; ++e) dl += a[e] * r[e];
    dl *= -wk;
    if (wave_max(lane < nsp ? fabs(dl) * wacc : 0.0) > 1.0) return 2;
    if (lane < nsp) sol += dl;
  }
  // enthalpy check.  The right-hand side normalises by its own density
  // rho_gp = sum(rho Y) (finitechem.f90:121-133), not by RHO_: the quantity
  // its energy equation conserves is h(sol / rho_gp, sol_T / rho_gp), which
  // differs from h(sol / RHO_, sol_T / RHO_) when sum(Y) /= 1.  T1' from
  // h(Y0', T0') = h(Y1', T1') in those variables (Newton from the solved
  // T'), then mapped back to the unpacked temperature sol_T / RHO_ by
  // rho_gp / RHO_.
  const double rg = wave_sum(lane < nsp ? sol0 : 0.0);
  const double T0 = bcast(sol0, nsp) / rg;
  double hk, cpk;
  thermo_lane<NEQ>(tab, m.tmid, T0, lane, hk, cpk);
  const double h0 = wave_sum(lane < nsp ? (sol0 / rg) * hk : 0.0);
  const double y1g = lane < nsp ? sol / rg : 0.0;
  const double Tsg = bcast(sol, nsp) / rg;
  double Thg = Tsg;
  for (int it = 0; it < 8; ++it) {
    thermo_lane<NEQ>(tab, m.tmid, Thg, lane, hk, cpk);
    const double hs = wave_sum(y1g * hk);
    const double cs = wave_sum(y1g * cpk);
    const double dT = (hs - h0) / cs;
    Thg -= dT;
    if (!(fabs(dT) > 1.0e-12 * Thg)) break;
  }
  const double to_unpacked = rg / rho;
  const double Th = Thg * to_unpacked;
  const double Ts = Tsg * to_unpacked;
  const double y1 = lane < nsp ? sol / rho : 0.0;
  if (!(fabs(Th - Ts) <= P.tol_t)) return 3;
  // unpack and diagnostics as the plain path (finitechem.f90:1041-1058)
  const double soln = lane == nsp ? rho * Th : sol;
  const double v = lane == nsp ? Th : y1;
  if (lane < nsp) p.sc[(int64_t)(p.isc_sc0 + lane) * p.plane + off] = v;
  else if (lane == nsp) p.sc[(int64_t)p.isc_T0 * p.plane + off] = v;
  thermo_lane<NEQ>(tab, m.tmid, Th, lane, hk, cpk);
  const double cpm = wave_sum(lane < nsp ? v * cpk : 0.0);
  if (p.hr && lane == nsp) p.hr[off] = cpm * (soln - sol0) / p.dt;
  if (p.chemsrc) {
    const double src = (soln - sol0) / p.dt;
    const int sp3[3] = {p.idx_h2o, p.idx_co, p.idx_co2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (sp3[q] >= 0 && lane == sp3[q]) p.chemsrc[q * p.plane + off] = src;
      if (sp3[q] < 0 && lane == nsp) p.chemsrc[q * p.plane + off] = 0.0;
    }
  }
  return 0;
}

template <int NEQ, bool AD, bool NET, bool AJ>
__global__ __launch_bounds__(kSurrBlock) void surr_kernel(SurrParams P) {
  constexpr TabLayout L = tab_layout<NEQ>();
  __shared__ __attribute__((aligned(16))) char tab_s[L.bytes];
  __shared__ RhsLds<NEQ> rl[kSurrWaves];
  __shared__ RateCache<NEQ> rcs[kSurrWaves];
  stage_tables((TGR_LDS char*)tab_s, P.s.m.tab, L.bytes);
  __syncthreads();
  const DevMech m = P.s.m;
  const int lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double* ws = P.s.ws + ((int64_t)blockIdx.x * kSurrWaves + w) * NEQ * kWave;
  for (;;) {
    unsigned int c = 0;
    if (lane == 0) c = atomicAdd(P.s.queue, 1u);
    const int64_t cell = (int64_t)(unsigned int)__builtin_amdgcn_readfirstlane((int)c);
    if (cell >= P.s.ncell) break;
    int nit = 0, nfe = 0, nlu = 0, nsub = 0, nje = 0;
    const int st = surr_cell<NEQ, AD, NET, AJ>(P, m, (tchar*)tab_s, (TGR_LDS RhsLds<NEQ>*)&rl[w],
                                               (TGR_LDS RateCache<NEQ>*)&rcs[w], cell, ws, nit, nfe, nlu, nsub, nje);
    if (st < 0) continue;
    if (lane == 0) {
      if (st != 0) {
        const unsigned int r = atomicAdd(P.rej_count, 1u);
        P.rej_list[r] = (int)cell;
      }
      atomicAdd(&P.stotals[st == 0 ? 0 : 1], 1ull);
      atomicAdd(&P.stotals[2], (unsigned long long)nit);
      atomicAdd(&P.stotals[3], (unsigned long long)nfe);
      if (st > 0) atomicAdd(&P.stotals[3 + st], 1ull);
      atomicAdd(&P.stotals[7], (unsigned long long)nlu);
      atomicAdd(&P.stotals[8], (unsigned long long)nsub);
      atomicAdd(&P.stotals[9], (unsigned long long)nje);
    }
  }
}

}  // namespace tgr

using namespace tgr;

extern "C" int tgr_surr_block_waves(void) { return kSurrWaves; }

template <int NEQ>
static hipError_t surr_occ(int* nb, bool adaptive) {
  // the instances differ in registers only a little; the default one decides
  return adaptive ? hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, surr_kernel<NEQ, true, false, false>, kSurrBlock, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, surr_kernel<NEQ, false, false, false>, kSurrBlock, 0);
}

extern "C" int tgr_surr_occupancy(int neq, int adaptive) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (neq) {
    case 10: e = surr_occ<10>(&nb, adaptive != 0); break;
    case 37: e = surr_occ<37>(&nb, adaptive != 0); break;
    case 40: e = surr_occ<40>(&nb, adaptive != 0); break;
    case 46: e = surr_occ<46>(&nb, adaptive != 0); break;
    default: break;
  }
  return e == hipSuccess ? nb : 0;
}

extern "C" int tgr_launch_surr_features(const StepParams* p, int neq, int64_t rows, int in_pad, const float* shift,
                                        const float* scale, float log_dt, uint16_t* X, hipStream_t s) {
  if (rows <= 0) return 0;
  const dim3 grid((unsigned)((rows + 255) / 256)), block(256);
  switch (neq) {
    case 10: hipLaunchKernelGGL(surr_features_kernel<10>, grid, block, 0, s, *p, rows, in_pad, shift, scale, log_dt, X); break;
    case 37: hipLaunchKernelGGL(surr_features_kernel<37>, grid, block, 0, s, *p, rows, in_pad, shift, scale, log_dt, X); break;
    case 40: hipLaunchKernelGGL(surr_features_kernel<40>, grid, block, 0, s, *p, rows, in_pad, shift, scale, log_dt, X); break;
    case 46: hipLaunchKernelGGL(surr_features_kernel<46>, grid, block, 0, s, *p, rows, in_pad, shift, scale, log_dt, X); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int tgr_launch_surr_set_column(uint16_t* X, int64_t rows, int in_pad, int col, float value, hipStream_t s) {
  if (rows <= 0) return 0;
  hipLaunchKernelGGL(surr_set_column_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, X, rows, in_pad,
                     col, value);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int NEQ, bool AJ>
static void launch_surr_j(const SurrParams& P, bool adaptive, int grid, hipStream_t s) {
  const bool net = P.use_guess != 0;
  if (adaptive && net)
    surr_kernel<NEQ, true, true, AJ><<<dim3(grid), dim3(kSurrBlock), 0, s>>>(P);
  else if (adaptive)
    surr_kernel<NEQ, true, false, AJ><<<dim3(grid), dim3(kSurrBlock), 0, s>>>(P);
  else if (net)
    surr_kernel<NEQ, false, true, AJ><<<dim3(grid), dim3(kSurrBlock), 0, s>>>(P);
  else
    surr_kernel<NEQ, false, false, AJ><<<dim3(grid), dim3(kSurrBlock), 0, s>>>(P);
}

template <int NEQ>
static void launch_surr(const SurrParams& P, bool adaptive, int grid, hipStream_t s) {
  if (P.s.jac) launch_surr_j<NEQ, true>(P, adaptive, grid, s);
  else launch_surr_j<NEQ, false>(P, adaptive, grid, s);
}

extern "C" int tgr_launch_surr(const SurrParams* P, int neq, int grid, hipStream_t s) {
  hipLaunchKernelGGL(surr_reset_kernel, dim3(1), dim3(64), 0, s, P->s.queue, P->rej_count, P->stotals);
  const bool ad = P->adaptive != 0;
  switch (neq) {
    case 10: launch_surr<10>(*P, ad, grid, s); break;
    case 37: launch_surr<37>(*P, ad, grid, s); break;
    case 40: launch_surr<40>(*P, ad, grid, s); break;
    case 46: launch_surr<46>(*P, ad, grid, s); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

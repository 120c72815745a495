// HIP kernels of the chemistry step (gfx950): the persistent step kernel,
// the component kernels used by the parity tests, and their launchers.
//
// Every kernel runs kWavesPerBlock independent wavefronts per workgroup; the
// workgroup copies the mechanism tables into LDS once, and each wave owns its
// right-hand-side scratch and Nordsieck history in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bdf_wave.h"
#include "chem_kernels.h"
#include "lane_rhs_dme39.h"
#include "lane_rhs_ethyred.h"
#include "lane_rhs_gri30.h"
#include "lane_rhs_h2.h"

namespace tgr {

// 8 waves = 2 per SIMD: the integrator needs up to 256 VGPRs per lane.
constexpr int kWavesPerBlock = 8;
constexpr int kBlock = 64 * kWavesPerBlock;

__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// finitechem_source_chemistry on a Fortran-layout block.  Persistent: each
// wave pulls cells from a device counter until all are taken, so that cells
// whose step counts differ by 10-100x keep every wave busy.
template <int NEQ, bool AJ>
__device__ __forceinline__ void step_body(const StepParams& p) {
  constexpr TabLayout L = tab_layout<NEQ>();
  __shared__ __attribute__((aligned(16))) char tab_s[L.bytes];
  __shared__ RhsLds<NEQ> rl[kWavesPerBlock];
  __shared__ double zn_s[kWavesPerBlock][kQmaxAbs + 2][kWave];  // Nordsieck columns + tq, l, tau
  stage_tables((TGR_LDS char*)tab_s, p.m.tab, L.bytes);
  __syncthreads();
  const DevMech m = p.m;
  const int lane = lane_id();
  const int w = wave_in_block();
  double* ws = p.ws + ((int64_t)blockIdx.x * kWavesPerBlock + w) * NEQ * kWave;
  const int64_t ntake = p.cell_list ? (int64_t)*p.nlist : p.ncell;
  for (;;) {
    unsigned int c = 0;
    if (lane == 0) c = atomicAdd(p.queue, 1u);
    const int64_t idx = (int64_t)(unsigned int)__builtin_amdgcn_readfirstlane((int)c);
    if (idx >= ntake) break;
    const int64_t cell = p.cell_list ? (int64_t)p.cell_list[idx] : idx;
    int cnt[TGR_NCOUNT];
    if (!process_cell<NEQ, AJ>(p, m, (tchar*)tab_s, (TGR_LDS RhsLds<NEQ>*)&rl[w], cell,
                           *(TGR_LDS double(*)[kQmaxAbs + 2][kWave]) & zn_s[w], ws, cnt))
      continue;
    if (lane == 0) {
      if (p.cell_stats) {
#pragma unroll
        for (int q = 0; q < TGR_NCOUNT; ++q) p.cell_stats[cell * TGR_NCOUNT + q] = cnt[q];
      }
      if (p.totals) {
#pragma unroll
        for (int q = 0; q < TGR_NCOUNT - 1; ++q) atomicAdd(&p.totals[q], (unsigned long long)cnt[q]);
        if (cnt[TGR_NCOUNT - 1] != 0) atomicAdd(&p.totals[TGR_NCOUNT - 1], 1ull);
      }
    }
  }
}

// two waves per SIMD: up to 256 VGPRs per lane (every mechanism)
template <int NEQ, bool AJ>
__global__ __launch_bounds__(kBlock) void step_kernel(StepParams p) {
  step_body<NEQ, AJ>(p);
}

// four waves per SIMD (two workgroups per CU): at most 128 VGPRs per lane, for
// mechanisms whose tables and per-wave scratch leave room for a second
// workgroup's copy in LDS (H2), with the difference-quotient Jacobian (the
// analytic one needs more registers); the autotuner times it against the above
template <int NEQ, bool AJ>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void step_kernel_w4(StepParams p) {
  step_body<NEQ, AJ>(p);
}

// ---------------------------------------------------------------------------
// component kernels (one wave per state, grid-stride over states):
// production rates and the right-hand side
// ---------------------------------------------------------------------------
template <int NEQ>
__global__ __launch_bounds__(kBlock) void prodrates_kernel(DevMech m, int64_t n, const double* C, const double* T,
                                                           const double* P, double* cdot) {
  constexpr TabLayout L = tab_layout<NEQ>();
  __shared__ __attribute__((aligned(16))) char tab_s[L.bytes];
  __shared__ RhsLds<NEQ> rl[kWavesPerBlock];
  stage_tables((TGR_LDS char*)tab_s, m.tab, L.bytes);
  __syncthreads();
  const int lane = lane_id();
  const int w = wave_in_block();
  for (int64_t c = (int64_t)blockIdx.x * kWavesPerBlock + w; c < n; c += (int64_t)gridDim.x * kWavesPerBlock) {
    const double Ck = lane < m.nsp ? C[lane * n + c] : 0.0;
    const double ctot = wave_sum(Ck);
    const double v = prodrates_wave<NEQ>((tchar*)tab_s, *(TGR_LDS RhsLds<NEQ>*)&rl[w], Ck, ctot, T[c], P[c]);
    if (lane < m.nsp) cdot[lane * n + c] = v;
  }
}

template <int NEQ>
__global__ __launch_bounds__(kBlock) void rhs_kernel(DevMech m, int64_t n, double p_gp, const double* y, double* f) {
  constexpr TabLayout L = tab_layout<NEQ>();
  __shared__ __attribute__((aligned(16))) char tab_s[L.bytes];
  __shared__ RhsLds<NEQ> rl[kWavesPerBlock];
  stage_tables((TGR_LDS char*)tab_s, m.tab, L.bytes);
  __syncthreads();
  const int lane = lane_id();
  const int w = wave_in_block();
  for (int64_t c = (int64_t)blockIdx.x * kWavesPerBlock + w; c < n; c += (int64_t)gridDim.x * kWavesPerBlock) {
    const double yv = lane < NEQ ? y[lane * n + c] : 0.0;
    const double v = rhs_wave<NEQ>((tchar*)tab_s, *(TGR_LDS RhsLds<NEQ>*)&rl[w], m.tmid, yv, p_gp);
    if (lane < NEQ) f[lane * n + c] = v;
  }
}

// Lane per state: each lane evaluates a whole right-hand side with the
// generated code of the mechanism (lane_rhs_<mech>.h); component-major
// arrays, so lane c's loads and stores of component k are coalesced.
template <int NEQ>
__global__ __launch_bounds__(64) void rhs_lanes_kernel(int64_t n, double p_gp, const double* y, double* f) {
  const int64_t c = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (c >= n) return;
  double s[NEQ], o[NEQ];
#pragma unroll
  for (int k = 0; k < NEQ; ++k) s[k] = y[k * n + c];
  LaneRhs<NEQ>::eval(s, p_gp, o);
#pragma unroll
  for (int k = 0; k < NEQ; ++k) f[k * n + c] = o[k];
}

// Temperature from mixture enthalpy, lane per state (no routine in the
// reference: the FP64 inversion of COMPTHERMODATA's h(T), HYDROGENF.f90:514-673,
// that a surrogate's enthalpy check needs).  Newton on
// sum_k Y_k h_k(T) = h with sum_k Y_k cp_k(T) as the derivative, the range
// chosen by T > tmid, iterates clamped to [50, 5000] K (finitechem.f90:7-8),
// stop at |dT| <= 1e-10 T, at most 50 iterations; T0 <= 0 starts at 1500 K;
// three consecutive crossings of the range switch end at T = tmid.
// iters: iterations on convergence, -1 otherwise.
constexpr double kTLow = 50.0, kTHigh = 5000.0, kTGuess = 1500.0;
constexpr int kTMaxIt = 50;

template <int NEQ>
__global__ __launch_bounds__(64) void temperature_kernel(DevMech m, int64_t n, const double* Y, const double* hmix,
                                                         const double* T0, double* T, int* iters) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int64_t c = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (c >= n) return;
  double y[nsp];
#pragma unroll
  for (int k = 0; k < nsp; ++k) y[k] = Y[k * n + c];
  const double ht = hmix[c];
  double t = T0 ? T0[c] : 0.0;
  if (!(t > 0.0) || !isfinite(t)) t = kTGuess;
  int res = -1, flips = 0;
  if (isfinite(ht)) {
    for (int it = 1; it <= kTMaxIt; ++it) {
      const TGR_GLOBAL double* th = (const TGR_GLOBAL double*)(m.tab + L.th + (t > m.tmid ? 0 : 13 * kRow));
      double hs = 0.0, cs = 0.0;
#pragma unroll
      for (int k = 0; k < nsp; ++k) {
        const TGR_GLOBAL double* cf = th + k;  // coefficient q of species k at cf[64 q]
        const double hk =
            cf[0] * (t * (cf[64] + t * (cf[128] + t * (cf[192] + t * (cf[256] + t * cf[320])))) + cf[384]);
        const double ck = cf[448] * (cf[512] + t * (cf[576] + t * (cf[640] + t * (cf[704] + t * cf[768]))));
        hs += y[k] * hk;
        cs += y[k] * ck;
      }
      double tn = fmin(fmax(t - (hs - ht) / cs, kTLow), kTHigh);
      bool done = fabs(tn - t) <= 1.0e-10 * tn;
      // h(T) jumps at the range switch: an h inside the jump has no root and
      // Newton alternates across the switch; the answer is then the switch
      flips = ((tn > m.tmid) != (t > m.tmid)) ? flips + 1 : 0;
      if (flips >= 3) {
        tn = m.tmid;
        done = true;
      }
      t = tn;
      if (done) {
        res = it;
        break;
      }
    }
  } else {
    t = ht;
  }
  T[c] = t;
  if (iters) iters[c] = res;
}

// Clears the work counter and the per-call totals: a kernel rather than a
// memset, so that a caller may capture the whole step in a HIP graph.
__global__ void reset_counter_kernel(unsigned int* q, unsigned long long* totals) {
  if (threadIdx.x == 0) *q = 0u;
  if (totals && threadIdx.x < TGR_NCOUNT) totals[threadIdx.x] = 0ull;
}

// grid of the component kernels: enough workgroups to fill the GPU, each
// looping over states (the tables are staged once per workgroup)
inline unsigned component_grid(int64_t n) {
  const int64_t need = (n + kWavesPerBlock - 1) / kWavesPerBlock;
  return (unsigned)(need < 2048 ? need : 2048);
}

}  // namespace tgr

using namespace tgr;

extern "C" int tgr_step_block_waves(void) { return kWavesPerBlock; }

extern "C" int tgr_step_waves_supported(int neq, int waves) { return waves == 2 || (waves == 4 && neq == 10); }

extern "C" int tgr_step_occupancy_w(int neq, int waves) {
  if (waves == 4) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (neq == 10) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel_w4<10, false>, kBlock, 0);
    return e == hipSuccess ? nb : 0;
  }
  return tgr_step_occupancy(neq);
}

extern "C" int tgr_step_occupancy(int neq) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (neq) {
    case 10: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<10, false>, kBlock, 0); break;
    case 37: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<37, false>, kBlock, 0); break;
    case 40: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<40, false>, kBlock, 0); break;
    case 46: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<46, false>, kBlock, 0); break;
    default: break;
  }
  return e == hipSuccess ? nb : 0;
}

extern "C" int tgr_launch_step(const StepParams* p, int neq, int grid, hipStream_t s) {
  if (!p->queue || !p->ws) return -1;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, p->queue, p->totals);
  switch (neq) {
    case 10:
      if (p->waves == 4 && !p->jac) {
        hipLaunchKernelGGL((step_kernel_w4<10, false>), dim3(grid), dim3(kBlock), 0, s, *p);
      } else if (p->jac) {
        hipLaunchKernelGGL((step_kernel<10, true>), dim3(grid), dim3(kBlock), 0, s, *p);
      } else {
        hipLaunchKernelGGL((step_kernel<10, false>), dim3(grid), dim3(kBlock), 0, s, *p);
      }
      break;
    case 37:
      if (p->jac) hipLaunchKernelGGL((step_kernel<37, true>), dim3(grid), dim3(kBlock), 0, s, *p);
      else hipLaunchKernelGGL((step_kernel<37, false>), dim3(grid), dim3(kBlock), 0, s, *p);
      break;
    case 40:
      if (p->jac) hipLaunchKernelGGL((step_kernel<40, true>), dim3(grid), dim3(kBlock), 0, s, *p);
      else hipLaunchKernelGGL((step_kernel<40, false>), dim3(grid), dim3(kBlock), 0, s, *p);
      break;
    case 46:
      if (p->jac) hipLaunchKernelGGL((step_kernel<46, true>), dim3(grid), dim3(kBlock), 0, s, *p);
      else hipLaunchKernelGGL((step_kernel<46, false>), dim3(grid), dim3(kBlock), 0, s, *p);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int tgr_launch_prodrates(const DevMech* m, int64_t n, const double* C, const double* T, const double* P,
                                    double* cdot, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 g(component_grid(n)), b(kBlock);
  switch (m->nsp + 1) {
    case 10: hipLaunchKernelGGL(prodrates_kernel<10>, g, b, 0, s, *m, n, C, T, P, cdot); break;
    case 37: hipLaunchKernelGGL(prodrates_kernel<37>, g, b, 0, s, *m, n, C, T, P, cdot); break;
    case 40: hipLaunchKernelGGL(prodrates_kernel<40>, g, b, 0, s, *m, n, C, T, P, cdot); break;
    case 46: hipLaunchKernelGGL(prodrates_kernel<46>, g, b, 0, s, *m, n, C, T, P, cdot); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int tgr_launch_rhs_lanes(int neq, int64_t n, double p_gp, const double* y, double* f, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 grid((unsigned)((n + 63) / 64));
  switch (neq) {
    case 10: hipLaunchKernelGGL(rhs_lanes_kernel<10>, grid, dim3(64), 0, s, n, p_gp, y, f); break;
    case 37: hipLaunchKernelGGL(rhs_lanes_kernel<37>, grid, dim3(64), 0, s, n, p_gp, y, f); break;
    case 40: hipLaunchKernelGGL(rhs_lanes_kernel<40>, grid, dim3(64), 0, s, n, p_gp, y, f); break;
    case 46: hipLaunchKernelGGL(rhs_lanes_kernel<46>, grid, dim3(64), 0, s, n, p_gp, y, f); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int tgr_launch_rhs(const DevMech* m, int64_t n, double p_gp, const double* y, double* f, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 g(component_grid(n)), b(kBlock);
  switch (m->nsp + 1) {
    case 10: hipLaunchKernelGGL(rhs_kernel<10>, g, b, 0, s, *m, n, p_gp, y, f); break;
    case 37: hipLaunchKernelGGL(rhs_kernel<37>, g, b, 0, s, *m, n, p_gp, y, f); break;
    case 40: hipLaunchKernelGGL(rhs_kernel<40>, g, b, 0, s, *m, n, p_gp, y, f); break;
    case 46: hipLaunchKernelGGL(rhs_kernel<46>, g, b, 0, s, *m, n, p_gp, y, f); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int tgr_launch_temperature(const DevMech* m, int64_t n, const double* Y, const double* hmix,
                                      const double* T0, double* T, int* iters, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 grid((unsigned)((n + 63) / 64));
  switch (m->nsp + 1) {
    case 10: hipLaunchKernelGGL(temperature_kernel<10>, grid, dim3(64), 0, s, *m, n, Y, hmix, T0, T, iters); break;
    case 37: hipLaunchKernelGGL(temperature_kernel<37>, grid, dim3(64), 0, s, *m, n, Y, hmix, T0, T, iters); break;
    case 40: hipLaunchKernelGGL(temperature_kernel<40>, grid, dim3(64), 0, s, *m, n, Y, hmix, T0, T, iters); break;
    case 46: hipLaunchKernelGGL(temperature_kernel<46>, grid, dim3(64), 0, s, *m, n, Y, hmix, T0, T, iters); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

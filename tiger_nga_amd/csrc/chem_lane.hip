// Lane-per-cell step kernel of the chemistry step (gfx950): one thread per
// cell, for mechanisms small enough that a cell's state, Nordsieck history and
// iteration matrix fit one thread (H2: neq 10).  The integration is the same
// CVODE-semantics BDF as the wave-per-cell step kernel (LaneBdf, bdf_lane.h);
// the autotuner times both mappings per mechanism ("step_mapping").
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bdf_lane.h"
#include "chem_kernels.h"
#include "lane_rhs_h2.h"

namespace tgr {

constexpr int kLaneBlock = 64;

// mixture cp of mass fractions v at T from the table block (COMPTHERMODATA)
template <int NEQ>
__device__ __forceinline__ double lane_cp_mix(const DevMech& m, const double (&v)[NEQ], double T) {
  constexpr TabLayout L = tab_layout<NEQ>();
  const TGR_GLOBAL double* th = (const TGR_GLOBAL double*)(m.tab + L.th + (T > m.tmid ? 0 : 13 * kRow));
  double cs = 0.0;
#pragma unroll
  for (int k = 0; k < NEQ - 1; ++k) {
    const TGR_GLOBAL double* cf = th + k;
    cs += v[k] * (cf[448] * (cf[512] + T * (cf[576] + T * (cf[640] + T * (cf[704] + T * cf[768])))));
  }
  return cs;
}

// finitechem_source_chemistry (finitechem.f90:962-1071), one cell per thread:
// pack, integrate, unpack, diagnostics (as process_cell in bdf_wave.h).  Each
// thread takes its next cell from the device counter when its cell is done.
template <int NEQ>
__global__ __launch_bounds__(kLaneBlock) void step_lane_kernel(StepParams p) {
  constexpr int nsp = NEQ - 1;
  __shared__ double m_lds[NEQ * NEQ * kWave];  // the iteration matrices of the wave's threads
  const int lane = threadIdx.x & 63;
  // Jacobian workspace of this thread: NEQ * NEQ doubles, [element][lane]
  TGR_GLOBAL double* jw = (TGR_GLOBAL double*)p.ws + (int64_t)blockIdx.x * NEQ * NEQ * kWave + lane;
  const int64_t ntake = p.cell_list ? (int64_t)*p.nlist : p.ncell;
  int64_t tot[TGR_NCOUNT] = {0};
  LaneBdf<NEQ> b;
  b.p_gp = p.p_gp;
  b.rtol = p.rtol;
  b.atol = p.atol;
  b.qmax = (p.qmax >= 1 && p.qmax <= kQmaxAbs) ? p.qmax : kQmaxAbs;
  b.max_steps = p.max_steps;
  b.jw = jw;
  b.Ml = (TGR_LDS double*)m_lds + lane;
  for (;;) {
    const int64_t idx = (int64_t)atomicAdd(p.queue, 1u);
    if (idx >= ntake) break;
    const int64_t cell = p.cell_list ? (int64_t)p.cell_list[idx] : idx;
    const int i = (int)(cell % p.nx), j = (int)((cell / p.nx) % p.ny), k = (int)(cell / ((int64_t)p.nx * p.ny));
    if (p.mask && p.mask[p.mask_base + (int64_t)j * p.mask_ld + i] == 1) continue;
    const int64_t off = p.base + i * p.sx + j * p.sy + k * p.sz;
    const double rho = p.rho[off];
    double sol[NEQ], sol0[NEQ];
#pragma unroll
    for (int q = 0; q < nsp; ++q) sol[q] = fmax(rho * p.sc[(int64_t)(p.isc_sc0 + q) * p.plane + off], 0.0);
    sol[nsp] = rho * p.sc[(int64_t)p.isc_T0 * p.plane + off];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) sol0[q] = sol[q];
    const int st = b.integrate(p.dt, sol);
    double v[NEQ];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) v[q] = sol[q] / rho;
#pragma unroll
    for (int q = 0; q < nsp; ++q) p.sc[(int64_t)(p.isc_sc0 + q) * p.plane + off] = v[q];
    p.sc[(int64_t)p.isc_T0 * p.plane + off] = v[nsp];
    if (p.hr) p.hr[off] = lane_cp_mix<NEQ>(p.m, v, v[nsp]) * (sol[nsp] - sol0[nsp]) / p.dt;
    if (p.chemsrc) {
      const int sp3[3] = {p.idx_h2o, p.idx_co, p.idx_co2};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < nsp; ++r)
          if (r == sp3[q]) s = (sol[r] - sol0[r]) / p.dt;
        p.chemsrc[q * p.plane + off] = s;
      }
    }
    const int cnt[TGR_NCOUNT] = {b.nst, b.c_nfe, b.c_nfe_dq, b.c_nje, b.c_nsetups, b.c_nni, b.c_netf, b.c_ncfn,
                                 b.c_nlu, st};
    if (p.cell_stats) {
#pragma unroll
      for (int q = 0; q < TGR_NCOUNT; ++q) p.cell_stats[cell * TGR_NCOUNT + q] = cnt[q];
    }
#pragma unroll
    for (int q = 0; q < TGR_NCOUNT - 1; ++q) tot[q] += cnt[q];
    tot[TGR_NCOUNT - 1] += st != 0;
  }
  if (p.totals) {
    // one atomic per counter per wave: the lanes' totals summed across the wave
#pragma unroll
    for (int q = 0; q < TGR_NCOUNT; ++q) {
      double s = (double)tot[q];  // exact: counts stay far below 2^53
      s = wave_sum(s);
      if (lane == 0) atomicAdd(&p.totals[q], (unsigned long long)s);
    }
  }
}

// clears the work counter and the totals (a kernel, so the call can be captured in a graph)
__global__ void reset_lane_counter_kernel(unsigned int* q, unsigned long long* totals) {
  if (threadIdx.x == 0) *q = 0u;
  if (totals && threadIdx.x < TGR_NCOUNT) totals[threadIdx.x] = 0ull;
}

}  // namespace tgr

using namespace tgr;

extern "C" int tgr_step_lane_supported(int neq) { return neq == 10; }

// workspace doubles per workgroup of the lane kernel
extern "C" int64_t tgr_step_lane_ws(int neq) { return (int64_t)neq * neq * kWave; }

extern "C" int tgr_step_lane_occupancy(int neq) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (neq == 10) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_lane_kernel<10>, kLaneBlock, 0);
  return e == hipSuccess ? nb : 0;
}

extern "C" int tgr_launch_step_lane(const StepParams* p, int neq, int grid, hipStream_t s) {
  hipLaunchKernelGGL(reset_lane_counter_kernel, dim3(1), dim3(64), 0, s, p->queue, p->totals);
  if (neq == 10) hipLaunchKernelGGL(step_lane_kernel<10>, dim3(grid), dim3(kLaneBlock), 0, s, *p);
  else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

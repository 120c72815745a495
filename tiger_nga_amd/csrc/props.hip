// Transport and equation-of-state properties on a Fortran-layout block
// (gfx950): the device restatement of
//   finitechem_W           combustion/finitechem.f90:65-74
//   finitechem_density     finitechem.f90:324-345  RHO = P_thermo W_mix / (R T)
//   finitechem_viscosity   finitechem.f90:351-411  Wilke mixing of
//                          eta_k = mucoeff_k sqrt(T) / omegamu(T kOverEps_k)
//   finitechem_diffusivity finitechem.f90:414-461  lambda from the mixture
//                          conductivities, DIFF = lambda / cp / Lewis
//   finitechem_omegamu     finitechem.f90:466-486
// over every cell of the block (ghosts included, as the reference loops over
// imino_:imaxo_ ...) whose mask is not 1.  Two mappings:
//  * props_kernel: one wavefront per cell, lane k = species k; the Wilke
//    double sum runs over j with the pair factors that depend on the molar
//    masses only, (1/sqrt 8)(1 + W_k/W_j)^-1/2 and (W_j/W_k)^1/4, precomputed
//    on the host;
//  * props_lane_kernel: one thread per cell, every lane busy in the double
//    sum.  With a = (1/sqrt 8)(1 + W_k/W_j)^-1/2 and b = (W_j/W_k)^1/4 the
//    reference's pair term a (1 + sqrt(eta_k/eta_j) b)^2 is written
//    (sqrt(a) + u sqrt(a) b)^2, u = sqrt(eta_k) / sqrt(eta_j), so each pair
//    costs a product, two fused multiply-adds and a weight product; the
//    factor pairs (sqrt a, sqrt a b) sit in LDS as one 16-byte broadcast read
//    per pair, with zeros on the diagonal (the reference skips j = k).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chem_device.h"
#include "chem_kernels.h"


namespace tgr {

constexpr double kRcst = 8314.34;  // R_cst, finitechem.f90:15 (exact double)

__device__ __forceinline__ double omegamu(double t) {
  const double m1 = 3.3530622607, m2 = 2.53272006, m3 = 2.9024238575, m4 = 0.11186138893, m5 = 0.8662326188,
               m6 = 1.3913958626, m7 = 3.158490576, m8 = 0.18973411754, m9 = 0.00018682962894;
  const double num = m1 + t * (m2 + t * (m3 + t * m4));
  const double den = m5 + t * (m6 + t * (m7 + t * (m8 + t * m9)));
  return num / den;
}

// omegamu(t) as its numerator and denominator (both positive for t > 0), so
// that a caller dividing by omegamu multiplies by the denominator instead
__device__ __forceinline__ void omegamu_nd(double t, double& num, double& den) {
  const double m1 = 3.3530622607, m2 = 2.53272006, m3 = 2.9024238575, m4 = 0.11186138893, m5 = 0.8662326188,
               m6 = 1.3913958626, m7 = 3.158490576, m8 = 0.18973411754, m9 = 0.00018682962894;
  num = m1 + t * (m2 + t * (m3 + t * m4));
  den = m5 + t * (m6 + t * (m7 + t * (m8 + t * m9)));
}

template <int NEQ>
__global__ __launch_bounds__(256) void props_kernel(PropsParams p) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const bool sp = lane < nsp;
  const TGR_GLOBAL double* pt = (const TGR_GLOBAL double*)p.ptab;
  const double mu = sp ? pt[lane] : 0.0, koe = sp ? pt[64 + lane] : 0.0, Wk = sp ? pt[128 + lane] : 1.0;
  const double lew = (sp && p.lewis) ? p.lewis[lane] : 1.0;
  const int64_t ncell = (int64_t)p.nx * p.ny * p.nz;
  const int wpb = blockDim.x >> 6;
  for (int64_t c = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6); c < ncell; c += (int64_t)gridDim.x * wpb) {
    const int i = (int)(c % p.nx), j = (int)((c / p.nx) % p.ny), k = (int)(c / ((int64_t)p.nx * p.ny));
    if (p.mask && p.mask[p.mask_base + (int64_t)j * p.mask_ld + i] == 1) continue;
    const int64_t off = c;  // the block is the whole array
    const double Y = sp ? p.sc[(int64_t)(p.isc_sc0 + lane) * p.plane + off] : 0.0;
    const double T = p.sc[(int64_t)p.isc_T0 * p.plane + off];
    // finitechem_W: mass fractions clipped to [0, 1]
    const double yt = fmin(fmax(Y, 0.0), 1.0);
    const double Wmix = wave_sum(yt) / wave_sum(sp ? yt / Wk : 0.0);
    if (p.rho && lane == 0) p.rho[off] = p.p_thermo * Wmix / (kRcst * T);
    const double temp = fmin(fmax(T, 50.0), 5000.0);  // Tmin, Tmax (finitechem.f90:7-8)
    const double eta = sp ? mu * sqrt(temp) / omegamu(temp * koe) : 1.0;
    if (p.visc) {
      // per pair: (eta_k/eta_j)^1/2 as sqrt(eta_k) / sqrt(eta_j) and the
      // weight Y_j W_mix / W_j formed once per species (rounding-level
      // differences from the reference's expression order)
      const double se = sqrt(eta), ise = 1.0 / se, yw = sp ? Y * Wmix / Wk : 0.0;
      double sum1 = 0.0;
      for (int q = 0; q < nsp; ++q) {
        const double iseq = bcast(ise, q), ywq = bcast(yw, q);
        const double a = pt[192 + q * 64 + lane], b = pt[192 + nsp * 64 + q * 64 + lane];
        const double t1 = se * iseq * b;
        const double phi = a * ((1.0 + t1) * (1.0 + t1));
        sum1 = q != lane ? fma(ywq, phi, sum1) : sum1;
      }
      sum1 = sum1 + Y * Wmix / Wk;
      const double v = wave_sum((sp && Y > 0.0) ? (Y * Wmix * eta / (Wk * sum1)) : 0.0);
      if (lane == 0) p.visc[off] = v;
    }
    if (p.diff) {
      // COMPTHERMODATA cp at the clipped temperature (table rows in HBM)
      const TGR_GLOBAL double* th =
          (const TGR_GLOBAL double*)(p.mtab + L.th + (temp > p.tmid ? 0 : 13 * kRow)) + lane;
      const double cpk = sp ? th[448] * (th[512] + temp * (th[576] + temp * (th[640] + temp * (th[704] + temp * th[768]))))
                            : 0.0;
      const double cond = eta * (cpk + 1.25 * kRcst / Wk);
      const double s1 = wave_sum(sp ? Y * Wmix / (cond * Wk) : 0.0);
      const double s2 = wave_sum(sp ? Y * cond * Wmix / Wk : 0.0);
      const double lambda = 0.5 * (s2 + (1.0 / s1));
      const double cpm = wave_sum(Y * cpk);
      const double D = lambda / cpm;
      if (sp) p.diff[(int64_t)(p.isc_sc0 + lane) * p.plane + off] = D / lew;
      if (lane == 0) p.diff[(int64_t)p.isc_T0 * p.plane + off] = D;
      if (lane == 0 && p.isc_z0 >= 0) p.diff[(int64_t)p.isc_z0 * p.plane + off] = D;
    }
  }
}

// One thread per cell (the block's cells in their storage order, so every
// field access is coalesced).  ise / yw: 1/sqrt(eta_j) and Y_j W_mix / W_j of
// all species in registers; the loop over k keeps one species' eta, cp and
// pair sums at a time.  Sums run in species order (the wave kernel's are
// tree reductions) and eta = mucoeff sqrt(T) den / num takes omegamu's
// numerator and denominator with one division: rounding-level differences.
template <int NEQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void props_lane_kernel(PropsParams p) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  constexpr int kPair = 192 + 2 * nsp * 64;  // [k][j] (sqrt a, sqrt a b), host table
  __shared__ double2 pair[nsp * nsp];
  const TGR_GLOBAL double* src = (const TGR_GLOBAL double*)(p.ptab + kPair);
  for (int i = threadIdx.x; i < nsp * nsp; i += blockDim.x) {
    const double a = src[2 * i], b = src[2 * i + 1];
    pair[i].x = a;
    pair[i].y = b;
  }
  __syncthreads();
  const int64_t ncell = (int64_t)p.nx * p.ny * p.nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  if (p.mask && p.mask[p.mask_base + (int64_t)((c / p.nx) % p.ny) * p.mask_ld + (int)(c % p.nx)] == 1) return;
  const TGR_CONST double* pt = (const TGR_CONST double*)p.ptab;
  const TGR_GLOBAL double* ys = (const TGR_GLOBAL double*)p.sc + (int64_t)p.isc_sc0 * p.plane + c;
  const double T = p.sc[(int64_t)p.isc_T0 * p.plane + c];
  const TGR_CONST double* iw = (const TGR_CONST double*)(p.mtab + L.iw);  // 1 / W_k
  // finitechem_W: mass fractions clipped to [0, 1]; yw first holds Y_j / W_j
  // (each Y read once), scaled by W_mix below
  double ise[nsp], yw[nsp];
  double sy = 0.0, syw = 0.0;
#pragma unroll
  for (int q = 0; q < nsp; ++q) {
    const double Y = ys[q * p.plane];
    const double yt = fmin(fmax(Y, 0.0), 1.0);
    sy += yt;
    syw = fma(yt, iw[q], syw);
    yw[q] = Y * iw[q];
  }
  const double Wmix = sy / syw;
  const double temp = fmin(fmax(T, 50.0), 5000.0);  // Tmin, Tmax (finitechem.f90:7-8)
  const double st = sqrt(temp);
  if (p.visc) {
#pragma unroll
    for (int q = 0; q < nsp; ++q) {
      double num, den;
      omegamu_nd(temp * pt[64 + q], num, den);
      ise[q] = sqrt(num / (pt[q] * st * den));  // 1 / sqrt(eta_j)
      yw[q] *= Wmix;
    }
  }
  // COMPTHERMODATA cp rows (both temperature ranges; each lane picks its own)
  const TGR_CONST double* thi = (const TGR_CONST double*)(p.mtab + L.th);
  const TGR_CONST double* tlo = thi + 13 * 64;
  const bool hi = temp > p.tmid;
  double visc = 0.0, s1 = 0.0, s2 = 0.0, cpm = 0.0;
#pragma unroll 1
  for (int k = 0; k < nsp; ++k) {
    const double Y = ys[k * p.plane];
    double num, den;
    omegamu_nd(temp * pt[64 + k], num, den);
    const double eta = pt[k] * st * den / num;
    const double ywk = Y * iw[k] * Wmix;
    if (p.visc) {
      const double se = sqrt(eta);
      // the factor pairs in chunks of kCh, each chunk's LDS reads issued
      // while the previous chunk computes.  The row offset passes through an
      // empty asm statement that takes the running sum as an input, so the
      // compiler cannot hoist every read of the row into registers at once.
      constexpr int kCh = 4, nch = (nsp + kCh - 1) / kCh;
      double2 buf[2][kCh];
      int row = k * nsp;
#pragma unroll
      for (int i = 0; i < kCh; ++i)
        if (i < nsp) buf[0][i] = pair[row + i];
      double s = 0.0;
#pragma unroll
      for (int ch = 0; ch < nch; ++ch) {
        asm volatile("" : "+v"(row) : "v"(s));
#pragma unroll
        for (int i = 0; i < kCh; ++i)
          if ((ch + 1) * kCh + i < nsp) buf[(ch + 1) & 1][i] = pair[row + (ch + 1) * kCh + i];
#pragma unroll
        for (int i = 0; i < kCh; ++i) {
          const int q = ch * kCh + i;
          if (q < nsp) {
            const double f = fma(se * ise[q], buf[ch & 1][i].y, buf[ch & 1][i].x);
            s = fma(yw[q] * f, f, s);
          }
        }
      }
      s += ywk;
      visc += Y > 0.0 ? ywk * eta / s : 0.0;
    }
    if (p.diff) {
      const double ac = hi ? thi[7 * 64 + k] : tlo[7 * 64 + k], d1 = hi ? thi[8 * 64 + k] : tlo[8 * 64 + k],
                   d2 = hi ? thi[9 * 64 + k] : tlo[9 * 64 + k], d3 = hi ? thi[10 * 64 + k] : tlo[10 * 64 + k],
                   d4 = hi ? thi[11 * 64 + k] : tlo[11 * 64 + k], d5 = hi ? thi[12 * 64 + k] : tlo[12 * 64 + k];
      const double cpk = ac * (d1 + temp * (d2 + temp * (d3 + temp * (d4 + temp * d5))));
      const double cond = eta * fma(1.25 * kRcst, iw[k], cpk);
      s1 += ywk / cond;
      s2 += ywk * cond;
      cpm += Y * cpk;
    }
  }
  if (p.rho) p.rho[c] = p.p_thermo * Wmix / (kRcst * T);
  if (p.visc) p.visc[c] = visc;
  if (p.diff) {
    const double D = 0.5 * (s2 + (1.0 / s1)) / cpm;
#pragma unroll 1
    for (int k = 0; k < nsp; ++k) p.diff[(int64_t)(p.isc_sc0 + k) * p.plane + c] = p.lewis ? D / p.lewis[k] : D;
    p.diff[(int64_t)p.isc_T0 * p.plane + c] = D;
    if (p.isc_z0 >= 0) p.diff[(int64_t)p.isc_z0 * p.plane + c] = D;
  }
}

}  // namespace tgr

using namespace tgr;

extern "C" int tgr_launch_props(const PropsParams* p, int neq, int lane_map, hipStream_t s) {
  const int64_t ncell = (int64_t)p->nx * p->ny * p->nz;
  if (ncell <= 0) return 0;
  if (lane_map) {
    const int64_t nb = (ncell + 255) / 256;
    if (nb > 0x7fffffff) return -1;
    const dim3 grid((unsigned)nb), block(256);
    switch (neq) {
      case 10: hipLaunchKernelGGL(props_lane_kernel<10>, grid, block, 0, s, *p); break;
      case 37: hipLaunchKernelGGL(props_lane_kernel<37>, grid, block, 0, s, *p); break;
      case 40: hipLaunchKernelGGL(props_lane_kernel<40>, grid, block, 0, s, *p); break;
      case 46: hipLaunchKernelGGL(props_lane_kernel<46>, grid, block, 0, s, *p); break;
      default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  const int64_t need = (ncell + 3) / 4;
  const dim3 grid((unsigned)(need < 8192 ? need : 8192)), block(256);
  switch (neq) {
    case 10: hipLaunchKernelGGL(props_kernel<10>, grid, block, 0, s, *p); break;
    case 37: hipLaunchKernelGGL(props_kernel<37>, grid, block, 0, s, *p); break;
    case 40: hipLaunchKernelGGL(props_kernel<40>, grid, block, 0, s, *p); break;
    case 46: hipLaunchKernelGGL(props_kernel<46>, grid, block, 0, s, *p); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

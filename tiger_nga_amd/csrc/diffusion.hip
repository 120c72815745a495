// Molar-diffusion correction of finite-rate chemistry on the GPU (gfx950):
// the body of finitechem_diffusion (reference combustion/finitechem.f90:538-923)
// and the source update of finitechem_source_transport (492-528).
//
// Stencil work over the caller's Fortran-layout fields, one lane per cell
// (i fastest, so every field and operator access of a wave is coalesced):
//   diff_cell_kernel  every cell of the ghosted block: mixture molar mass
//                     W_gp (finitechem_W) and heat capacity cp_mix (cv when
//                     compressible / constant volume);
//   diff_face_kernel  every face cell of the stencil range: the correction
//                     velocities corr_vel_{x,y,z} = sum over species of the
//                     Fickian (and molar) face fluxes;
//   diff_main_kernel  every interior cell: per species the corrected flux at
//                     the faces the divergence needs (recomputed per cell
//                     instead of stored: HBM-bound work, the recomputation
//                     is cheap arithmetic on cached neighbours), the species
//                     term div F_k, the enthalpy flux cp_k(T) interp_m(F_k)
//                     grad_m(T), the H2O / CO / CO2 dump, gas radiation, the
//                     cp-gradient term, / cp, * dt, and the optional update of
//                     the transport source.
// Readings where the reference text does not conform (the CPU oracle's too):
// SC(i-st2:j+st1) at 652 read as i+st1; the z sections k-st2:k-st1 at
// 634-635 and 706 read as k-st2:k+st1 (identical at second order); Cp_sp(isc)
// at 720-730 read as the species' own cp.  The burner heat loss (796-802) and
// the external ignition source (842-859) belong to single cases' meshes and
// are left to the caller.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "chem_device.h"
#include "chem_kernels.h"

namespace tgr {


namespace {

__device__ __forceinline__ int64_t c3(const DiffGeom& g, int i, int j, int k) {
  return (int64_t)(i - g.imino_) + (int64_t)g.nxo * ((int64_t)(j - g.jmino_) + (int64_t)g.nyo * (k - g.kmino_));
}

__device__ __forceinline__ bool wall(const DiffGeom& g, const int* mask, int i, int j) {
  return mask && mask[(i - g.imino) + (int64_t)g.mx * (j - g.jmino)] == 1;
}

// operator coefficient (i, j, st), first index from i0 (ni entries), second
// from j0 (nj), stencil from s0
__device__ __forceinline__ double op(const double* a, int i, int i0, int ni, int j, int j0, int nj, int st, int s0) {
  return a[(int64_t)(i - i0) + (int64_t)ni * ((int64_t)(j - j0) + (int64_t)nj * (st - s0))];
}

// thermo row (range, coefficient) of species k from the mechanism table block
__device__ __forceinline__ void cp_species(const char* tab, const TabLayout& L, double tmid, double T, int k,
                                           double& cp) {
  const double* c = (const double*)(tab + L.th + (T > tmid ? 0 : 13 * kRow)) + k;
  const int r = kRow / 8;
  const double ac = c[7 * r], d1 = c[8 * r], d2 = c[9 * r], d3 = c[10 * r], d4 = c[11 * r], d5 = c[12 * r];
  cp = ac * (d1 + T * (d2 + T * (d3 + T * (d4 + T * d5))));
}

}  // namespace

template <int NEQ>
__global__ __launch_bounds__(256) void diff_cell_kernel(DiffParams p) {
  constexpr TabLayout L = tab_layout<NEQ>();
  constexpr int N = NEQ - 1;
  const DiffGeom& g = p.gm;
  const int64_t n = (int64_t)g.nxo * g.nyo * g.nzo;
  const double* W = (const double*)(p.mech.tab + L.w);
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.imino_ + (int)(c % g.nxo), j = g.jmino_ + (int)((c / g.nxo) % g.nyo);
    double w = 1.0, cp = 0.0;
    if (!wall(g, p.mask, i, j)) {
      const double T = p.sc[(int64_t)g.isc_T0 * g.plane + c];
      double s1 = 0.0, s2 = 0.0;
      for (int k = 0; k < N; ++k) {
        const double yk = p.sc[(int64_t)(g.isc_sc0 + k) * g.plane + c];
        const double y = fmin(fmax(yk, 0.0), 1.0);
        s1 += y;
        s2 += y / W[k];
        double cpk;
        cp_species(p.mech.tab, L, p.mech.tmid, T, k, cpk);
        cp += yk * cpk;
      }
      w = s1 / s2;
      if (p.cv) cp = cp - 8314.34 / w;
    }
    p.wg[c] = w;
    p.cpm[c] = cp;
  }
}

template <int NEQ>
__global__ __launch_bounds__(256) void diff_face_kernel(DiffParams p) {
  constexpr int N = NEQ - 1;
  const DiffGeom& g = p.gm;
  const tgr_metric_dev& m = p.m;
  const int st1 = m.st1, st2 = m.st2;
  const int fi0 = g.imin_ - st1, fni = g.imax_ + st2 - fi0 + 1;
  const int fj0 = g.jmin_ - st1, fnj = g.jmax_ + st2 - fj0 + 1;
  const int fk0 = g.kmin_ - st1, fnk = g.kmax_ + st2 - fk0 + 1;
  const int64_t n = (int64_t)fni * fnj * fnk;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = fi0 + (int)(q % fni), j = fj0 + (int)((q / fni) % fnj), k = fk0 + (int)(q / ((int64_t)fni * fnj));
    const int64_t o = c3(g, i, j, k);
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (!wall(g, p.mask, i, j)) {
      for (int sp = 0; sp < N; ++sp) {
        const double* Y = p.sc + (int64_t)(g.isc_sc0 + sp) * g.plane;
        const double* D = p.diff + (int64_t)(g.isc_sc0 + sp) * g.plane;
        double a = 0, b = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
        double ix = 0, gx = 0, iy = 0, gy = 0, iz = 0, gz = 0;
        for (int s = -st2; s <= st1; ++s) {
          const int64_t ox = c3(g, i + s, j, k), oy = c3(g, i, j + s, k), oz = c3(g, i, j, k + s);
          const double wx = op(m.interp_sc_x, i, fi0, fni, j, g.jmino_, g.nyo, s, -st2);
          const double wy = op(m.interp_sc_y, i, g.imino_, g.nxo, j, fj0, fnj, s, -st2);
          const double wz = op(m.interp_sc_z, i, g.imino_, g.nxo, j, fj0, fnj, s, -st2);
          const double ggx = op(m.grad_x, i, fi0, fni, j, fj0, fnj, s, -st2);
          const double ggy = op(m.grad_y, i, fi0, fni, j, fj0, fnj, s, -st2);
          const double ggz = op(m.grad_z, i, fi0, fni, j, fj0, fnj, s, -st2);
          a += wx * D[ox];
          b += ggx * Y[ox];
          a2 += wy * D[oy];
          b2 += ggy * Y[oy];
          a3 += wz * D[oz];
          b3 += ggz * Y[oz];
          if (p.molar) {
            ix += wx * (Y[ox] / p.wg[ox]);
            gx += ggx * p.wg[ox];
            iy += wy * (Y[oy] / p.wg[oy]);
            gy += ggy * p.wg[oy];
            iz += wz * (Y[oz] / p.wg[oz]);
            gz += ggz * p.wg[oz];
          }
        }
        vx += a * b;
        vy += a2 * b2;
        vz += a3 * b3;
        if (p.molar) {
          vx += a * ix * gx;
          vy += a2 * iy * gy;
          vz += a3 * iz * gz;
        }
      }
    }
    p.cvel[o] = vx;
    p.cvel[g.plane + o] = vy;
    p.cvel[2 * g.plane + o] = vz;
  }
}

namespace {

// Corrected (and molar) flux and Fickian flux of species sp at the face
// cell (i, j, k) in direction d (0 x, 1 y, 2 z); 0 at a wall face.
struct Flux {
  double corr, fick;
};

__device__ __forceinline__ Flux face_flux(const DiffParams& p, int d, int i, int j, int k, const double* Y,
                                          const double* D) {
  const DiffGeom& g = p.gm;
  const tgr_metric_dev& m = p.m;
  if (wall(g, p.mask, i, j)) return {0.0, 0.0};
  const int st1 = m.st1, st2 = m.st2;
  const int fi0 = g.imin_ - st1, fni = g.imax_ + st2 - fi0 + 1;
  const int fj0 = g.jmin_ - st1, fnj = g.jmax_ + st2 - fj0 + 1;
  double sy = 0, a = 0, b = 0, iyw = 0, gw = 0;
  for (int s = -st2; s <= st1; ++s) {
    const int64_t o = d == 0 ? c3(g, i + s, j, k) : (d == 1 ? c3(g, i, j + s, k) : c3(g, i, j, k + s));
    const double w = d == 0 ? op(m.interp_sc_x, i, fi0, fni, j, g.jmino_, g.nyo, s, -st2)
                            : op(d == 1 ? m.interp_sc_y : m.interp_sc_z, i, g.imino_, g.nxo, j, fj0, fnj, s, -st2);
    const double gg = op(d == 0 ? m.grad_x : (d == 1 ? m.grad_y : m.grad_z), i, fi0, fni, j, fj0, fnj, s, -st2);
    sy += w * Y[o];
    a += w * D[o];
    b += gg * Y[o];
    if (p.molar) {
      iyw += w * (Y[o] / p.wg[o]);
      gw += gg * p.wg[o];
    }
  }
  Flux f;
  f.corr = -p.cvel[(int64_t)d * g.plane + c3(g, i, j, k)] * sy;
  if (p.molar) f.corr += a * iyw * gw;
  f.fick = a * b;
  return f;
}

// Gas radiation of the temperature equation (finitechem.f90:738-792):
// optically thin loss of H2O, CO2, CH4 and CO with the reference's Planck-mean
// absorption fits, Tb = 298 K; a species the mechanism lacks contributes 0.
__device__ __forceinline__ double gas_radiation(const DiffParams& p, double T, int64_t o) {
  const DiffGeom& g = p.gm;
  const double tm1 = 1000.0 / T;
  const double T2 = T * T;
  const double aH2O =
      -0.23093 + tm1 * (-1.12390 + tm1 * (9.41530 + tm1 * (-2.99880 + tm1 * (0.51382 + tm1 * -1.86840E-05))));
  const double aCO2 = 18.741 + tm1 * (-121.310 + tm1 * (273.500 + tm1 * (-194.050 + tm1 * (56.310 + tm1 * -5.8169))));
  const double aCH4 = 6.6334 - 0.0035686 * T + 1.6682e-08 * T2 + 2.5611e-10 * T2 * T - 2.6558e-14 * T2 * T2;
  const bool lo = T <= 750.0;
  const double aCO = lo ? 4.7869 + T * (-0.06953 + T * (2.95775e-4 + T * (-4.25732e-7 + T * 2.02894e-10)))
                        : 10.09 + T * (-0.01183 + T * (4.7753e-6 + T * (-5.87209e-10 + T * -2.5334e-14)));
  const double* S = p.sc;
  double sum = 0.0;
  if (p.i_co2 >= 0) sum += aCO2 * S[(int64_t)(g.isc_sc0 + p.i_co2) * g.plane + o] / p.w_co2;
  if (p.i_h2o >= 0) sum += aH2O * S[(int64_t)(g.isc_sc0 + p.i_h2o) * g.plane + o] / p.w_h2o;
  if (p.i_ch4 >= 0) sum += aCH4 * S[(int64_t)(g.isc_sc0 + p.i_ch4) * g.plane + o] / p.w_ch4;
  if (p.i_co >= 0) sum += aCO * S[(int64_t)(g.isc_sc0 + p.i_co) * g.plane + o] / p.w_co;
  const double radconst = 4.0 * 5.67051e-8 * 8314.34 / 101320.0;
  const double T4 = T2 * T2, Tb = 298.0;
  return -radconst * p.rho[o] * T * (T4 - Tb * Tb * Tb * Tb) * sum;
}

}  // namespace

template <int NEQ>
__global__ __launch_bounds__(256) void diff_main_kernel(DiffParams p) {
  constexpr TabLayout L = tab_layout<NEQ>();
  constexpr int N = NEQ - 1;
  const DiffGeom& g = p.gm;
  const tgr_metric_dev& m = p.m;
  const int st1 = m.st1, st2 = m.st2, stp = st1 + st2;
  const int nx = g.imax_ - g.imin_ + 1, ny = g.jmax_ - g.jmin_ + 1, nz = g.kmax_ - g.kmin_ + 1;
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.imin_ + (int)(q % nx), j = g.jmin_ + (int)((q / nx) % ny), k = g.kmin_ + (int)(q / ((int64_t)nx * ny));
    const int64_t o = c3(g, i, j, k);
    if (wall(g, p.mask, i, j)) continue;
    const int a0 = i - g.imin_, b0 = j - g.jmin_;
    const double* Tf = p.sc + (int64_t)g.isc_T0 * g.plane;
    const double T = Tf[o];
    // cell gradients of T and cp_mix
    double tx = 0, ty = 0, tz = 0, cx = 0, cy = 0, cz = 0;
    for (int s = -stp; s <= stp; ++s) {
      const double gxm = op(m.grad_xm, i, g.imin_, nx, j, g.jmin_, ny, s, -stp);
      const double gym = op(m.grad_ym, i, g.imin_, nx, j, g.jmin_, ny, s, -stp);
      const double gzm = op(m.grad_zm, i, g.imin_, nx, j, g.jmin_, ny, s, -stp);
      tx += gxm * Tf[c3(g, i + s, j, k)];
      ty += gym * Tf[c3(g, i, j + s, k)];
      tz += gzm * Tf[c3(g, i, j, k + s)];
      cx += gxm * p.cpm[c3(g, i + s, j, k)];
      cy += gym * p.cpm[c3(g, i, j + s, k)];
      cz += gzm * p.cpm[c3(g, i, j, k + s)];
    }
    double eT = 0.0;  // temperature term before the cp-gradient part and / cp
    for (int sp = 0; sp < N; ++sp) {
      const double* Y = p.sc + (int64_t)(g.isc_sc0 + sp) * g.plane;
      const double* D = p.diff + (int64_t)(g.isc_sc0 + sp) * g.plane;
      double dcor = 0.0, dfick = 0.0, fxm = 0.0, fym = 0.0, fzm = 0.0;
      for (int s = -st1; s <= st2; ++s) {
        const Flux fx = face_flux(p, 0, i + s, j, k, Y, D);
        const Flux fy = face_flux(p, 1, i, j + s, k, Y, D);
        const Flux fz = face_flux(p, 2, i, j, k + s, Y, D);
        const double du = op(m.div_u, i, g.imin_, nx, j, g.jmin_, ny, s, -st1);
        const double dv = op(m.div_v, i, g.imin_, nx, j, g.jmin_, ny, s, -st1);
        const double dw = op(m.div_w, i, g.imin_, nx, j, g.jmin_, ny, s, -st1);
        dcor += du * fx.corr + dv * fy.corr + dw * fz.corr;
        dfick += du * fx.fick + dv * fy.fick + dw * fz.fick;
        fxm += op(m.interp_u_xm, i, g.imin_, nx, j, g.jmin_, ny, s, -st1) * (fx.corr + fx.fick);
        fym += op(m.interp_v_ym, i, g.imin_, nx, j, g.jmin_, ny, s, -st1) * (fy.corr + fy.fick);
        fzm += op(m.interp_w_zm, i, g.imin_, nx, j, g.jmin_, ny, s, -st1) * (fz.corr + fz.fick);
      }
      double cpk;
      cp_species(p.mech.tab, L, p.mech.tmid, T, sp, cpk);
      eT = eT + cpk * fxm * tx;
      eT = eT + cpk * fym * ty;
      eT = eT + cpk * fzm * tz;
      if (p.corr) p.corr[(int64_t)sp * g.plane + o] = dcor * p.dt;
      if (p.src) p.src[(int64_t)(g.isc_sc0 + sp) * g.plane + o] += dcor * p.dt;
      if (p.dgp) {
        const int slot = sp == p.i_h2o ? 0 : (sp == p.i_co ? 1 : (sp == p.i_co2 ? 2 : -1));
        if (slot >= 0) p.dgp[(int64_t)slot * g.plane + o] = dfick + dcor;
      }
    }
    if (p.gasrad) eT += gas_radiation(p, T, o);
    const double dT = p.diff[(int64_t)g.isc_T0 * g.plane + o];
    eT = eT + dT * cx * tx;
    eT = eT + dT * cy * ty;
    eT = eT + dT * cz * tz;
    eT = eT / p.cpm[o] * p.dt;
    if (p.corr) p.corr[(int64_t)N * g.plane + o] = eT;
    if (p.src) p.src[(int64_t)g.isc_T0 * g.plane + o] += eT;
  }
}

// ---- second order (st1 = 0, st2 = 1) ----------------------------------------
// The same sums as the kernels above, in the same order, with what does not
// depend on the species hoisted out of the species loop: the operator
// coefficients, the molar factors w / W_gp and sum grad * W_gp (the molar
// term's w (Y / W_gp) becomes (w / W_gp) Y: no division per species), the
// correction velocity at each face; per species only Y and D at the cell and its six
// neighbours are loaded (each once), instead of once per face and stencil
// point.

// correction velocities at the face cells (diff_face_kernel, st1 = 0, st2 = 1:
// a face interpolates its cell and the one before it in each direction)
template <int NEQ>
__global__ __launch_bounds__(256) void diff_face2_kernel(DiffParams p) {
  constexpr int N = NEQ - 1;
  const DiffGeom& g = p.gm;
  const tgr_metric_dev& m = p.m;
  const int fi0 = g.imin_, fni = g.imax_ + 1 - fi0 + 1;
  const int fj0 = g.jmin_, fnj = g.jmax_ + 1 - fj0 + 1;
  const int fk0 = g.kmin_, fnk = g.kmax_ + 1 - fk0 + 1;
  const int64_t n = (int64_t)fni * fnj * fnk;
  const int64_t sy = g.nxo, sz = (int64_t)g.nxo * g.nyo;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = fi0 + (int)(q % fni), j = fj0 + (int)((q / fni) % fnj), k = fk0 + (int)(q / ((int64_t)fni * fnj));
    const int64_t o = c3(g, i, j, k);
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (!wall(g, p.mask, i, j)) {
      // coefficients of the two stencil points (s = -1, 0) per direction
      double w[3][2], gg[3][2], wi[3][2], gw[3];
      const int64_t nb[3] = {o - 1, o - sy, o - sz};
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int s = t - 1;
        w[0][t] = op(m.interp_sc_x, i, fi0, fni, j, g.jmino_, g.nyo, s, -1);
        w[1][t] = op(m.interp_sc_y, i, g.imino_, g.nxo, j, fj0, fnj, s, -1);
        w[2][t] = op(m.interp_sc_z, i, g.imino_, g.nxo, j, fj0, fnj, s, -1);
        gg[0][t] = op(m.grad_x, i, fi0, fni, j, fj0, fnj, s, -1);
        gg[1][t] = op(m.grad_y, i, fi0, fni, j, fj0, fnj, s, -1);
        gg[2][t] = op(m.grad_z, i, fi0, fni, j, fj0, fnj, s, -1);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        gw[d] = 0.0;
        if (p.molar) {
          const double w0 = p.wg[nb[d]], w1 = p.wg[o];
          gw[d] += gg[d][0] * w0;
          gw[d] += gg[d][1] * w1;
          wi[d][0] = 1.0 / w0;
          wi[d][1] = 1.0 / w1;
        }
      }
      double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 2
      for (int sp = 0; sp < N; ++sp) {
        const double* Y = p.sc + (int64_t)(g.isc_sc0 + sp) * g.plane;
        const double* D = p.diff + (int64_t)(g.isc_sc0 + sp) * g.plane;
        const double yc = Y[o], dc = D[o];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double y0 = Y[nb[d]], d0 = D[nb[d]];
          double a = 0, b = 0, ii = 0;
          a += w[d][0] * d0;
          b += gg[d][0] * y0;
          a += w[d][1] * dc;
          b += gg[d][1] * yc;
          acc[d] += a * b;
          if (p.molar) {
            ii += w[d][0] * (y0 * wi[d][0]);
            ii += w[d][1] * (yc * wi[d][1]);
            acc[d] += a * ii * gw[d];
          }
        }
      }
      vx = acc[0];
      vy = acc[1];
      vz = acc[2];
    }
    p.cvel[o] = vx;
    p.cvel[g.plane + o] = vy;
    p.cvel[2 * g.plane + o] = vz;
  }
}

// the per-cell terms (diff_main_kernel, st1 = 0, st2 = 1: faces at the cell
// and the next one in each direction, each interpolating two cells).  The
// species loop reads the operator coefficients through restrict pointers, so
// that the compiler keeps the loop-invariant ones in registers as far as they
// fit; per species it loads Y and D at the cell and its six neighbours once.
template <int NEQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void diff_main2_kernel(DiffParams p) {
  constexpr TabLayout L = tab_layout<NEQ>();
  constexpr int N = NEQ - 1;
  const DiffGeom& g = p.gm;
  const int nx = g.imax_ - g.imin_ + 1, ny = g.jmax_ - g.jmin_ + 1, nz = g.kmax_ - g.kmin_ + 1;
  const int fi0 = g.imin_, fni = nx + 1, fj0 = g.jmin_, fnj = ny + 1;
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t sy = g.nxo, sz = (int64_t)g.nxo * g.nyo;
  const double* __restrict__ isx = p.m.interp_sc_x;
  const double* __restrict__ isy = p.m.interp_sc_y;
  const double* __restrict__ isz = p.m.interp_sc_z;
  const double* __restrict__ grx = p.m.grad_x;
  const double* __restrict__ gry = p.m.grad_y;
  const double* __restrict__ grz = p.m.grad_z;
  const double* __restrict__ dvu = p.m.div_u;
  const double* __restrict__ dvv = p.m.div_v;
  const double* __restrict__ dvw = p.m.div_w;
  const double* __restrict__ imx = p.m.interp_u_xm;
  const double* __restrict__ imy = p.m.interp_v_ym;
  const double* __restrict__ imz = p.m.interp_w_zm;
  const double* __restrict__ sc = p.sc;
  const double* __restrict__ df = p.diff;
  const double* __restrict__ wgp = p.wg;
  const double* __restrict__ cpm = p.cpm;
  const double* __restrict__ cvel = p.cvel;
  const char* __restrict__ tab = p.mech.tab;
  double* __restrict__ corr = p.corr;
  double* __restrict__ src = p.src;
  double* __restrict__ dgp = p.dgp;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.imin_ + (int)(q % nx), j = g.jmin_ + (int)((q / nx) % ny), k = g.kmin_ + (int)(q / ((int64_t)nx * ny));
    const int64_t o = c3(g, i, j, k);
    if (wall(g, p.mask, i, j)) continue;
    const double* Tf = sc + (int64_t)g.isc_T0 * g.plane;
    const double T = Tf[o];
    // cell gradients of T and cp_mix (stp = 1)
    double tx = 0, ty = 0, tz = 0, cx = 0, cy = 0, cz = 0;
#pragma unroll
    for (int s = -1; s <= 1; ++s) {
      const double gxm = op(p.m.grad_xm, i, g.imin_, nx, j, g.jmin_, ny, s, -1);
      const double gym = op(p.m.grad_ym, i, g.imin_, nx, j, g.jmin_, ny, s, -1);
      const double gzm = op(p.m.grad_zm, i, g.imin_, nx, j, g.jmin_, ny, s, -1);
      tx += gxm * Tf[o + s];
      ty += gym * Tf[o + s * sy];
      tz += gzm * Tf[o + s * sz];
      cx += gxm * cpm[o + s];
      cy += gym * cpm[o + s * sy];
      cz += gzm * cpm[o + s * sz];
    }
    const int64_t step[3] = {1, sy, sz};
    // wall faces carry no flux (the z faces of a non-wall cell are never walls)
    const bool live[3][2] = {{true, !wall(g, p.mask, i + 1, j)}, {true, !wall(g, p.mask, i, j + 1)}, {true, true}};
    // per face (d, f) what does not depend on the species: the interpolation
    // and gradient weights of its two points, the correction velocity and,
    // for the molar term, w_m / W_gp,m and sum_m grad_m W_gp,m (the
    // reference's SC / W_gp then reads (w_m / W_gp,m) Y_m: no division in the
    // species loop)
    double fw0[3][2], fw1[3][2], fg0[3][2], fg1[3][2], fcv[3][2], fu0[3][2], fu1[3][2], fgw[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        // the face cell is (i, j, k) + f along d; its two points are
        // (-1, 0) for f = 0 and (0, +1) for f = 1
        const int fi = d == 0 ? i + f : i, fj = d == 1 ? j + f : j;
        const int64_t fo = o + f * step[d];
        double w0, w1;
        if (d == 0) {
          w0 = op(isx, fi, fi0, fni, fj, g.jmino_, g.nyo, -1, -1);
          w1 = op(isx, fi, fi0, fni, fj, g.jmino_, g.nyo, 0, -1);
        } else {
          const double* ws = d == 1 ? isy : isz;
          w0 = op(ws, fi, g.imino_, g.nxo, fj, fj0, fnj, -1, -1);
          w1 = op(ws, fi, g.imino_, g.nxo, fj, fj0, fnj, 0, -1);
        }
        const double* gs = d == 0 ? grx : (d == 1 ? gry : grz);
        const double g0 = op(gs, fi, fi0, fni, fj, fj0, fnj, -1, -1);
        const double g1 = op(gs, fi, fi0, fni, fj, fj0, fnj, 0, -1);
        fw0[d][f] = w0;
        fw1[d][f] = w1;
        fg0[d][f] = g0;
        fg1[d][f] = g1;
        fcv[d][f] = -cvel[(int64_t)d * g.plane + fo];
        if (p.molar) {
          const double wa = wgp[fo - step[d]], wb = wgp[fo];
          fu0[d][f] = w0 / wa;
          fu1[d][f] = w1 / wb;
          fgw[d][f] = g0 * wa + g1 * wb;
        } else {
          fu0[d][f] = fu1[d][f] = fgw[d][f] = 0.0;
        }
      }
    }
    double eT = 0.0;  // temperature term before the cp-gradient part and / cp
#pragma unroll 1
    for (int sp = 0; sp < N; ++sp) {
      const double* Y = sc + (int64_t)(g.isc_sc0 + sp) * g.plane;
      const double* D = df + (int64_t)(g.isc_sc0 + sp) * g.plane;
      const double yc = Y[o], dc = D[o];
      double dcor = 0.0, dfick = 0.0, fm[3] = {0.0, 0.0, 0.0};
      double fc[3][2], ff[3][2];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double ym = Y[o - step[d]], yp = Y[o + step[d]], dm = D[o - step[d]], dp = D[o + step[d]];
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          const double y0 = f ? yc : ym, y1 = f ? yp : yc;
          const double d0 = f ? dc : dm, d1 = f ? dp : dc;
          const double w0 = fw0[d][f], w1 = fw1[d][f], g0 = fg0[d][f], g1 = fg1[d][f];
          double syf = 0, a = 0, b = 0;
          syf += w0 * y0;
          a += w0 * d0;
          b += g0 * y0;
          syf += w1 * y1;
          a += w1 * d1;
          b += g1 * y1;
          double c = fcv[d][f] * syf;
          if (p.molar) c += a * (fu0[d][f] * y0 + fu1[d][f] * y1) * fgw[d][f];
          fc[d][f] = live[d][f] ? c : 0.0;
          ff[d][f] = live[d][f] ? a * b : 0.0;
        }
      }
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const double du = op(dvu, i, g.imin_, nx, j, g.jmin_, ny, f, 0);
        const double dv = op(dvv, i, g.imin_, nx, j, g.jmin_, ny, f, 0);
        const double dw = op(dvw, i, g.imin_, nx, j, g.jmin_, ny, f, 0);
        dcor += du * fc[0][f] + dv * fc[1][f] + dw * fc[2][f];
        dfick += du * ff[0][f] + dv * ff[1][f] + dw * ff[2][f];
        fm[0] += op(imx, i, g.imin_, nx, j, g.jmin_, ny, f, 0) * (fc[0][f] + ff[0][f]);
        fm[1] += op(imy, i, g.imin_, nx, j, g.jmin_, ny, f, 0) * (fc[1][f] + ff[1][f]);
        fm[2] += op(imz, i, g.imin_, nx, j, g.jmin_, ny, f, 0) * (fc[2][f] + ff[2][f]);
      }
      double cpk;
      cp_species(tab, L, p.mech.tmid, T, sp, cpk);
      eT = eT + cpk * fm[0] * tx;
      eT = eT + cpk * fm[1] * ty;
      eT = eT + cpk * fm[2] * tz;
      if (corr) corr[(int64_t)sp * g.plane + o] = dcor * p.dt;
      if (src) src[(int64_t)(g.isc_sc0 + sp) * g.plane + o] += dcor * p.dt;
      if (dgp) {
        const int slot = sp == p.i_h2o ? 0 : (sp == p.i_co ? 1 : (sp == p.i_co2 ? 2 : -1));
        if (slot >= 0) dgp[(int64_t)slot * g.plane + o] = dfick + dcor;
      }
    }
    if (p.gasrad) eT += gas_radiation(p, T, o);
    const double dT = df[(int64_t)g.isc_T0 * g.plane + o];
    eT = eT + dT * cx * tx;
    eT = eT + dT * cy * ty;
    eT = eT + dT * cz * tz;
    eT = eT / cpm[o] * p.dt;
    if (corr) corr[(int64_t)N * g.plane + o] = eT;
    if (src) src[(int64_t)g.isc_T0 * g.plane + o] += eT;
  }
}

}  // namespace tgr

using namespace tgr;

extern "C" int tgr_launch_diffusion(const DiffParams* p, int neq, hipStream_t s) {
  const DiffGeom& g = p->gm;
  const int64_t ncell = (int64_t)g.nxo * g.nyo * g.nzo;
  const int st1 = p->m.st1, st2 = p->m.st2;
  const int64_t nface = (int64_t)(g.imax_ - g.imin_ + 1 + st1 + st2) * (g.jmax_ - g.jmin_ + 1 + st1 + st2) *
                        (g.kmax_ - g.kmin_ + 1 + st1 + st2);
  const int64_t nint = (int64_t)(g.imax_ - g.imin_ + 1) * (g.jmax_ - g.jmin_ + 1) * (g.kmax_ - g.kmin_ + 1);
  auto blocks = [](int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65535); };
  // second order (the reference's default stencil) has its own pair of
  // kernels; other orders take the general ones
  const bool o2 = st1 == 0 && st2 == 1;
#define TGR_DIFF_LAUNCH(NEQ_)                                                                  \
  hipLaunchKernelGGL(diff_cell_kernel<NEQ_>, dim3(blocks(ncell)), dim3(256), 0, s, *p);       \
  if (o2) {                                                                                     \
    hipLaunchKernelGGL(diff_face2_kernel<NEQ_>, dim3(blocks(nface)), dim3(256), 0, s, *p);    \
    hipLaunchKernelGGL(diff_main2_kernel<NEQ_>, dim3(blocks(nint)), dim3(256), 0, s, *p);     \
  } else {                                                                                      \
    hipLaunchKernelGGL(diff_face_kernel<NEQ_>, dim3(blocks(nface)), dim3(256), 0, s, *p);     \
    hipLaunchKernelGGL(diff_main_kernel<NEQ_>, dim3(blocks(nint)), dim3(256), 0, s, *p);      \
  }
  switch (neq) {
    case 10: TGR_DIFF_LAUNCH(10) break;
    case 37: TGR_DIFF_LAUNCH(37) break;
    case 40: TGR_DIFF_LAUNCH(40) break;
    case 46: TGR_DIFF_LAUNCH(46) break;
    default: return -1;
  }
#undef TGR_DIFF_LAUNCH
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

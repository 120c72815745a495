/* CPU oracle, part 4: the molar-diffusion correction of finite-rate
 * chemistry.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates finitechem_diffusion (reference combustion/finitechem.f90:538-923)
 * loop for loop, with the metric operators of metric_generic
 * (metric/metric_generic.f90:46-84, stencil widths st2 = order/2,
 * st1 = st2 - 1, stp = st1 + st2, :683-685) passed in by the caller:
 *   W_gp       = finitechem_W of every cell (574-597)
 *   corr_vel   = sum_k interp(DIFF_k) grad(Y_k) [+ molar: interp(DIFF_k)
 *                interp(Y_k / W_gp) grad(W_gp)] at the faces   (599-640)
 *   F_k        = -corr_vel interp(Y_k) [+ the molar term]       (642-681)
 *   diff_corr_k = div F_k                                        (683-694)
 *   F_k       += interp(DIFF_k) grad(Y_k)  (total flux)          (696-711)
 *   diff_corr_T += cp_k(T) interp_m(F_k) grad_m(T)               (713-735)
 *   gas radiation (optional)                                      (738-792)
 *   cp (cv) of the mixture; the cp-gradient term; / cp           (804-840)
 *   diff_gp(1:3) = div(Fickian flux) + diff_corr of H2O, CO, CO2 (863-916)
 *   diff_corr *= dt                                               (920)
 * Readings chosen where the reference text does not conform (the same as the
 * HIP kernels'):
 *   * 652: SC(i-st2:j+st1, ...) is read SC(i-st2:i+st1, ...);
 *   * 634-635 and 706: the z sections k-st2:k-st1 are read k-st2:k+st1 (the
 *     x and y lines' form; identical at second order, st1 = 0);
 *   * 720-730: Cp_sp(isc) is read as the species' own cp (identical when the
 *     species block starts at scalar 1);
 *   * gas radiation: a species the mechanism lacks contributes 0.
 * Not restated: the burner heat loss (796-802: fixed cell indices of one
 * case's mesh) and the external ignition source (842-859: a region of one
 * case's geometry).
 * Face arrays are 0 where mask = 1 (never assigned there in the reference,
 * whose F and corr_vel start at 0).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_int.h"

typedef struct {
  const orc_grid *g;
  int nxo, nyo, nzo, mx;
  size_t plane;
} ctx_t;

static size_t c3(const ctx_t *c, int i, int j, int k) {
  const orc_grid *g = c->g;
  return (size_t)(i - g->imino_) + (size_t)c->nxo * ((size_t)(j - g->jmino_) + (size_t)c->nyo * (k - g->kmino_));
}

static int wall(const ctx_t *c, const int *mask, int i, int j) {
  return mask && mask[(i - c->g->imino) + c->mx * (j - c->g->jmino)] == 1;
}

/* operator coefficient (i, j, st): first index from i0, second from j0, st from s0 */
static double op(const double *a, int i, int i0, int ni, int j, int j0, int nj, int st, int s0) {
  return a[(size_t)(i - i0) + (size_t)ni * ((size_t)(j - j0) + (size_t)nj * (st - s0))];
}

static int sp_index(const tgr_mech_tables *t, const char *name) {
  for (int k = 0; k < t->nsp; ++k)
    if (strcmp(t->species[k], name) == 0) return k;
  return -1;
}

int orc_diffusion(int mech, const orc_grid *g, const orc_metric *m, double dt, int molar, int gasrad, int cv,
                  const double *RHO_, const double *SC_, const double *DIFF_, const int *mask, double *diff_corr,
                  double *diff_gp) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t) return -1;
  const int N = t->nsp, st1 = m->st1, st2 = m->st2, stp = st1 + st2;
  ctx_t c = {g, g->imaxo_ - g->imino_ + 1, g->jmaxo_ - g->jmino_ + 1, g->kmaxo_ - g->kmino_ + 1,
             g->imaxo - g->imino + 1, 0};
  c.plane = (size_t)c.nxo * c.nyo * c.nzo;
  const size_t P = c.plane;
  const double R_cst = 8314.34; /* finitechem.f90:15 */
  /* operator extents */
  const int fi0 = g->imin_ - st1, fni = g->imax_ + st2 - fi0 + 1;   /* face i range */
  const int fj0 = g->jmin_ - st1, fnj = g->jmax_ + st2 - fj0 + 1;   /* face j range */
  const int oi0 = g->imino_, oni = c.nxo, oj0 = g->jmino_, onj = c.nyo;
  const int ii0 = g->imin_, ini = g->imax_ - g->imin_ + 1, ij0 = g->jmin_, inj = g->jmax_ - g->jmin_ + 1;
#define ISX(i, j, s) op(m->interp_sc_x, i, fi0, fni, j, oj0, onj, s, -st2)
#define ISY(i, j, s) op(m->interp_sc_y, i, oi0, oni, j, fj0, fnj, s, -st2)
#define ISZ(i, j, s) op(m->interp_sc_z, i, oi0, oni, j, fj0, fnj, s, -st2)
#define GX(i, j, s) op(m->grad_x, i, fi0, fni, j, fj0, fnj, s, -st2)
#define GY(i, j, s) op(m->grad_y, i, fi0, fni, j, fj0, fnj, s, -st2)
#define GZ(i, j, s) op(m->grad_z, i, fi0, fni, j, fj0, fnj, s, -st2)
#define DU(i, j, s) op(m->div_u, i, ii0, ini, j, ij0, inj, s, -st1)
#define DV(i, j, s) op(m->div_v, i, ii0, ini, j, ij0, inj, s, -st1)
#define DW(i, j, s) op(m->div_w, i, ii0, ini, j, ij0, inj, s, -st1)
#define IUM(i, j, s) op(m->interp_u_xm, i, ii0, ini, j, ij0, inj, s, -st1)
#define IVM(i, j, s) op(m->interp_v_ym, i, ii0, ini, j, ij0, inj, s, -st1)
#define IWM(i, j, s) op(m->interp_w_zm, i, ii0, ini, j, ij0, inj, s, -st1)
#define GXM(i, j, s) op(m->grad_xm, i, ii0, ini, j, ij0, inj, s, -stp)
#define GYM(i, j, s) op(m->grad_ym, i, ii0, ini, j, ij0, inj, s, -stp)
#define GZM(i, j, s) op(m->grad_zm, i, ii0, ini, j, ij0, inj, s, -stp)
  const int isc_sc = g->isc_sc, isc_T = g->isc_T;
#define SC(i, j, k, s) SC_[c3(&c, i, j, k) + (size_t)((s) - 1) * P]
#define DIFF(i, j, k, s) DIFF_[c3(&c, i, j, k) + (size_t)((s) - 1) * P]
#define DC(i, j, k, n) diff_corr[c3(&c, i, j, k) + (size_t)((n) - 1) * P]
  double *W_gp = malloc(sizeof(double) * P), *Cp_mix = malloc(sizeof(double) * P);
  double *cvx = calloc(P, sizeof(double)), *cvy = calloc(P, sizeof(double)), *cvz = calloc(P, sizeof(double));
  double *FX = calloc(P, sizeof(double)), *FY = calloc(P, sizeof(double)), *FZ = calloc(P, sizeof(double));
  if (!W_gp || !Cp_mix || !cvx || !cvy || !cvz || !FX || !FY || !FZ) return -2;
  memset(diff_corr, 0, sizeof(double) * P * (N + 1));
  memset(diff_gp, 0, sizeof(double) * P * 3);
  double h[64], cp[64];
  /* mixture molar mass of every cell (574-597) */
  for (int k = g->kmino_; k <= g->kmaxo_; ++k)
    for (int j = g->jmino_; j <= g->jmaxo_; ++j)
      for (int i = g->imino_; i <= g->imaxo_; ++i) {
        double w = 1.0;
        if (!wall(&c, mask, i, j)) {
          double s1 = 0.0, s2 = 0.0;
          for (int q = 0; q < N; ++q) {
            const double y = fmin(fmax(SC(i, j, k, isc_sc + q), 0.0), 1.0);
            s1 += y;
            s2 += y / t->W[q];
          }
          w = s1 / s2;
        }
        W_gp[c3(&c, i, j, k)] = w;
      }
#define WG(i, j, k) W_gp[c3(&c, i, j, k)]
  /* correction velocities (599-640) */
  for (int isc = isc_sc; isc < isc_sc + N; ++isc)
    for (int k = g->kmin_ - st1; k <= g->kmax_ + st2; ++k)
      for (int j = g->jmin_ - st1; j <= g->jmax_ + st2; ++j)
        for (int i = g->imin_ - st1; i <= g->imax_ + st2; ++i) {
          if (wall(&c, mask, i, j)) continue;
          double a = 0, b = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
          for (int s = -st2; s <= st1; ++s) {
            a += ISX(i, j, s) * DIFF(i + s, j, k, isc);
            b += GX(i, j, s) * SC(i + s, j, k, isc);
            a2 += ISY(i, j, s) * DIFF(i, j + s, k, isc);
            b2 += GY(i, j, s) * SC(i, j + s, k, isc);
            a3 += ISZ(i, j, s) * DIFF(i, j, k + s, isc);
            b3 += GZ(i, j, s) * SC(i, j, k + s, isc);
          }
          const size_t o = c3(&c, i, j, k);
          cvx[o] += a * b;
          cvy[o] += a2 * b2;
          cvz[o] += a3 * b3;
          if (molar) {
            double ix = 0, gx = 0, iy = 0, gy = 0, iz = 0, gz = 0;
            for (int s = -st2; s <= st1; ++s) {
              ix += ISX(i, j, s) * (SC(i + s, j, k, isc) / WG(i + s, j, k));
              gx += GX(i, j, s) * WG(i + s, j, k);
              iy += ISY(i, j, s) * (SC(i, j + s, k, isc) / WG(i, j + s, k));
              gy += GY(i, j, s) * WG(i, j + s, k);
              iz += ISZ(i, j, s) * (SC(i, j, k + s, isc) / WG(i, j, k + s));
              gz += GZ(i, j, s) * WG(i, j, k + s);
            }
            cvx[o] += a * ix * gx;
            cvy[o] += a2 * iy * gy;
            cvz[o] += a3 * iz * gz;
          }
        }
  /* per species: fluxes, species term, total flux, enthalpy flux */
  for (int isc = isc_sc; isc < isc_sc + N; ++isc) {
    const int ks = isc - isc_sc;
    for (int k = g->kmin_ - st1; k <= g->kmax_ + st2; ++k)
      for (int j = g->jmin_ - st1; j <= g->jmax_ + st2; ++j)
        for (int i = g->imin_ - st1; i <= g->imax_ + st2; ++i) {
          if (wall(&c, mask, i, j)) continue;
          const size_t o = c3(&c, i, j, k);
          double sx = 0, sy = 0, sz = 0;
          for (int s = -st2; s <= st1; ++s) {
            sx += ISX(i, j, s) * SC(i + s, j, k, isc);
            sy += ISY(i, j, s) * SC(i, j + s, k, isc);
            sz += ISZ(i, j, s) * SC(i, j, k + s, isc);
          }
          FX[o] = -cvx[o] * sx;
          FY[o] = -cvy[o] * sy;
          FZ[o] = -cvz[o] * sz;
          if (molar) {
            double a = 0, ix = 0, gx = 0, a2 = 0, iy = 0, gy = 0, a3 = 0, iz = 0, gz = 0;
            for (int s = -st2; s <= st1; ++s) {
              a += ISX(i, j, s) * DIFF(i + s, j, k, isc);
              ix += ISX(i, j, s) * (SC(i + s, j, k, isc) / WG(i + s, j, k));
              gx += GX(i, j, s) * WG(i + s, j, k);
              a2 += ISY(i, j, s) * DIFF(i, j + s, k, isc);
              iy += ISY(i, j, s) * (SC(i, j + s, k, isc) / WG(i, j + s, k));
              gy += GY(i, j, s) * WG(i, j + s, k);
              a3 += ISZ(i, j, s) * DIFF(i, j, k + s, isc);
              iz += ISZ(i, j, s) * (SC(i, j, k + s, isc) / WG(i, j, k + s));
              gz += GZ(i, j, s) * WG(i, j, k + s);
            }
            FX[o] += a * ix * gx;
            FY[o] += a2 * iy * gy;
            FZ[o] += a3 * iz * gz;
          }
        }
    for (int k = g->kmin_; k <= g->kmax_; ++k)
      for (int j = g->jmin_; j <= g->jmax_; ++j)
        for (int i = g->imin_; i <= g->imax_; ++i) {
          if (wall(&c, mask, i, j)) continue;
          double dx = 0, dy = 0, dz = 0;
          for (int s = -st1; s <= st2; ++s) {
            dx += DU(i, j, s) * FX[c3(&c, i + s, j, k)];
            dy += DV(i, j, s) * FY[c3(&c, i, j + s, k)];
            dz += DW(i, j, s) * FZ[c3(&c, i, j, k + s)];
          }
          DC(i, j, k, ks + 1) = dx + dy + dz;
        }
    for (int k = g->kmin_ - st1; k <= g->kmax_ + st2; ++k)
      for (int j = g->jmin_ - st1; j <= g->jmax_ + st2; ++j)
        for (int i = g->imin_ - st1; i <= g->imax_ + st2; ++i) {
          if (wall(&c, mask, i, j)) continue;
          double a = 0, b = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
          for (int s = -st2; s <= st1; ++s) {
            a += ISX(i, j, s) * DIFF(i + s, j, k, isc);
            b += GX(i, j, s) * SC(i + s, j, k, isc);
            a2 += ISY(i, j, s) * DIFF(i, j + s, k, isc);
            b2 += GY(i, j, s) * SC(i, j + s, k, isc);
            a3 += ISZ(i, j, s) * DIFF(i, j, k + s, isc);
            b3 += GZ(i, j, s) * SC(i, j, k + s, isc);
          }
          const size_t o = c3(&c, i, j, k);
          FX[o] += a * b;
          FY[o] += a2 * b2;
          FZ[o] += a3 * b3;
        }
    for (int k = g->kmin_; k <= g->kmax_; ++k)
      for (int j = g->jmin_; j <= g->jmax_; ++j)
        for (int i = g->imin_; i <= g->imax_; ++i) {
          if (wall(&c, mask, i, j)) continue;
          orc_thermo_t(t, SC(i, j, k, isc_T), h, cp);
          double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
          for (int s = -st1; s <= st2; ++s) {
            fx += IUM(i, j, s) * FX[c3(&c, i + s, j, k)];
            fy += IVM(i, j, s) * FY[c3(&c, i, j + s, k)];
            fz += IWM(i, j, s) * FZ[c3(&c, i, j, k + s)];
          }
          for (int s = -stp; s <= stp; ++s) {
            tx += GXM(i, j, s) * SC(i + s, j, k, isc_T);
            ty += GYM(i, j, s) * SC(i, j + s, k, isc_T);
            tz += GZM(i, j, s) * SC(i, j, k + s, isc_T);
          }
          DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + cp[ks] * fx * tx;
          DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + cp[ks] * fy * ty;
          DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + cp[ks] * fz * tz;
        }
  }
  /* gas radiation (738-792) */
  if (gasrad) {
    const double Tb = 298.0;
    const double radconst = 4.0 * 5.67051e-8 * R_cst / 101320.0;
    const int iH2O = sp_index(t, "H2O"), iCO2 = sp_index(t, "CO2"), iCH4 = sp_index(t, "CH4"),
              iCO = sp_index(t, "CO");
    for (int k = g->kmino_; k <= g->kmaxo_; ++k)
      for (int j = g->jmino_; j <= g->jmaxo_; ++j)
        for (int i = g->imino_; i <= g->imaxo_; ++i) {
          if (wall(&c, mask, i, j)) continue;
          const double T = SC(i, j, k, isc_T);
          const double tm1 = 1000.0 / T;
          const double T2 = pow(T, 2.0);
          const double aH2O = -0.23093 + tm1 * (-1.12390 + tm1 * (9.41530 + tm1 * (-2.99880 + tm1 * (0.51382 +
                              tm1 * -1.86840E-05))));
          const double aCO2 = 18.741 + tm1 * (-121.310 + tm1 * (273.500 + tm1 * (-194.050 + tm1 * (56.310 +
                              tm1 * -5.8169))));
          const double aCH4 = 6.6334 - 0.0035686 * T + 1.6682e-08 * T2 + 2.5611e-10 * T2 * T - 2.6558e-14 * T2 * T2;
          double c0, c1, c2, cc3, c4;
          if (T <= 750.0) {
            c0 = 4.7869, c1 = -0.06953, c2 = 2.95775e-4, cc3 = -4.25732e-7, c4 = 2.02894e-10;
          } else {
            c0 = 10.09, c1 = -0.01183, c2 = 4.7753e-6, cc3 = -5.87209e-10, c4 = -2.5334e-14;
          }
          const double aCO = c0 + T * (c1 + T * (c2 + T * (cc3 + T * c4)));
          double sum = 0.0;
          if (iCO2 >= 0) sum += aCO2 * SC(i, j, k, isc_sc + iCO2) / t->W[iCO2];
          if (iH2O >= 0) sum += aH2O * SC(i, j, k, isc_sc + iH2O) / t->W[iH2O];
          if (iCH4 >= 0) sum += aCH4 * SC(i, j, k, isc_sc + iCH4) / t->W[iCH4];
          if (iCO >= 0) sum += aCO * SC(i, j, k, isc_sc + iCO) / t->W[iCO];
          const double Qrad = -radconst * RHO_[c3(&c, i, j, k)] * T * (pow(T, 4.0) - pow(Tb, 4.0)) * sum;
          if (i >= g->imin_ && i <= g->imax_ && j >= g->jmin_ && j <= g->jmax_ && k >= g->kmin_ && k <= g->kmax_)
            DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + Qrad;
        }
  }
  /* mixture heat capacity (cv for compressible / constant volume) (804-815) */
  for (int k = g->kmino_; k <= g->kmaxo_; ++k)
    for (int j = g->jmino_; j <= g->jmaxo_; ++j)
      for (int i = g->imino_; i <= g->imaxo_; ++i) {
        double v = 0.0;
        if (!wall(&c, mask, i, j)) {
          orc_thermo_t(t, SC(i, j, k, isc_T), h, cp);
          for (int q = 0; q < N; ++q) v += SC(i, j, k, isc_sc + q) * cp[q];
          if (cv) v = v - R_cst / WG(i, j, k);
        }
        Cp_mix[c3(&c, i, j, k)] = v;
      }
  /* cp-gradient term, then / cp (817-840) */
  for (int k = g->kmin_; k <= g->kmax_; ++k)
    for (int j = g->jmin_; j <= g->jmax_; ++j)
      for (int i = g->imin_; i <= g->imax_; ++i) {
        if (wall(&c, mask, i, j)) continue;
        double cx = 0, tx = 0, cy = 0, ty = 0, cz = 0, tz = 0;
        for (int s = -stp; s <= stp; ++s) {
          cx += GXM(i, j, s) * Cp_mix[c3(&c, i + s, j, k)];
          tx += GXM(i, j, s) * SC(i + s, j, k, isc_T);
          cy += GYM(i, j, s) * Cp_mix[c3(&c, i, j + s, k)];
          ty += GYM(i, j, s) * SC(i, j + s, k, isc_T);
          cz += GZM(i, j, s) * Cp_mix[c3(&c, i, j, k + s)];
          tz += GZM(i, j, s) * SC(i, j, k + s, isc_T);
        }
        const double dT = DIFF(i, j, k, isc_T);
        DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + dT * cx * tx;
        DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + dT * cy * ty;
        DC(i, j, k, N + 1) = DC(i, j, k, N + 1) + dT * cz * tz;
        DC(i, j, k, N + 1) = DC(i, j, k, N + 1) / Cp_mix[c3(&c, i, j, k)];
      }
  /* H2O, CO, CO2: Fickian divergence plus the correction (863-916) */
  const char *names[3] = {"H2O", "CO", "CO2"};
  for (int q = 0; q < 3; ++q) {
    const int ks = sp_index(t, names[q]);
    if (ks < 0) continue;
    const int isc = isc_sc + ks;
    memset(cvx, 0, sizeof(double) * P);
    memset(cvy, 0, sizeof(double) * P);
    memset(cvz, 0, sizeof(double) * P);
    for (int k = g->kmin_ - st1; k <= g->kmax_ + st2; ++k)
      for (int j = g->jmin_ - st1; j <= g->jmax_ + st2; ++j)
        for (int i = g->imin_ - st1; i <= g->imax_ + st2; ++i) {
          if (wall(&c, mask, i, j)) continue;
          double a = 0, b = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
          for (int s = -st2; s <= st1; ++s) {
            a += ISX(i, j, s) * DIFF(i + s, j, k, isc);
            b += GX(i, j, s) * SC(i + s, j, k, isc);
            a2 += ISY(i, j, s) * DIFF(i, j + s, k, isc);
            b2 += GY(i, j, s) * SC(i, j + s, k, isc);
            a3 += ISZ(i, j, s) * DIFF(i, j, k + s, isc);
            b3 += GZ(i, j, s) * SC(i, j, k + s, isc);
          }
          const size_t o = c3(&c, i, j, k);
          cvx[o] = a * b;
          cvy[o] = a2 * b2;
          cvz[o] = a3 * b3;
        }
    for (int k = g->kmin_; k <= g->kmax_; ++k)
      for (int j = g->jmin_; j <= g->jmax_; ++j)
        for (int i = g->imin_; i <= g->imax_; ++i) {
          if (wall(&c, mask, i, j)) continue;
          double dx = 0, dy = 0, dz = 0;
          for (int s = -st1; s <= st2; ++s) {
            dx += DU(i, j, s) * cvx[c3(&c, i + s, j, k)];
            dy += DV(i, j, s) * cvy[c3(&c, i, j + s, k)];
            dz += DW(i, j, s) * cvz[c3(&c, i, j, k + s)];
          }
          diff_gp[c3(&c, i, j, k) + (size_t)q * P] = dx + dy + dz + DC(i, j, k, ks + 1);
        }
  }
  /* diff_corr * dt (920) */
  for (size_t o = 0; o < P * (N + 1); ++o) diff_corr[o] *= dt;
  free(W_gp);
  free(Cp_mix);
  free(cvx);
  free(cvy);
  free(cvz);
  free(FX);
  free(FY);
  free(FZ);
  return 0;
}

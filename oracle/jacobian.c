/* CPU oracle, part 3: analytic Jacobian of the chemistry right-hand side.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no Jacobian routine: CVODE builds the dense Jacobian of
 * finitechem_compute_rhs (combustion/finitechem.f90:93-175) by difference
 * quotients (FCVDENSE without FCVDENSESETJAC, finitechem.f90:1028).  This is
 * the exact derivative of the same function, differentiated term by term
 * from the PRODRATES / GETLINDRATECOEFF / COMPTHERMODATA forms the oracle
 * restates in mech_eval.c (HYDROGENF.f90:4-316, 514-673 and the same
 * routines of grimech30F.f90 / DME39F.f90 / EthyRedF.f90):
 *
 *   k(T)        = A exp(b ln T + e/(R T))           dk/dT = k (b - e/(R T)) / T
 *   q_r         = (k_f prod C_reac - k_b prod C_prod) [* M_r]      (third body)
 *   fall-off    k = k_inf F Pr / (1 + Pr),  Pr = k_0 conc / max(k_inf, 1e-60),
 *               conc = M_r if M_r > 0 else P / (R T),
 *               ln F = ln Fc / (1 + x^2),  x = L / (n - d L),
 *               L = log10 Pr + c,  c = -0.4 - 0.67 log10 Fc,  n = 0.75 - 1.27 log10 Fc,
 *               d ln F = d ln Fc / (1 + x^2) - ln Fc 2 x dx / (1 + x^2)^2,
 *               dx = (n dL - L dn) / (n - d L)^2
 *   M_r         = sum_k eff_rk C_k                  dM_r/dC_k = eff_rk
 *   rhs_k       = W_k cdot_k(C, T),  C_k = y_k / W_k,  T = y_T / rho,  rho = sum_k y_k
 *   rhs_T       = -sum_k h_k rhs_k / cp_mix,  cp_mix = sum_k (y_k / rho) cp_k(T)
 *
 * Pinned by central differences of orc_rhs (tests/test_oracle_jacobian.py). */
#include <math.h>
#include <string.h>

#include "oracle_int.h"

#define F4(x) ((double)(x##f))
#define LN10 2.302585092994045684

/* k and its derivatives for one fall-off rate: dk_dT at fixed conc (through
 * k_inf, k_0 and Fc), dk_dconc at fixed T */
static double troe_d(double kinf, double dlkinf, double k0, double dlk0, double fc, double dfc, double conc,
                     double *dk_dT, double *dk_dconc, double dlconc) {
  const double ntmp = 0.75 - F4(1.27) * log10(fc);
  const double ccoeff = -F4(0.4) - F4(0.67) * log10(fc);
  const double dcoeff = F4(0.14);
  const double kin = fmax(kinf, 1.0e-60);
  const double pr = k0 * conc / kin;
  const double L = log10(pr) + ccoeff;
  const double D = ntmp - dcoeff * L;
  const double x = L / D;
  const double lnfc = log(fc);
  const double ix2 = 1.0 / (1.0 + x * x);
  const double F = exp(lnfc * ix2);
  const double k = kinf * F * pr / (1.0 + pr);
  /* d ln Pr / dT (k_inf below the clamp does not enter Pr) */
  const double dlpr = dlk0 + dlconc - (kinf > 1.0e-60 ? dlkinf : 0.0);
  const double dlfc = dfc / fc;
  const double dL = dlpr / LN10 - F4(0.67) * dlfc / LN10;
  const double dn = -F4(1.27) * dlfc / LN10;
  const double dx = (ntmp * dL - L * dn) / (D * D);
  const double dlF = dlfc * ix2 - lnfc * 2.0 * x * dx * ix2 * ix2;
  *dk_dT = k * (dlkinf + dlF + dlpr / (1.0 + pr));
  /* at fixed T: dk/dPr = k_inf F (dlnF/dlnPr / (1 + Pr) + 1 / (1 + Pr)^2), dPr/dconc = k0 / kin */
  const double dxp = ntmp / (LN10 * D * D);
  const double dlFp = -lnfc * 2.0 * x * dxp * ix2 * ix2;
  const double dkdpr = kinf * F * (dlFp / (1.0 + pr) + 1.0 / ((1.0 + pr) * (1.0 + pr)));
  *dk_dconc = dkdpr * k0 / kin;
  return k;
}

static double arrh_d(const double *a, double lt, double rt, double T, double *dl) {
  *dl = (a[1] - a[2] / rt) / T;
  return a[0] * exp(a[1] * lt + a[2] / rt);
}

/* cdot[nsp]; dcdC[j * nsp + k] = d cdot_k / d C_j; dcdT[k] = d cdot_k / dT at fixed C */
int orc_dprodrates_t(const tgr_mech_tables *t, const double *C, double T, double P, double *cdot, double *dcdC,
                     double *dcdT) {
  const int nsp = t->nsp;
  double M[64], dq[64];
  const double lt = log(T);
  const double rt = F4(8314.34) * T;
  for (int j = 0; j < t->n_third; ++j) {
    double s = 0.0;
    for (int k = 0; k < nsp; ++k) {
      const double e = t->eff[j * nsp + k];
      if (e != 0.0) s += e * C[k];
    }
    M[j] = s;
  }
  for (int k = 0; k < nsp; ++k) cdot[k] = dcdT[k] = 0.0;
  memset(dcdC, 0, sizeof(double) * nsp * nsp);
  for (int i = 0; i < t->nrxn; ++i) {
    const tgr_rxn *r = &t->rxn[i];
    double kf, kb = 0.0, dkf = 0.0, dkb = 0.0;    /* d/dT */
    double dkfc = 0.0, dkbc = 0.0;                /* d/dconc (fall-off) */
    int mdep = -1;                                /* third-body sum the rate depends on */
    if (r->kind == 2) {
      double fc = 0.0, dfc = 0.0;
      for (int q = 0; q < r->nfc; ++q) {
        if (r->fc_kind[q] == 0) {
          const double e = r->fc_c[q] * exp(-(T / r->fc_d[q]));
          fc += e;
          dfc += -e / r->fc_d[q];
        } else {
          const double e = r->fc_c[q] * exp(-(r->fc_d[q] / T));
          fc += e;
          dfc += e * r->fc_d[q] / (T * T);
        }
      }
      const int usem = r->m >= 0 && M[r->m] > 0.0;
      const double conc = usem ? M[r->m] : P / rt;
      const double dlconc = usem ? 0.0 : -1.0 / T;
      double dli, dl0;
      double ki = arrh_d(r->f, lt, rt, T, &dli), k0 = arrh_d(r->f0, lt, rt, T, &dl0);
      kf = troe_d(ki, dli, k0, dl0, fc, dfc, conc, &dkf, &dkfc, dlconc);
      if (r->rev) {
        ki = arrh_d(r->b, lt, rt, T, &dli);
        k0 = arrh_d(r->b0, lt, rt, T, &dl0);
        kb = troe_d(ki, dli, k0, dl0, fc, dfc, conc, &dkb, &dkbc, dlconc);
      }
      if (usem) mdep = r->m;
    } else {
      double dl;
      kf = arrh_d(r->f, lt, rt, T, &dl);
      dkf = kf * dl;
      if (r->rev) {
        kb = arrh_d(r->b, lt, rt, T, &dl);
        dkb = kb * dl;
      }
      if (r->kind == 1) mdep = r->m;
    }
    if (!r->rev) kb = dkb = dkbc = 0.0;
    double pf = 1.0, pb = 1.0;
    for (int q = 0; q < r->nr; ++q) pf *= C[r->reac[q]];
    for (int q = 0; q < r->np; ++q) pb *= C[r->prod[q]];
    const double mm = r->kind == 1 ? M[r->m] : 1.0;
    const double qv = (kf * pf - kb * pb) * mm;
    const double qT = (dkf * pf - dkb * pb) * mm;
    /* d q / d C_j */
    for (int j = 0; j < nsp; ++j) dq[j] = 0.0;
    for (int q = 0; q < r->nr; ++q) {
      double p = kf * mm;
      for (int s = 0; s < r->nr; ++s)
        if (s != q) p *= C[r->reac[s]];
      dq[r->reac[q]] += p;
    }
    for (int q = 0; q < r->np; ++q) {
      double p = kb * mm;
      for (int s = 0; s < r->np; ++s)
        if (s != q) p *= C[r->prod[s]];
      dq[r->prod[q]] -= p;
    }
    if (mdep >= 0) {
      const double g = r->kind == 1 ? kf * pf - kb * pb : dkfc * pf - dkbc * pb;
      for (int j = 0; j < nsp; ++j) dq[j] += g * t->eff[mdep * nsp + j];
    }
    for (int n = 0; n < r->nnu; ++n) {
      const tgr_nu *u = &t->nu[r->nu0 + n];
      cdot[u->sp] += u->nu * qv;
      dcdT[u->sp] += u->nu * qT;
      for (int j = 0; j < nsp; ++j) dcdC[j * nsp + u->sp] += u->nu * dq[j];
    }
  }
  return 0;
}

/* f[neq] = finitechem_compute_rhs(sol); J[j * neq + i] = d f_i / d sol_j */
void orc_jacobian_t(const tgr_mech_tables *t, double p_gp, const double *sol, double *f, double *J) {
  const int nsp = t->nsp, neq = nsp + 1;
  double h[64], cp[64], dcp[64], C[64], cdot[64], dcdT[64], fT[64];
  static __thread double dcdC[64 * 64];
  double rho = 0.0;
  for (int k = 0; k < nsp; ++k) rho += sol[k];
  const double T = sol[nsp] / rho;
  orc_thermo_t(t, T, h, cp);
  const double *base = t->thermo + (T > t->tmid ? 0 : nsp * 13);
  for (int k = 0; k < nsp; ++k) {
    const double *c = base + 13 * k;
    dcp[k] = c[7] * (c[9] + T * (2.0 * c[10] + T * (3.0 * c[11] + T * 4.0 * c[12])));
  }
  double cpm = 0.0, cpmT = 0.0; /* cp_mix and d cp_mix / dT at fixed Y */
  for (int k = 0; k < nsp; ++k) {
    cpm += (sol[k] / rho) * cp[k];
    cpmT += (sol[k] / rho) * dcp[k];
  }
  for (int k = 0; k < nsp; ++k) C[k] = sol[k] / t->W[k];
  orc_dprodrates_t(t, C, T, p_gp, cdot, dcdC, dcdT);
  double g = 0.0;
  for (int k = 0; k < nsp; ++k) {
    f[k] = cdot[k] * t->W[k];
    fT[k] = dcdT[k] * t->W[k]; /* d f_k / dT at fixed C */
    g += h[k] * f[k];
  }
  f[nsp] = -g / cpm;
  for (int j = 0; j < neq; ++j) {
    double *col = J + (size_t)j * neq;
    /* dT/dy_j and, for the species, the fixed-T part of df_k/dy_j */
    const double dTdy = j < nsp ? -T / rho : 1.0 / rho;
    double dg = 0.0;
    for (int k = 0; k < nsp; ++k) {
      const double fixT = j < nsp ? t->W[k] * dcdC[j * nsp + k] / t->W[j] : 0.0;
      col[k] = fixT + fT[k] * dTdy;
      dg += cp[k] * dTdy * f[k] + h[k] * col[k];
    }
    const double dcpm = (j < nsp ? (cp[j] - cpm) / rho : 0.0) + cpmT * dTdy;
    col[nsp] = -dg / cpm + g / (cpm * cpm) * dcpm;
  }
}

int orc_jacobian(int mech, double p_gp, const double *sol, double *f, double *J) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t) return -1;
  orc_jacobian_t(t, p_gp, sol, f, J);
  return 0;
}

int orc_dprodrates(int mech, const double *C, double T, double P, double *cdot, double *dcdC, double *dcdT) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t) return -1;
  return orc_dprodrates_t(t, C, T, P, cdot, dcdC, dcdT);
}

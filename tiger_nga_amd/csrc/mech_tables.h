/* Table layout of a compiled chemistry mechanism.
 *
 * One record per reaction (a forward rate slot plus, when reversible, its
 * backward slot), mirroring the generated PRODRATES of the reference
 * (e.g. combustion/HYDROGENF.f90:4-277):
 *   k   = A * exp(b*ln T + e/(R T))          R = 8314.34 (binary32 value)
 *   W_f = k_f * prod C(reac) [* M(m)]        W_b = k_b * prod C(prod) [* M(m)]
 *   CDOT(s) = sum_r nu(s,r) * (W_f - W_b)
 * Fall-off reactions (kind 2) hold k0/kinf pairs for both directions, the
 * shared Troe centre Fc = sum_i c_i * exp(-T/d_i) (fc_kind 0) or
 * c_i * exp(-d_i/T) (fc_kind 1), and the third-body sum m that
 * GETLINDRATECOEFF receives (m = -1: P/(R T)).
 * Thermo rows: [0..nsp) high range (T > tmid), [nsp..2 nsp) low range; each row
 *   a_h, c1..c5, c6, a_cp, d1..d5 with
 *   h  = a_h  * (T*(c1 + T*(c2 + T*(c3 + T*(c4 + T*c5)))) + c6)   [J/kg]
 *   cp = a_cp * (d1 + T*(d2 + T*(d3 + T*(d4 + T*d5))))           [J/kg/K]
 * Plain C so that the HIP library and the C oracle share it. */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int sp;    /* species index, 0-based */
  double nu; /* net stoichiometric coefficient (products positive) */
} tgr_nu;

typedef struct {
  int kind;       /* 0 elementary, 1 third-body (W carries M(m)), 2 fall-off */
  int rev;        /* 1: has a backward slot */
  int m;          /* third-body sum index (kind 1, 2) or -1 */
  int nr, np;     /* concentration factors of W_f / W_b (repeats allowed) */
  int nfc;        /* number of Fc terms (kind 2) */
  int reac[3], prod[3];
  int nu0, nnu;   /* range in the nu table */
  int slot_f, slot_b; /* reference rate-slot indices (0-based), slot_b -1 if irreversible */
  int lind;       /* 1: Fc == 1 exactly (Lindemann form) */
  double f[3], b[3];   /* A, b, e : forward / backward (kind 2: k_inf) */
  double f0[3], b0[3]; /* kind 2: k_0 forward / backward */
  double fc_c[3];
  int fc_kind[3];
  double fc_d[3];
} tgr_rxn;

typedef struct {
  const char *name;
  int nsp, nrxn, nslot, n_third, n_nu;
  double tmid;
  const char *const *species;
  const double *W;      /* [nsp] kg/kmol */
  const double *thermo; /* [2*nsp][13] */
  const double *eff;    /* [n_third][nsp] third-body efficiencies */
  const tgr_rxn *rxn;   /* [nrxn] */
  const tgr_nu *nu;     /* [n_nu] */
  const double *mucoeff; /* [nsp] GETMUCOEFF, kg/(m s) */
  const double *koveps;  /* [nsp] GETKOVEREPS, 1/K */
  const double *elem;    /* [TGR_NELEM][nsp] atoms of C, H, O, N, AR per molecule */
} tgr_mech_tables;

#define TGR_NELEM 5

#ifdef __cplusplus
}
#endif

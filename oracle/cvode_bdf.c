/* CPU oracle, part 2: the stiff integrator behind FCVODE -- TEST INFRASTRUCTURE ONLY.
 *
 * combustion/finitechem.f90:1014-1034 configures SUNDIALS CVODE 2.5.0 (an
 * un-vendored third-party library: Makefile.in:40-41) as: BDF, Newton
 * iteration, dense direct linear solver with internal difference-quotient
 * Jacobian (FCVDENSE without FCVDENSESETJAC), scalar rtol/atol, stop time =
 * tout = dt, MAX_NSTEPS = 10000, every other option at its default.  This file
 * restates that configuration of CVODE from its published algorithm
 * description (Hindmarsh et al., "SUNDIALS: Suite of nonlinear and
 * differential/algebraic equation solvers", ACM TOMS 31(3) 2005; CVODE v2.5
 * user guide, "Mathematical considerations"): variable-coefficient BDF of
 * order 1..5 in Nordsieck form, modified Newton on M = I - gamma J with
 * J reused up to 50 steps and M refactored every 20 steps or on a 30% gamma
 * change, WRMS error/convergence tests, step/order selection with the
 * documented bias factors, the h0 heuristic, and tstop handling.  Library
 * defaults: maxord 5, maxcor 3, nlscoef 0.1, maxnef 7, maxncf 10.
 *
 * Parity of this restatement is pinned by tests against scipy's Radau/BDF at
 * tight tolerance and by invariants, not against CVODE 2.5.0 itself, which is
 * absent from the reference tree.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_int.h"


#define QMAX_ABS 5
#define UROUND DBL_EPSILON

/* algorithm constants (CVODE defaults) */
#define MXNCF 10
#define MXNEF 7
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define NLS_MAXCOR 3
#define CRDOWN 0.3
#define DGMAX 0.3
#define RDIV 2.0
#define MSBP 20
#define MSBJ 50
#define DGMAX_J 0.2
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 1.0e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define SMALL_NST 10
#define THRESH 1.5
#define NLSCOEF 0.1
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define HIN_ITERS 4
#define FUZZ_FACTOR 100.0
#define MIN_INC_MULT 1000.0

enum { ST_OK = 0, ST_TOO_MUCH_WORK = -1, ST_TOO_MUCH_ACC = -2, ST_ERR_FAIL = -3, ST_CONV_FAIL = -4,
       ST_SETUP_FAIL = -5, ST_BAD_INPUT = -6 };
enum { NF_FIRST = 0, NF_PREV_CONV = 1, NF_PREV_ERR = 2, NF_OK = 3, NF_CONV_FAIL = 4, NF_TRY_AGAIN = 5 };
enum { CF_NONE = 0, CF_BAD_J = 1, CF_OTHER = 2 };

typedef struct {
  const tgr_mech_tables *t;
  double p_gp;
  int n, qmax, jac_an;
  double rtol, atol;
  double *zn[QMAX_ABS + 1];
  double *ewt, *acor, *y, *tempv, *ftemp, *fdq;
  double *J, *M;
  int *piv;
  double tn, h, hprime, hscale, eta, etamax, etaq, etaqm1, etaqp1;
  double rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5;
  double tq[6], l[QMAX_ABS + 1], tau[QMAX_ABS + 2];
  int q, qprime, qwait, L, jcur, max_steps;
  long nst, nstlp, nstlj;
  orc_stats *st;
} bdf;

static void f_eval(bdf *b, const double *y, double *f) { orc_rhs_t(b->t, b->p_gp, y, f); }

static double wrms(const bdf *b, const double *x) {
  double s = 0.0;
  for (int i = 0; i < b->n; ++i) {
    double v = x[i] * b->ewt[i];
    s += v * v;
  }
  return sqrt(s / b->n);
}

static void set_ewt(bdf *b, const double *y) {
  for (int i = 0; i < b->n; ++i) b->ewt[i] = 1.0 / (b->rtol * fabs(y[i]) + b->atol);
}

/* LU with partial pivoting, column-major a[j*n + i] (row i, column j) */
static int getrf(int n, double *a, int *p) {
  for (int k = 0; k < n; ++k) {
    double *ck = a + (size_t)k * n;
    int l = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(ck[i]) > fabs(ck[l])) l = i;
    p[k] = l;
    if (ck[l] == 0.0) return k + 1;
    if (l != k)
      for (int j = 0; j < n; ++j) {
        double tmp = a[(size_t)j * n + l];
        a[(size_t)j * n + l] = a[(size_t)j * n + k];
        a[(size_t)j * n + k] = tmp;
      }
    double mult = 1.0 / ck[k];
    for (int i = k + 1; i < n; ++i) ck[i] *= mult;
    for (int j = k + 1; j < n; ++j) {
      double *cj = a + (size_t)j * n;
      double akj = cj[k];
      if (akj != 0.0)
        for (int i = k + 1; i < n; ++i) cj[i] -= akj * ck[i];
    }
  }
  return 0;
}

static void getrs(int n, const double *a, const int *p, double *bv) {
  for (int k = 0; k < n; ++k) {
    int pk = p[k];
    if (pk != k) {
      double tmp = bv[k];
      bv[k] = bv[pk];
      bv[pk] = tmp;
    }
  }
  for (int k = 0; k < n - 1; ++k) {
    const double *ck = a + (size_t)k * n;
    double bk = bv[k];
    for (int i = k + 1; i < n; ++i) bv[i] -= ck[i] * bk;
  }
  for (int k = n - 1; k > 0; --k) {
    const double *ck = a + (size_t)k * n;
    bv[k] /= ck[k];
    double bk = bv[k];
    for (int i = 0; i < k; ++i) bv[i] -= ck[i] * bk;
  }
  bv[0] /= a[0];
}

/* difference-quotient Jacobian, column j perturbs y_j (CVODE dense DQ rule) */
static void dq_jac(bdf *b, double *y, const double *fy) {
  const int n = b->n;
  double srur = sqrt(UROUND);
  double fnorm = wrms(b, fy);
  double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(b->h) * UROUND * n * fnorm) : 1.0;
  for (int j = 0; j < n; ++j) {
    double yj = y[j];
    double inc = fmax(srur * fabs(yj), minInc / b->ewt[j]);
    y[j] += inc;
    f_eval(b, y, b->fdq);
    b->st->nfe_dq++;
    y[j] = yj;
    double ii = 1.0 / inc;
    double *col = b->J + (size_t)j * n;
    for (int i = 0; i < n; ++i) col[i] = ii * b->fdq[i] - ii * fy[i];
  }
}

/* linear-solver setup: returns 0 ok, 1 singular */
static int lsetup(bdf *b, int convfail) {
  const int n = b->n;
  double dgamma = fabs(b->gamma / b->gammap - 1.0);
  int jbad = (b->nst == 0) || (b->nst > b->nstlj + MSBJ) || ((convfail == CF_BAD_J) && (dgamma < DGMAX_J)) ||
             (convfail == CF_OTHER);
  if (!jbad) {
    b->jcur = 0;
  } else {
    b->st->nje++;
    b->nstlj = b->nst;
    b->jcur = 1;
    if (b->jac_an) orc_jacobian_t(b->t, b->p_gp, b->zn[0], b->fdq, b->J); /* analytic (option) */
    else dq_jac(b, b->zn[0], b->ftemp);
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double v = -b->gamma * b->J[(size_t)j * n + i];
      if (i == j) v += 1.0;
      b->M[(size_t)j * n + i] = v;
    }
  b->st->nlu++;
  return getrf(n, b->M, b->piv) ? 1 : 0;
}

/* BDF coefficients l[] and error-test constants tq[] for the current step
 * history (variable-coefficient Nordsieck BDF). */
static void set_bdf(bdf *b) {
  const int q = b->q;
  double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = b->h;
  b->l[0] = b->l[1] = 1.0;
  for (int i = 2; i <= q; ++i) b->l[i] = 0.0;
  if (q > 1) {
    for (int j = 2; j < q; ++j) {
      hsum += b->tau[j - 1];
      xi_inv = b->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) b->l[i] += b->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -b->l[1] - alpha0;
    hsum += b->tau[q - 1];
    xi_inv = b->h / hsum;
    alpha0_hat = -b->l[1] - xi_inv;
    for (int i = q; i >= 1; --i) b->l[i] += b->l[i - 1] * xistar_inv;
  }
  /* error constants */
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + q * A1;
  b->tq[2] = fabs(A1 / (alpha0 * A2));
  b->tq[5] = fabs(A2 * xistar_inv / (b->l[q] * xi_inv));
  if (b->qwait == 1) {
    if (q > 1) {
      double C = xistar_inv / b->l[q];
      double A3 = alpha0 + 1.0 / q;
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      b->tq[1] = fabs(C * Cpinv);
    } else {
      b->tq[1] = 1.0;
    }
    hsum += b->tau[q];
    xi_inv = b->h / hsum;
    double A5 = alpha0 - (1.0 / (q + 1));
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    b->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  b->tq[4] = NLSCOEF / b->tq[2];
}

static void set_coeffs(bdf *b) {
  set_bdf(b);
  b->rl1 = 1.0 / b->l[1];
  b->gamma = b->h * b->rl1;
  if (b->nst == 0) b->gammap = b->gamma;
  b->gamrat = (b->nst > 0) ? b->gamma / b->gammap : 1.0;
}

static void predict(bdf *b, double tstop) {
  const int n = b->n;
  b->tn += b->h;
  if ((b->tn - tstop) * b->h > 0.0) b->tn = tstop;
  for (int k = 1; k <= b->q; ++k)
    for (int j = b->q; j >= k; --j)
      for (int i = 0; i < n; ++i) b->zn[j - 1][i] += b->zn[j][i];
}

static void restore(bdf *b, double saved_t) {
  const int n = b->n;
  b->tn = saved_t;
  for (int k = 1; k <= b->q; ++k)
    for (int j = b->q; j >= k; --j)
      for (int i = 0; i < n; ++i) b->zn[j - 1][i] -= b->zn[j][i];
}

static void rescale(bdf *b) {
  double factor = b->eta;
  for (int j = 1; j <= b->q; ++j) {
    for (int i = 0; i < b->n; ++i) b->zn[j][i] *= factor;
    factor *= b->eta;
  }
  b->h = b->hscale * b->eta;
  b->hscale = b->h;
}

static void increase_order(bdf *b) {
  const int n = b->n, q = b->q, L = b->L;
  double l[QMAX_ABS + 2];
  for (int i = 0; i <= b->qmax + 1; ++i) l[i] = 0.0;
  double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = b->hscale;
  l[2] = 1.0;
  if (q > 1) {
    for (int j = 1; j < q; ++j) {
      hsum += b->tau[j + 1];
      double xi = hsum / b->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
      xiold = xi;
    }
  }
  double A1 = (-alpha0 - alpha1) / prod;
  /* zn[L] = A1 * (saved acor in zn[qmax]); L <= qmax here */
  for (int i = 0; i < n; ++i) b->zn[L][i] = A1 * b->zn[b->qmax][i];
  for (int j = 2; j <= q; ++j)
    for (int i = 0; i < n; ++i) b->zn[j][i] += l[j] * b->zn[L][i];
}

static void decrease_order(bdf *b) {
  const int n = b->n, q = b->q;
  double l[QMAX_ABS + 2];
  for (int i = 0; i <= b->qmax + 1; ++i) l[i] = 0.0;
  l[2] = 1.0;
  double hsum = 0.0;
  for (int j = 1; j <= q - 2; ++j) {
    hsum += b->tau[j];
    double xi = hsum / b->hscale;
    for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
  }
  for (int j = 2; j < q; ++j)
    for (int i = 0; i < n; ++i) b->zn[j][i] -= l[j] * b->zn[q][i];
}

static void adjust_order(bdf *b, int dq) {
  if (b->q == 2 && dq != 1) return;
  if (dq == 1) increase_order(b);
  else decrease_order(b);
}

static void adjust_params(bdf *b) {
  if (b->qprime != b->q) {
    adjust_order(b, b->qprime - b->q);
    b->q = b->qprime;
    b->L = b->q + 1;
    b->qwait = b->L;
  }
  rescale(b);
}

/* Newton iteration on the corrector; returns NF_OK / NF_CONV_FAIL / NF_TRY_AGAIN */
static int newton_iter(bdf *b) {
  const int n = b->n;
  int m = 0;
  double delp = 0.0;
  for (;;) {
    for (int i = 0; i < n; ++i) {
      double r = b->rl1 * b->zn[1][i] + b->acor[i];
      b->tempv[i] = b->gamma * b->ftemp[i] - r;
    }
    getrs(n, b->M, b->piv, b->tempv);
    if (b->gamrat != 1.0) {
      double s = 2.0 / (1.0 + b->gamrat);
      for (int i = 0; i < n; ++i) b->tempv[i] *= s;
    }
    b->st->nni++;
    double del = wrms(b, b->tempv);
    for (int i = 0; i < n; ++i) {
      b->acor[i] += b->tempv[i];
      b->y[i] = b->zn[0][i] + b->acor[i];
    }
    if (m > 0) b->crate = fmax(CRDOWN * b->crate, del / delp);
    double dcon = del * fmin(1.0, b->crate) / b->tq[4];
    if (dcon <= 1.0) {
      b->acnrm = (m == 0) ? del : wrms(b, b->acor);
      b->jcur = 0;
      return NF_OK;
    }
    ++m;
    if (m == NLS_MAXCOR || (m >= 2 && del > RDIV * delp)) return b->jcur ? NF_CONV_FAIL : NF_TRY_AGAIN;
    delp = del;
    f_eval(b, b->y, b->ftemp);
    b->st->nfe++;
  }
}

static int nls(bdf *b, int nflag) {
  int convfail = (nflag == NF_FIRST || nflag == NF_PREV_ERR) ? CF_NONE : CF_OTHER;
  int callSetup = (nflag == NF_PREV_CONV) || (nflag == NF_PREV_ERR) || (b->nst == 0) ||
                  (b->nst >= b->nstlp + MSBP) || (fabs(b->gamrat - 1.0) > DGMAX);
  for (;;) {
    f_eval(b, b->zn[0], b->ftemp);
    b->st->nfe++;
    if (callSetup) {
      int ier = lsetup(b, convfail);
      b->st->nsetups++;
      callSetup = 0;
      b->gamrat = b->crate = 1.0;
      b->gammap = b->gamma;
      b->nstlp = b->nst;
      if (ier) return NF_CONV_FAIL;
    }
    for (int i = 0; i < b->n; ++i) {
      b->acor[i] = 0.0;
      b->y[i] = b->zn[0][i];
    }
    int ier = newton_iter(b);
    if (ier != NF_TRY_AGAIN) return ier;
    callSetup = 1;
    convfail = CF_BAD_J;
  }
}

/* 0: do error test, 1: predict again, <0: failure */
static int handle_nflag(bdf *b, int *nflag, double saved_t, int *ncf) {
  if (*nflag == NF_OK) return 0;
  b->st->ncfn++;
  restore(b, saved_t);
  (*ncf)++;
  b->etamax = 1.0;
  if (*ncf == MXNCF) return ST_CONV_FAIL;
  b->eta = ETACF;
  *nflag = NF_PREV_CONV;
  rescale(b);
  return 1;
}

/* 0: pass, 1: try again, <0: failure */
static int error_test(bdf *b, int *nflag, double saved_t, int *nef, double *dsm_out) {
  double dsm = b->acnrm * b->tq[2];
  *dsm_out = dsm;
  if (dsm <= 1.0) return 0;
  (*nef)++;
  b->st->netf++;
  *nflag = NF_PREV_ERR;
  restore(b, saved_t);
  if (*nef == MXNEF) return ST_ERR_FAIL;
  b->etamax = 1.0;
  if (*nef <= MXNEF1) {
    b->eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / b->L) + ADDON);
    b->eta = fmax(ETAMIN, b->eta);
    if (*nef >= SMALL_NEF) b->eta = fmin(b->eta, ETAMXF);
    rescale(b);
    return 1;
  }
  if (b->q > 1) {
    b->eta = ETAMIN;
    adjust_order(b, -1);
    b->L = b->q;
    b->q--;
    b->qwait = b->L;
    rescale(b);
    return 1;
  }
  /* order 1 after repeated failures: restart from f(tn, y) */
  b->eta = ETAMIN;
  b->h *= b->eta;
  b->hscale = b->h;
  b->qwait = LONG_WAIT;
  f_eval(b, b->zn[0], b->tempv);
  b->st->nfe++;
  for (int i = 0; i < b->n; ++i) b->zn[1][i] = b->h * b->tempv[i];
  return 1;
}

static void complete_step(bdf *b) {
  const int n = b->n;
  b->nst++;
  b->st->nst++;
  for (int i = b->q; i >= 2; --i) b->tau[i] = b->tau[i - 1];
  if (b->q == 1 && b->nst > 1) b->tau[2] = b->tau[1];
  b->tau[1] = b->h;
  for (int j = 0; j <= b->q; ++j)
    for (int i = 0; i < n; ++i) b->zn[j][i] += b->l[j] * b->acor[i];
  b->qwait--;
  if (b->qwait == 1 && b->q != b->qmax) {
    memcpy(b->zn[b->qmax], b->acor, sizeof(double) * n);
    b->saved_tq5 = b->tq[5];
  }
}

static void set_eta(bdf *b) {
  if (b->eta < THRESH) {
    b->eta = 1.0;
    b->hprime = b->h;
  } else {
    b->eta = fmin(b->eta, b->etamax);
    b->hprime = b->h * b->eta;
  }
}

static void prepare_next(bdf *b, double dsm) {
  const int n = b->n;
  if (b->etamax == 1.0) {
    if (b->qwait < 2) b->qwait = 2;
    b->qprime = b->q;
    b->hprime = b->h;
    b->eta = 1.0;
    return;
  }
  b->etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / b->L) + ADDON);
  if (b->qwait != 0) {
    b->eta = b->etaq;
    b->qprime = b->q;
    set_eta(b);
    return;
  }
  b->qwait = 2;
  b->etaqm1 = 0.0;
  if (b->q > 1) {
    double ddn = wrms(b, b->zn[b->q]) * b->tq[1];
    b->etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / b->q) + ADDON);
  }
  b->etaqp1 = 0.0;
  if (b->q != b->qmax && b->saved_tq5 != 0.0) {
    double cquot = (b->tq[5] / b->saved_tq5) * pow(b->h / b->tau[2], (double)b->L);
    for (int i = 0; i < n; ++i) b->tempv[i] = b->acor[i] - cquot * b->zn[b->qmax][i];
    double dup = wrms(b, b->tempv) * b->tq[3];
    b->etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (b->L + 1)) + ADDON);
  }
  /* choose the order with the largest step ratio; ties keep q */
  double etam = fmax(b->etaqm1, fmax(b->etaq, b->etaqp1));
  if (etam < THRESH) {
    b->eta = 1.0;
    b->qprime = b->q;
  } else if (etam == b->etaq) {
    b->eta = b->etaq;
    b->qprime = b->q;
  } else if (etam == b->etaqm1) {
    b->eta = b->etaqm1;
    b->qprime = b->q - 1;
  } else {
    b->eta = b->etaqp1;
    b->qprime = b->q + 1;
    memcpy(b->zn[b->qmax], b->acor, sizeof(double) * n);
  }
  set_eta(b);
}

/* one internal step; 0 ok, <0 failure */
static int bdf_step(bdf *b, double tstop) {
  double saved_t = b->tn, dsm = 0.0;
  int ncf = 0, nef = 0, nflag = NF_FIRST;
  if (b->nst > 0 && b->hprime != b->h) adjust_params(b);
  for (;;) {
    predict(b, tstop);
    set_coeffs(b);
    nflag = nls(b, nflag);
    int k = handle_nflag(b, &nflag, saved_t, &ncf);
    if (k == 1) continue;
    if (k < 0) return k;
    int e = error_test(b, &nflag, saved_t, &nef, &dsm);
    if (e == 1) continue;
    if (e < 0) return e;
    break;
  }
  complete_step(b);
  prepare_next(b, dsm);
  b->etamax = (b->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  double s = 1.0 / b->tq[2];
  for (int i = 0; i < b->n; ++i) b->acor[i] *= s;
  return 0;
}

/* initial step size: geometric mean of bounds refined by a second-derivative
 * estimate along the forward-Euler direction */
static double initial_h(bdf *b, double tout) {
  const int n = b->n;
  double tdist = fabs(tout - b->tn);
  double tround = UROUND * fmax(fabs(b->tn), fabs(tout));
  double hlb = HLB_FACTOR * tround;
  /* upper bound from |y0'| relative to |y0| */
  double hub_inv = 0.0;
  for (int i = 0; i < n; ++i) {
    double den = HUB_FACTOR * fabs(b->zn[0][i]) + 1.0 / b->ewt[i];
    double r = fabs(b->zn[1][i]) / den;
    if (r > hub_inv) hub_inv = r;
  }
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) return hg;
  double hnew = hg;
  int hnewOK = 0;
  for (int it = 1; it <= HIN_ITERS; ++it) {
    /* |y''| estimate at hg */
    for (int i = 0; i < n; ++i) b->y[i] = hg * b->zn[1][i] + b->zn[0][i];
    f_eval(b, b->y, b->tempv);
    b->st->nfe++;
    for (int i = 0; i < n; ++i) b->tempv[i] = (b->tempv[i] - b->zn[1][i]) * (1.0 / hg);
    double yddnrm = wrms(b, b->tempv);
    if (hnewOK || it == HIN_ITERS) {
      hnew = hg;
      break;
    }
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) hnewOK = 1;
    if (it > 1 && hrat > 2.0) {
      hnew = hg;
      hnewOK = 1;
    }
    hg = hnew;
  }
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  return h0;
}

static int all_finite(const double *x, int n) {
  for (int i = 0; i < n; ++i)
    if (!isfinite(x[i])) return 0;
  return 1;
}

static int integrate_cell(bdf *b, double tout, double *sol) {
  const int n = b->n;
  if (!all_finite(sol, n)) return ST_BAD_INPUT;
  b->tn = 0.0;
  b->q = 1;
  b->L = 2;
  b->qwait = b->L;
  b->qprime = 1;
  b->etamax = ETAMX1;
  b->nst = b->nstlp = b->nstlj = 0;
  b->saved_tq5 = 0.0;
  b->crate = 1.0;
  b->gammap = b->gamrat = 1.0;
  b->jcur = 0;
  b->eta = 1.0;
  for (int j = 0; j <= QMAX_ABS + 1; ++j) b->tau[j] = 0.0;
  for (int j = 0; j <= QMAX_ABS; ++j) memset(b->zn[j], 0, sizeof(double) * n);
  memcpy(b->zn[0], sol, sizeof(double) * n);
  set_ewt(b, b->zn[0]);
  f_eval(b, b->zn[0], b->zn[1]);
  b->st->nfe++;
  double tstop = tout;
  if (tout - b->tn == 0.0 || fabs(tout - b->tn) < 2.0 * UROUND * fabs(tout)) return ST_BAD_INPUT;
  b->h = initial_h(b, tout);
  if ((b->tn + b->h - tstop) * b->h > 0.0) b->h = (tstop - b->tn) * (1.0 - 4.0 * UROUND);
  b->hscale = b->h;
  b->hprime = b->h;
  for (int i = 0; i < n; ++i) b->zn[1][i] *= b->h;

  long nstloc = 0;
  for (;;) {
    if (b->nst > 0) set_ewt(b, b->zn[0]);
    if (nstloc >= b->max_steps) {
      memcpy(sol, b->zn[0], sizeof(double) * n);
      return ST_TOO_MUCH_WORK;
    }
    if (UROUND * wrms(b, b->zn[0]) > 1.0) {
      memcpy(sol, b->zn[0], sizeof(double) * n);
      return ST_TOO_MUCH_ACC;
    }
    int k = bdf_step(b, tstop);
    if (k < 0) {
      memcpy(sol, b->zn[0], sizeof(double) * n);
      return k;
    }
    nstloc++;
    double tret = -1.0;
    if ((b->tn - tout) * b->h >= 0.0) {
      tret = tout;
    } else {
      double troundoff = FUZZ_FACTOR * UROUND * (fabs(b->tn) + fabs(b->h));
      if (fabs(b->tn - tstop) <= troundoff) tret = tstop;
    }
    if (tret >= 0.0) {
      /* interpolate the Nordsieck history at tret */
      double s = (tret - b->tn) / b->h;
      for (int i = 0; i < n; ++i) {
        double v = b->zn[b->q][i];
        for (int j = b->q - 1; j >= 0; --j) v = b->zn[j][i] + s * v;
        sol[i] = v;
      }
      return 0;
    }
    if ((b->tn + b->hprime - tstop) * b->h > 0.0) {
      b->hprime = (tstop - b->tn) * (1.0 - 4.0 * UROUND);
      b->eta = b->hprime / b->h;
    }
  }
}

typedef struct {
  bdf b;
  double *mem;
  int *ipiv;
} bdf_ws;

static int ws_init(bdf_ws *w, const tgr_mech_tables *t, const orc_opts *o, double p_gp) {
  int n = t->nsp + 1;
  memset(w, 0, sizeof(*w));
  size_t nd = (size_t)(QMAX_ABS + 1 + 6) * n + 2 * (size_t)n * n;
  w->mem = (double *)calloc(nd, sizeof(double));
  w->ipiv = (int *)calloc(n, sizeof(int));
  if (!w->mem || !w->ipiv) return -1;
  bdf *b = &w->b;
  b->t = t;
  b->p_gp = p_gp;
  b->n = n;
  b->qmax = (o->max_order >= 1 && o->max_order <= QMAX_ABS) ? o->max_order : QMAX_ABS;
  b->jac_an = o->jac == 1;
  b->rtol = o->rtol;
  b->atol = o->atol;
  b->max_steps = o->max_steps > 0 ? o->max_steps : 500;
  double *p = w->mem;
  for (int j = 0; j <= QMAX_ABS; ++j, p += n) b->zn[j] = p;
  b->ewt = p; p += n;
  b->acor = p; p += n;
  b->y = p; p += n;
  b->tempv = p; p += n;
  b->ftemp = p; p += n;
  b->fdq = p; p += n;
  b->J = p; p += (size_t)n * n;
  b->M = p;
  b->piv = w->ipiv;
  return 0;
}

static void ws_free(bdf_ws *w) {
  free(w->mem);
  free(w->ipiv);
}

int orc_integrate(int mech, const orc_opts *o, double dt, double p_gp, double *sol, orc_stats *st) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t) return -1;
  bdf_ws w;
  if (ws_init(&w, t, o, p_gp)) return -1;
  memset(st, 0, sizeof(*st));
  w.b.st = st;
  st->status = integrate_cell(&w.b, dt, sol);
  ws_free(&w);
  return 0;
}

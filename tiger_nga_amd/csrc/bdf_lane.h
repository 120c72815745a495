// Lane-per-cell CVODE-semantics BDF integrator for CDNA4 (gfx950): one thread
// integrates one cell.
//
// The same integration as the wave-per-cell solver (bdf_wave.h) and the CPU
// oracle -- the SUNDIALS CVODE 2.5.0 BDF / Newton / dense difference-quotient
// Jacobian solve that finitechem_source_chemistry drives (reference
// combustion/finitechem.f90:1014-1034) -- laid out for small mechanisms,
// where a wave per cell leaves most lanes idle (H2: 10 of 64):
//   * every vector of length neq and the Nordsieck history live in the
//     thread's registers, indexed at compile time (run-time orders select with
//     guarded unrolled loops);
//   * the iteration matrix M = I - gamma J is held in registers; the saved
//     Jacobian J (reused up to 50 steps) lives in a per-thread slice of an HBM
//     workspace laid out [element][lane], so a wave's accesses are coalesced;
//   * the right-hand side is the generated straight-line code of the
//     mechanism (LaneRhs<NEQ>, lane_rhs_<mech>.h);
//   * each thread takes its next cell from a device counter as soon as it is
//     done, so cells of different stiffness share a wave without waiting for
//     one another's whole integration.
#pragma once
#include <stdint.h>

#include "bdf_wave.h"
#include "lane_rhs.h"

namespace tgr {

// One right-hand side out of line: the solver calls it from six places, and
// one copy of the generated code keeps the kernel within the instruction
// cache.  in / out are the caller's call buffers.
template <int NEQ>
__device__ __attribute__((noinline)) void lane_rhs_call(const double* in, double* out, double p_gp) {
  double s[NEQ], o[NEQ];
#pragma unroll
  for (int k = 0; k < NEQ; ++k) s[k] = in[k];
  LaneRhs<NEQ>::eval(s, p_gp, o);
#pragma unroll
  for (int k = 0; k < NEQ; ++k) out[k] = o[k];
}

template <int NEQ>
struct LaneBdf {
  static constexpr int Q = kQmaxAbs;
  double p_gp, rtol, atol;
  int qmax, max_steps;
  TGR_GLOBAL double* jw;  // this thread's Jacobian: jw[(c * NEQ + r) * 64] = J(r, c)
  double zn[Q + 1][NEQ];
  double ewt[NEQ], acor[NEQ], y[NEQ], tempv[NEQ], ftemp[NEQ];
  // iteration matrix M = I - gamma J, after getrf its LU factors (unit lower
  // L), in this thread's column of an LDS block: element (r, c) at
  // Ml[(r * NEQ + c) * 64]
  TGR_LDS double* Ml;
  __device__ __forceinline__ TGR_LDS double& Mx(int r, int c) const { return Ml[(r * NEQ + c) * kWave]; }
  int piv[NEQ];
  double tn, h, hprime, hscale, eta, etamax, rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5;
  double tq[6], l[Q + 1], tau[Q + 2];
  int q, qprime, qwait, Lq, jcur, nst, nstlp, nstlj;
  int c_nfe, c_nfe_dq, c_nje, c_nsetups, c_nni, c_netf, c_ncfn, c_nlu;

  // the right-hand side through local call buffers (only they have their
  // address taken; the solver state stays in registers)
  __device__ __forceinline__ void f(const double (&yv)[NEQ], double (&out)[NEQ]) {
    double fin[NEQ], fout[NEQ];
#pragma unroll
    for (int k = 0; k < NEQ; ++k) fin[k] = yv[k];
    lane_rhs_call<NEQ>(fin, fout, p_gp);
#pragma unroll
    for (int k = 0; k < NEQ; ++k) out[k] = fout[k];
  }

  __device__ __forceinline__ double wrms(const double (&x)[NEQ]) const {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      const double v = x[i] * ewt[i];
      s += v * v;
    }
    return sqrt(s / NEQ);
  }

  __device__ __forceinline__ void set_ewt(const double (&yv)[NEQ]) {
#pragma unroll
    for (int i = 0; i < NEQ; ++i) ewt[i] = 1.0 / (rtol * fabs(yv[i]) + atol);
  }

  // LU with partial pivoting, rows exchanged in place; returns 0 or k + 1 at
  // a zero pivot in column k
  __device__ __forceinline__ int getrf() {
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      int p = k;
      double amax = fabs(Mx(k, k));
#pragma unroll
      for (int r = k + 1; r < NEQ; ++r) {
        const double a = fabs(Mx(r, k));
        if (a > amax) {
          amax = a;
          p = r;
        }
      }
      piv[k] = p;
      if (!(amax > 0.0)) return k + 1;
      if (p != k) {
#pragma unroll
        for (int r = k + 1; r < NEQ; ++r)
          if (r == p) {
#pragma unroll
            for (int c = 0; c < NEQ; ++c) {
              const double t = Mx(k, c);
              Mx(k, c) = Mx(r, c);
              Mx(r, c) = t;
            }
          }
      }
      const double rp = 1.0 / Mx(k, k);
      double rowk[NEQ];
#pragma unroll
      for (int c = k + 1; c < NEQ; ++c) rowk[c] = Mx(k, c);
#pragma unroll
      for (int r = k + 1; r < NEQ; ++r) {
        const double lf = Mx(r, k) * rp;
        Mx(r, k) = lf;
#pragma unroll
        for (int c = k + 1; c < NEQ; ++c) Mx(r, c) = fma(-lf, rowk[c], Mx(r, c));
      }
    }
    return 0;
  }

  // b <- M^-1 b with the factors and row exchanges of getrf
  __device__ __forceinline__ void getrs(double (&b)[NEQ]) const {
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      const int p = piv[k];
      if (p != k) {
#pragma unroll
        for (int r = k + 1; r < NEQ; ++r)
          if (r == p) {
            const double t = b[k];
            b[k] = b[r];
            b[r] = t;
          }
      }
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k)
#pragma unroll
      for (int r = k + 1; r < NEQ; ++r) b[r] = fma(-Mx(r, k), b[k], b[r]);
#pragma unroll
    for (int k = NEQ - 1; k >= 0; --k) {
      b[k] = b[k] / Mx(k, k);
#pragma unroll
      for (int r = 0; r < k; ++r) b[r] = fma(-Mx(r, k), b[k], b[r]);
    }
  }

  // difference-quotient Jacobian at yv (fy = f(yv)), column j perturbs y_j
  // (a rolled loop: one right-hand-side call site)
  __device__ __forceinline__ void dq_jac(const double (&yv)[NEQ], const double (&fy)[NEQ]) {
    const double srur = sqrt(UROUND);
    const double fnorm = wrms(fy);
    const double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(h) * UROUND * NEQ * fnorm) : 1.0;
    double fin[NEQ], fout[NEQ];
#pragma nounroll
    for (int j = 0; j < NEQ; ++j) {
      const double yj = pick(yv, j);
      const double inc = fmax(srur * fabs(yj), minInc / pick(ewt, j));
#pragma unroll
      for (int k = 0; k < NEQ; ++k) fin[k] = (k == j) ? yj + inc : yv[k];
      lane_rhs_call<NEQ>(fin, fout, p_gp);
      c_nfe_dq++;
      const double ii = 1.0 / inc;
#pragma unroll
      for (int r = 0; r < NEQ; ++r) jw[(j * NEQ + r) * kWave] = ii * fout[r] - ii * fy[r];
    }
  }

  __device__ __forceinline__ int lsetup(int convfail) {
    const double dgamma = fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst > nstlj + MSBJ) || ((convfail == CF_BAD_J) && (dgamma < DGMAX_J)) ||
                      (convfail == CF_OTHER);
    if (!jbad) {
      jcur = 0;
    } else {
      c_nje++;
      nstlj = nst;
      jcur = 1;
      dq_jac(zn[0], ftemp);
    }
#pragma unroll
    for (int c = 0; c < NEQ; ++c)
#pragma unroll
      for (int r = 0; r < NEQ; ++r) Mx(r, c) = (r == c ? 1.0 : 0.0) - gamma * jw[(c * NEQ + r) * kWave];
    c_nlu++;
    return getrf() ? 1 : 0;
  }

  __device__ __forceinline__ void set_bdf() {
    double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = h;
    l[0] = l[1] = 1.0;
#pragma unroll
    for (int i = 2; i <= Q; ++i) l[i] = 0.0;
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < Q; ++j) {
        if (j < q) {
          hsum += pick(tau, j - 1);
          xi_inv = h / hsum;
          alpha0 -= 1.0 / j;
#pragma unroll
          for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
        }
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += pick(tau, q - 1);
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = Q; i >= 1; --i)
        if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    const double lq = pick(l, q);
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / lq;
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += pick(tau, q);
      xi_inv = h / hsum;
      const double A5 = alpha0 - (1.0 / (q + 1));
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }

  __device__ __forceinline__ void set_coeffs() {
    set_bdf();
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  __device__ __forceinline__ void predict(double tstop) {
    tn += h;
    if ((tn - tstop) * h > 0.0) tn = tstop;
#pragma unroll
    for (int k = 1; k <= Q; ++k)
#pragma unroll
      for (int j = Q; j >= k; --j)
        if (k <= q && j <= q)
#pragma unroll
          for (int i = 0; i < NEQ; ++i) zn[j - 1][i] += zn[j][i];
  }

  __device__ __forceinline__ void restore(double saved_t) {
    tn = saved_t;
#pragma unroll
    for (int k = 1; k <= Q; ++k)
#pragma unroll
      for (int j = Q; j >= k; --j)
        if (k <= q && j <= q)
#pragma unroll
          for (int i = 0; i < NEQ; ++i) zn[j - 1][i] -= zn[j][i];
  }

  __device__ __forceinline__ void rescale() {
    double factor = eta;
#pragma unroll
    for (int j = 1; j <= Q; ++j)
      if (j <= q) {
#pragma unroll
        for (int i = 0; i < NEQ; ++i) zn[j][i] *= factor;
        factor *= eta;
      }
    h = hscale * eta;
    hscale = h;
  }

  __device__ __forceinline__ void increase_order() {
    double lo[Q + 2];
#pragma unroll
    for (int i = 0; i < Q + 2; ++i) lo[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
    lo[2] = 1.0;
#pragma unroll
    for (int j = 1; j < Q; ++j) {
      if (j < q) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) lo[i] = lo[i] * xiold + lo[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      const double zl = A1 * zn[Q][i];  // zn[qmax] holds the saved acor (qmax = Q)
#pragma unroll
      for (int j = 0; j <= Q; ++j)
        if (j == Lq) zn[j][i] = zl;
#pragma unroll
      for (int j = 2; j <= Q; ++j)
        if (j <= q) zn[j][i] += lo[j] * zl;
    }
  }

  __device__ __forceinline__ void decrease_order() {
    double lo[Q + 2];
#pragma unroll
    for (int i = 0; i < Q + 2; ++i) lo[i] = 0.0;
    lo[2] = 1.0;
    double hsum = 0.0;
#pragma unroll
    for (int j = 1; j <= Q - 2; ++j) {
      if (j <= q - 2) {
        hsum += tau[j];
        const double xi = hsum / hscale;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) lo[i] = lo[i] * xi + lo[i - 1];
      }
    }
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      const double zq = pick_row(q, i);
#pragma unroll
      for (int j = 2; j < Q; ++j)
        if (j < q) zn[j][i] -= lo[j] * zq;
    }
  }

  // zn[j][i] for a run-time j
  __device__ __forceinline__ double pick_row(int j, int i) const {
    double r = zn[0][i];
#pragma unroll
    for (int k = 1; k <= Q; ++k) r = (j == k) ? zn[k][i] : r;
    return r;
  }

  __device__ __forceinline__ void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) increase_order();
    else decrease_order();
  }

  __device__ __forceinline__ void adjust_params() {
    if (qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      Lq = q + 1;
      qwait = Lq;
    }
    rescale();
  }

  __device__ __forceinline__ int newton_iter() {
    int mm = 0;
    double delp = 0.0;
    for (;;) {
      double b[NEQ];
#pragma unroll
      for (int i = 0; i < NEQ; ++i) b[i] = gamma * ftemp[i] - (rl1 * zn[1][i] + acor[i]);
      getrs(b);
      if (gamrat != 1.0) {
        const double s = 2.0 / (1.0 + gamrat);
#pragma unroll
        for (int i = 0; i < NEQ; ++i) b[i] *= s;
      }
      c_nni++;
      const double del = wrms(b);
#pragma unroll
      for (int i = 0; i < NEQ; ++i) {
        acor[i] += b[i];
        y[i] = zn[0][i] + acor[i];
      }
      if (mm > 0) crate = fmax(CRDOWN * crate, del / delp);
      const double dcon = del * fmin(1.0, crate) / tq[4];
      if (dcon <= 1.0) {
        acnrm = (mm == 0) ? del : wrms(acor);
        jcur = 0;
        return NF_OK;
      }
      ++mm;
      if (mm == NLS_MAXCOR || (mm >= 2 && del > RDIV * delp)) return jcur ? NF_CONV_FAIL : NF_TRY_AGAIN;
      delp = del;
      f(y, ftemp);
      c_nfe++;
    }
  }

  __device__ __forceinline__ int nls(int nflag) {
    int convfail = (nflag == NF_FIRST || nflag == NF_PREV_ERR) ? CF_NONE : CF_OTHER;
    bool callSetup = (nflag == NF_PREV_CONV) || (nflag == NF_PREV_ERR) || (nst == 0) || (nst >= nstlp + MSBP) ||
                     (fabs(gamrat - 1.0) > DGMAX);
    for (;;) {
      f(zn[0], ftemp);
      c_nfe++;
      if (callSetup) {
        const int ier = lsetup(convfail);
        c_nsetups++;
        callSetup = false;
        gamrat = crate = 1.0;
        gammap = gamma;
        nstlp = nst;
        if (ier) return NF_CONV_FAIL;
      }
#pragma unroll
      for (int i = 0; i < NEQ; ++i) {
        acor[i] = 0.0;
        y[i] = zn[0][i];
      }
      const int ier = newton_iter();
      if (ier != NF_TRY_AGAIN) return ier;
      callSetup = true;
      convfail = CF_BAD_J;
    }
  }

  __device__ __forceinline__ int handle_nflag(int& nflag, double saved_t, int& ncf) {
    if (nflag == NF_OK) return 0;
    c_ncfn++;
    restore(saved_t);
    ncf++;
    etamax = 1.0;
    if (ncf == MXNCF) return TGR_ST_CONV_FAIL;
    eta = ETACF;
    nflag = NF_PREV_CONV;
    rescale();
    return 1;
  }

  __device__ __forceinline__ int error_test(int& nflag, double saved_t, int& nef, double& dsm) {
    dsm = acnrm * tq[2];
    if (dsm <= 1.0) return 0;
    nef++;
    c_netf++;
    nflag = NF_PREV_ERR;
    restore(saved_t);
    if (nef == MXNEF) return TGR_ST_ERR_FAIL;
    etamax = 1.0;
    if (nef <= MXNEF1) {
      eta = 1.0 / (rpow(BIAS2 * dsm, 1.0 / Lq) + ADDON);
      eta = fmax(ETAMIN, eta);
      if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
      rescale();
      return 1;
    }
    if (q > 1) {
      eta = ETAMIN;
      adjust_order(-1);
      Lq = q;
      q--;
      qwait = Lq;
      rescale();
      return 1;
    }
    eta = ETAMIN;
    h *= eta;
    hscale = h;
    qwait = LONG_WAIT;
    f(zn[0], tempv);
    c_nfe++;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) zn[1][i] = h * tempv[i];
    return 1;
  }

  __device__ __forceinline__ void complete_step() {
    nst++;
#pragma unroll
    for (int i = Q; i >= 2; --i)
      if (i <= q) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
#pragma unroll
    for (int j = 0; j <= Q; ++j)
      if (j <= q)
#pragma unroll
        for (int i = 0; i < NEQ; ++i) zn[j][i] += l[j] * acor[i];
    qwait--;
    if (qwait == 1 && q != qmax) {
#pragma unroll
      for (int i = 0; i < NEQ; ++i) zn[Q][i] = acor[i];
      saved_tq5 = tq[5];
    }
  }

  __device__ __forceinline__ void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      hprime = h * eta;
    }
  }

  __device__ __forceinline__ void prepare_next(double dsm) {
    if (etamax == 1.0) {
      if (qwait < 2) qwait = 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    const double etaq = 1.0 / (rpow(BIAS2 * dsm, 1.0 / Lq) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    double etaqm1 = 0.0;
    if (q > 1) {
      double zq[NEQ];
#pragma unroll
      for (int i = 0; i < NEQ; ++i) zq[i] = pick_row(q, i);
      const double ddn = wrms(zq) * tq[1];
      etaqm1 = 1.0 / (rpow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    double etaqp1 = 0.0;
    if (q != qmax && saved_tq5 != 0.0) {
      const double cquot = (tq[5] / saved_tq5) * rpow(h / tau[2], (double)Lq);
#pragma unroll
      for (int i = 0; i < NEQ; ++i) tempv[i] = acor[i] - cquot * zn[Q][i];
      const double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (rpow(BIAS3 * dup, 1.0 / (Lq + 1)) + ADDON);
    }
    const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < THRESH) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
#pragma unroll
      for (int i = 0; i < NEQ; ++i) zn[Q][i] = acor[i];
    }
    set_eta();
  }

  __device__ __forceinline__ int step(double tstop) {
    const double saved_t = tn;
    double dsm = 0.0;
    int ncf = 0, nef = 0, nflag = NF_FIRST;
    if (nst > 0 && hprime != h) adjust_params();
    for (;;) {
      predict(tstop);
      set_coeffs();
      nflag = nls(nflag);
      const int k = handle_nflag(nflag, saved_t, ncf);
      if (k == 1) continue;
      if (k < 0) return k;
      const int e = error_test(nflag, saved_t, nef, dsm);
      if (e == 1) continue;
      if (e < 0) return e;
      break;
    }
    complete_step();
    prepare_next(dsm);
    etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    const double s = 1.0 / tq[2];
#pragma unroll
    for (int i = 0; i < NEQ; ++i) acor[i] *= s;
    return 0;
  }

  __device__ __forceinline__ double initial_h(double tout) {
    const double tdist = fabs(tout - tn);
    const double tround = UROUND * fmax(fabs(tn), fabs(tout));
    const double hlb = HLB_FACTOR * tround;
    double hub_inv = 0.0;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      const double den = HUB_FACTOR * fabs(zn[0][i]) + 1.0 / ewt[i];
      hub_inv = fmax(hub_inv, fabs(zn[1][i]) / den);
    }
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = sqrt(hlb * hub);
    if (hub < hlb) return hg;
    double hnew = hg;
    bool hnewOK = false;
    for (int it = 1; it <= HIN_ITERS; ++it) {
#pragma unroll
      for (int i = 0; i < NEQ; ++i) y[i] = hg * zn[1][i] + zn[0][i];
      f(y, tempv);
      c_nfe++;
      const double ih = 1.0 / hg;
#pragma unroll
      for (int i = 0; i < NEQ; ++i) tempv[i] = (tempv[i] - zn[1][i]) * ih;
      const double yddnrm = wrms(tempv);
      if (hnewOK || it == HIN_ITERS) {
        hnew = hg;
        break;
      }
      hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) hnewOK = true;
      if (it > 1 && hrat > 2.0) {
        hnew = hg;
        hnewOK = true;
      }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    return h0;
  }

  // integrate 0 -> tout; sol updated in place
  __device__ __forceinline__ int integrate(double tout, double (&sol)[NEQ]) {
    c_nfe = c_nfe_dq = c_nje = c_nsetups = c_nni = c_netf = c_ncfn = c_nlu = 0;
    nst = 0;
    bool finite = true;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) finite = finite && isfinite(sol[i]);
    if (!finite) return TGR_ST_BAD_INPUT;
    tn = 0.0;
    q = 1;
    Lq = 2;
    qwait = Lq;
    qprime = 1;
    etamax = ETAMX1;
    nstlp = nstlj = 0;
    saved_tq5 = 0.0;
    crate = 1.0;
    gammap = gamrat = 1.0;
    jcur = 0;
    eta = 1.0;
#pragma unroll
    for (int j = 0; j < Q + 2; ++j) tau[j] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) tq[j] = 0.0;
#pragma unroll
    for (int j = 0; j <= Q; ++j) l[j] = 0.0;
#pragma unroll
    for (int j = 0; j <= Q; ++j)
#pragma unroll
      for (int i = 0; i < NEQ; ++i) zn[j][i] = 0.0;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      acor[i] = y[i] = tempv[i] = 0.0;
      zn[0][i] = sol[i];
    }
    set_ewt(zn[0]);
    f(zn[0], zn[1]);
    c_nfe++;
    const double tstop = tout;
    if (!(fabs(tout - tn) >= 2.0 * UROUND * fabs(tout)) || tout <= 0.0) return TGR_ST_BAD_INPUT;
    h = initial_h(tout);
    if ((tn + h - tstop) * h > 0.0) h = (tstop - tn) * (1.0 - 4.0 * UROUND);
    hscale = h;
    hprime = h;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) zn[1][i] *= h;
    int nstloc = 0;
    for (;;) {
      if (nst > 0) set_ewt(zn[0]);
      if (nstloc >= max_steps) {
#pragma unroll
        for (int i = 0; i < NEQ; ++i) sol[i] = zn[0][i];
        return TGR_ST_TOO_MUCH_WORK;
      }
      if (UROUND * wrms(zn[0]) > 1.0) {
#pragma unroll
        for (int i = 0; i < NEQ; ++i) sol[i] = zn[0][i];
        return TGR_ST_TOO_MUCH_ACC;
      }
      const int k = step(tstop);
      if (k < 0) {
#pragma unroll
        for (int i = 0; i < NEQ; ++i) sol[i] = zn[0][i];
        return k;
      }
      nstloc++;
      double tret = -1.0;
      if ((tn - tout) * h >= 0.0) {
        tret = tout;
      } else {
        const double troundoff = FUZZ_FACTOR * UROUND * (fabs(tn) + fabs(h));
        if (fabs(tn - tstop) <= troundoff) tret = tstop;
      }
      if (tret >= 0.0) {
        const double s = (tret - tn) / h;
#pragma unroll
        for (int i = 0; i < NEQ; ++i) {
          double v = pick_row(q, i);
#pragma unroll
          for (int j = Q - 1; j >= 0; --j)
            if (j <= q - 1) v = zn[j][i] + s * v;
          sol[i] = v;
        }
        return TGR_ST_OK;
      }
      if ((tn + hprime - tstop) * h > 0.0) {
        hprime = (tstop - tn) * (1.0 - 4.0 * UROUND);
        eta = hprime / h;
      }
    }
  }
};

}  // namespace tgr

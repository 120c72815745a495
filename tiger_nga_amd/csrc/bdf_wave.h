// Wave-per-cell CVODE-semantics BDF integrator for CDNA4 (gfx950, wave64).
//
// Replaces the SUNDIALS CVODE 2.5.0 solve that finitechem_source_chemistry
// drives (reference combustion/finitechem.f90:1014-1034: BDF, Newton, dense
// difference-quotient Jacobian, scalar tolerances, tstop = tout = dt,
// MAX_NSTEPS).  The integrator follows the published CVODE algorithm step for
// step (the same restatement the CPU oracle holds), laid out for a 64-wide
// wavefront:
//   * vectors of length neq = nsp+1 live one component per lane; the
//     Nordsieck history lives in LDS (lane-private columns);
//   * the iteration matrix M = I - gamma*J lives one ROW per lane in
//     registers (NEQ doubles); the saved Jacobian J (reused for up to 50
//     steps) lives in a per-wave HBM workspace, column-major with the rows in
//     lanes, so each column store / row load is one coalesced access;
//   * LU with partial pivoting keeps rows in their lanes (implicit row
//     exchange): the pivot is a wave arg-max, the pivot row is broadcast with
//     v_readlane, and each lane records the elimination step of its row;
//   * all step-size / order / Newton control is wave-uniform scalar code.
#pragma once
#include <stdint.h>

#include "chem_device.h"
#include "chem_params.h"

namespace tgr {

// CVODE defaults and algorithm constants (same values as the CPU oracle)
constexpr int kQmaxAbs = 5;
constexpr int MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, NLS_MAXCOR = 3;
constexpr int MSBP = 20, MSBJ = 50, SMALL_NST = 10;
constexpr double CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0, DGMAX_J = 0.2;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1, ETACF = 0.25;
constexpr double ADDON = 1.0e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5, NLSCOEF = 0.1;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5, FUZZ_FACTOR = 100.0, MIN_INC_MULT = 1000.0;
constexpr int HIN_ITERS = 4;
constexpr double UROUND = 2.2204460492503131e-16;

enum { NF_FIRST = 0, NF_PREV_CONV = 1, NF_PREV_ERR = 2, NF_OK = 3, NF_CONV_FAIL = 4, NF_TRY_AGAIN = 5 };
enum { CF_NONE = 0, CF_BAD_J = 1, CF_OTHER = 2 };

// x^y for x >= 0 through the project's exp/log (step-size ratios only)
__device__ __forceinline__ double rpow(double x, double y) { return fexp(y * flog(x)); }

template <class T, int N>
__device__ __forceinline__ T pick(const T (&a)[N], int i) {
  T r = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (i == k) ? a[k] : r;
  return r;
}

template <class T, int N>
__device__ __forceinline__ void put(T (&a)[N], int i, T v) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = (i == k) ? v : a[k];
}

// Analytic Jacobian of finitechem_compute_rhs at y0 (this lane's component)
// into jw (jw[c * 64 + lane] = J(lane, c)); returns f(y0).
// With C_k = y_k / W_k, T = y_T / rho, rho = sum_k y_k (species) and
// f_k = W_k CDOT_k(C, T), f_T = -g / cp_mix, g = sum_k h_k f_k:
//   J(k, j)  = W_k dCDOT_k/dC_j / W_j + fT_k dT/dy_j,  fT_k = W_k dCDOT_k/dT
//   J(T, j)  = -(dT/dy_j sum_k cp_k f_k + sum_k h_k J(k, j)) / cp_mix
//              + g / cp_mix^2 dcp_mix/dy_j
//   dT/dy_j  = -T / rho (species), 1 / rho (y_T);
//   dcp_mix/dy_j = (cp_j - cp_mix) / rho (species) + dT/dy_j sum_k Y_k dcp_k/dT.
template <int NEQ>
__device__ __forceinline__ double jac_wave(tchar* tab, TGR_LDS RhsLds<NEQ>* S, double* jw, double tmid, double y0,
                                           double p_gp) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const bool sp = lane < nsp;
  const double rho = wave_sum(sp ? y0 : 0.0);
  const double T = bcast(y0, nsp) / rho;
  double hk, cpk, dcpk;
  thermo_lane_d<NEQ>(tab, tmid, T, lane, hk, cpk, dcpk);
  const double Wk = sp ? ld<double>(tab, L.w + lane * 8) : 0.0;
  const double Ck = sp ? y0 * ld<double>(tab, L.iw + lane * 8) : 0.0;
  const double cpm = wave_sum(sp ? (y0 / rho) * cpk : 0.0);
  const double cpmT = wave_sum(sp ? (y0 / rho) * dcpk : 0.0);
  const double ctot = wave_sum(Ck);
  JacLane jl[L.rounds];
  double dcdT;
  const double cdot = prodrates_jac_base<NEQ>(tab, *S, Ck, ctot, T, p_gp, jl, dcdT);
  const double fk = sp ? cdot * Wk : 0.0;
  const double fTk = sp ? dcdT * Wk : 0.0;
  const double tmp = wave_sum(-hk * fk);
  const double A = wave_sum(cpk * fk);
  const double fy = sp ? fk : (lane == nsp ? tmp / cpm : 0.0);
  const double g = -tmp;
  const double irho = 1.0 / rho;
  const double dTs = -T * irho;
  const double icpm = 1.0 / cpm;
  const double gc2 = g * icpm * icpm;
  TGR_GLOBAL double* jg = (TGR_GLOBAL double*)jw;  // stores only (fresh() needs a scalar pointer)
  for (int j = 0; j < nsp; ++j) {
    const double dc = jac_column<NEQ>(tab, *S, jl, j);
    const double iWj = ld<double>(tab, L.iw + j * 8);
    const double Jk = sp ? fma(Wk * dc, iWj, fTk * dTs) : 0.0;
    const double hJ = wave_sum(hk * Jk);
    const double cpj = bcast(cpk, j);
    const double dcpm = fma(cpmT, dTs, (cpj - cpm) * irho);
    const double JT = fma(gc2, dcpm, -fma(dTs, A, hJ) * icpm);
    jg[j * kWave + lane] = sp ? Jk : (lane == nsp ? JT : 0.0);
  }
  {
    const double Jk = sp ? fTk * irho : 0.0;
    const double hJ = wave_sum(hk * Jk);
    const double JT = fma(gc2, cpmT * irho, -fma(irho, A, hJ) * icpm);
    jg[nsp * kWave + lane] = sp ? Jk : (lane == nsp ? JT : 0.0);
  }
  return fy;
}

// Out of line, like the right-hand side: its per-reaction registers do not
// join the integrator's register allocation.
template <int NEQ>
__device__ __attribute__((noinline)) double jac_call(tchar* tab, TGR_LDS RhsLds<NEQ>* S, double* jw, double tmid,
                                                     double y0, double p_gp) {
  return jac_wave<NEQ>(tab, S, jw, tmid, y0, p_gp);
}

// AJ: the analytic Jacobian (jac_analytic) instead of difference quotients
// (a compile-time choice: the difference-quotient instance carries no call to
// it, and its register allocation is that of the plain solver)
template <int NEQ, bool AJ = false>
struct Bdf {
  const DevMech m;  // by value: fields stay in scalar registers
  tchar* tab;        // mechanism tables in LDS
  TGR_LDS RhsLds<NEQ>* S;  // this wave's right-hand-side scratch in LDS
  double p_gp, rtol, atol;
  int lane, n, qmax, max_steps;
  bool act;
  // per-lane vectors; the Nordsieck history lives in LDS (lane-private
  // columns), so that zn[q] with a run-time order q is a plain address
  TGR_LDS double (*zn)[kWave];  // [kQmaxAbs + 2] rows
  double ewt, acor, y, tempv, ftemp;
  double* jw;  // saved Jacobian: jw[c * 64 + lane] = J(lane, c)
  double M[NEQ];
  int pivl;
  double rdk;
  // wave-uniform control state
  double tn, h, hprime, hscale, eta, etamax, rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5;
  // small wave-uniform arrays indexed by the run-time order (LDS, every lane
  // writes the same value)
  TGR_LDS double* tq;
  TGR_LDS double* l;
  TGR_LDS double* tau;
  int q, qprime, qwait, Lq, jcur, nst, nstlp, nstlj;
  int c_nfe, c_nfe_dq, c_nje, c_nsetups, c_nni, c_netf, c_ncfn, c_nlu;
  // one-step SDIRK solves: rate cache for the split Jacobian (nullptr: the
  // plain difference-quotient Jacobian)
  TGR_LDS RateCache<NEQ>* rc = nullptr;

  __device__ Bdf(const DevMech& mech, tchar* tabl, TGR_LDS RhsLds<NEQ>* rl, TGR_LDS double (&znl)[kQmaxAbs + 2][kWave],
                 double* ws, double pgp, double rt, double at, int qm, int ms)
      : m(mech), tab(tabl), S(rl), p_gp(pgp), rtol(rt), atol(at), max_steps(ms), zn(znl), jw(ws) {
    tq = &znl[kQmaxAbs + 1][0];
    l = tq + 8;
    tau = tq + 16;
    lane = lane_id();
    n = NEQ;
    act = lane < NEQ;
    qmax = (qm >= 1 && qm <= kQmaxAbs) ? qm : kQmaxAbs;
  }

  // without the Nordsieck storage: for the one-step SDIRK solves only (the
  // BDF members that use zn, tq, l or tau must not be called)
  __device__ Bdf(const DevMech& mech, tchar* tabl, TGR_LDS RhsLds<NEQ>* rl, double* ws, double pgp, double rt, double at)
      : m(mech), tab(tabl), S(rl), p_gp(pgp), rtol(rt), atol(at), max_steps(0), zn(nullptr), jw(ws) {
    tq = l = tau = nullptr;
    lane = lane_id();
    n = NEQ;
    act = lane < NEQ;
    qmax = kQmaxAbs;
    pivl = lane;
  }

  __device__ __forceinline__ double f(double yv) { return rhs_call<NEQ>(tab, S, m.tmid, act ? yv : 0.0, p_gp); }

  __device__ __forceinline__ double wrms(double x) {
    double v = act ? x * ewt : 0.0;
    return sqrt((wave_sum(v * v)) / n);
  }

  __device__ __forceinline__ void set_ewt(double yv) { ewt = act ? 1.0 / (rtol * fabs(yv) + atol) : 0.0; }

  // ---- dense linear algebra, one matrix row per lane --------------------
  // getrf factors M with partial pivoting (rows stay in their lanes during
  // the elimination; the pivot row is broadcast through LDS, the right-hand-
  // side scratch being idle), then moves row p_k of the factors to lane k,
  // so that the triangular solves run in pivot order with compile-time lane
  // masks.  rdk: lane k holds 1 / pivot of step k; pivl: lane k holds p_k.
  __device__ __forceinline__ int getrf() {
    TGR_LDS double* prow = S->q;
    int pstep = act ? NEQ : -1;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      const bool cand = (pstep == NEQ);
      const double a = cand ? fabs(M[k]) : -1.0;
      const double amax = (wave_max(a));
      if (!(amax > 0.0)) return k + 1;
      const uint64_t bal = __ballot(cand && a == amax);
      const int p = __builtin_amdgcn_readfirstlane(bal ? (int)__builtin_ctzll(bal) : 0);
      if (lane == p) {
        pstep = k;
#pragma unroll
        for (int c = k; c < NEQ; ++c) prow[c] = M[c];
      }
      wave_sync();
      const double rp = 1.0 / prow[k];
      if (lane == k) {
        pivl = p;
        rdk = rp;
      }
      const bool elim = cand && lane != p;
      // multiplier, 0 on rows that are not eliminated (then the update
      // below leaves them unchanged exactly)
      const double lf = elim ? M[k] * rp : 0.0;
      if (elim) M[k] = lf;
#pragma unroll
      for (int c = k + 1; c < NEQ; ++c) M[c] = fma(-prow[c], lf, M[c]);
      wave_sync();
    }
#pragma unroll
    for (int c = 0; c < NEQ; ++c) M[c] = __shfl(M[c], pivl);
    return 0;
  }

  // solve M x = b with the factors from getrf; b, x: one component per lane
  // (b in equation order, x in unknown order)
  __device__ __forceinline__ double getrs(double b) {
    b = __shfl(b, pivl);  // pivot order
#pragma unroll
    for (int k = 0; k < NEQ - 1; ++k) {
      const double yk = bcast(b, k);
      if (lane > k) b = fma(-M[k], yk, b);  // lanes past NEQ: never read
    }
#pragma unroll
    for (int k = NEQ - 1; k >= 0; --k) {
      const double xk = bcast(b, k) * bcast(rdk, k);
      if (lane < k) b = fma(-M[k], xk, b);
    }
    // lane k's b is final from step k on (later steps update lanes below
    // it): its x_k is the same product the step broadcast
    return act ? b * rdk : 0.0;
  }

  // Gauss-Jordan inverse of M in place, partial pivoting, for the one-step
  // solves (about twenty solves per factorisation there: each one is then a
  // matrix-vector product of independent broadcasts instead of two dependent
  // triangular sweeps).  Rows stay in their lanes.  Step k exchanges unknown
  // k with equation p_k, the not yet pivoted row of largest |a_ik|: with
  // r = 1 / a_pk, row p_k becomes r in column k and -r a_pj elsewhere, every
  // other row i becomes a_ik r in column k and a_ij - a_ik r a_pj elsewhere
  // (the pivot row through LDS).  The columns rotate by one per step (the
  // updated column k+1 lands in M[0], column k's new values in M[NEQ-1]), so
  // that the pivot column is always M[0] and the steps run as a loop with
  // constant register indices; after the NEQ steps every column is back in
  // place.  Row p_k then expresses x_k with the coefficient of b_{p_j} in
  // column j; lane k receives row p_k, and pivl (lane k) holds p_k.  Returns
  // 0, or k + 1 when column k has no nonzero candidate.
  __device__ __forceinline__ int getri() {
    TGR_LDS double* prow = S->q;
    bool cand = act;
#pragma unroll 1
    for (int k = 0; k < NEQ; ++k) {
      const double a = cand ? fabs(M[0]) : -1.0;
      const double amax = (wave_max(a));
      if (!(amax > 0.0)) return k + 1;
      const uint64_t bal = __ballot(cand && a == amax);
      const int p = __builtin_amdgcn_readfirstlane(bal ? (int)__builtin_ctzll(bal) : 0);
      const bool isp = lane == p;
      if (isp) {
        cand = false;
#pragma unroll
        for (int c = 0; c < NEQ; ++c) prow[c] = M[c];
      }
      wave_sync();
      const double r = 1.0 / prow[0];
      if (lane == k) pivl = p;
      // pivot row: 0 * a_pc - r a_pc; other rows: a_ic - (a_ik r) a_pc (the
      // products by 0 and 1 are exact)
      const double fk = (isp ? 1.0 : M[0]) * r;
      const double g = isp ? 0.0 : 1.0;
#pragma unroll
      for (int c = 0; c + 1 < NEQ; ++c) M[c] = fma(-fk, prow[c + 1], g * M[c + 1]);
      M[NEQ - 1] = fk;
      wave_sync();
    }
#pragma unroll
    for (int c = 0; c < NEQ; ++c) M[c] = __shfl(M[c], pivl);
    return 0;
  }

  // x = A^-1 b with the inverse from getri (b in equation order, x in
  // unknown order; 0 on the lanes past NEQ): b, in pivot order, goes through
  // the idle LDS scratch and comes back to every lane as broadcast reads (LDS
  // instructions instead of a pair of v_readlane per component); two partial
  // sums
  __device__ __forceinline__ double getri_solve(double b) {
    TGR_LDS double* v = S->q;
    v[lane] = __shfl(b, pivl);  // lane j: b_{p_j}
    wave_sync();
    double x0 = 0.0, x1 = 0.0;
#pragma unroll
    for (int j = 0; j + 1 < NEQ; j += 2) {
      x0 = fma(M[j], v[j], x0);
      x1 = fma(M[j + 1], v[j + 1], x1);
    }
    if constexpr (NEQ % 2) x0 = fma(M[NEQ - 1], v[NEQ - 1], x0);
    wave_sync();
    return act ? x0 + x1 : 0.0;
  }

  // difference-quotient Jacobian, column j perturbs component j
  __device__ __forceinline__ void dq_jac(double yv, double fy) {
    const double srur = sqrt(UROUND);
    const double fnorm = wrms(fy);
    const double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(h) * UROUND * n * fnorm) : 1.0;
    for (int j = 0; j < NEQ; ++j) {
      const double yj = bcast(yv, j);
      const double ewj = bcast(ewt, j);
      const double inc = fmax(srur * fabs(yj), minInc / ewj);
      const double yp = (lane == j) ? yj + inc : yv;
      const double fp = f(yp);
      c_nfe_dq++;
      const double ii = 1.0 / inc;
      fresh(jw)[j * kWave + lane] = ii * fp - ii * fy;
    }
  }

  // The Jacobian with the temperature-dependent factors evaluated once: with
  // T = y_T / rho, J(:, j) = df/dy_j|_T + df/dT|_y dT/dy_j, dT/dy_j = -T / rho
  // for a species and 1 / rho for y_T.  The species columns are difference
  // quotients at fixed T (increments as in dq_jac, from the current h) that
  // reuse the base evaluation's rate constants and thermo (rc); df/dT is one
  // difference quotient in T.  jac_split_base evaluates f(yv) and fills the
  // rate cache; jac_split_cols (same yv, fy its result) stores J in jw.
  __device__ __forceinline__ double jac_split_base(double yv) {
    constexpr int nsp = NEQ - 1;
    const double y0 = act ? yv : 0.0;
    const double rho = wave_sum(lane < nsp ? y0 : 0.0);
    const double T = bcast(y0, nsp) / rho;
    const double fy = rhs_call_T<NEQ, 1>(tab, S, rc, m.tmid, y0, p_gp, T);
    c_nfe++;
    return fy;
  }

  __device__ __forceinline__ void jac_split_cols(double yv, double fy) {
    constexpr int nsp = NEQ - 1;
    const double y0 = act ? yv : 0.0;
    const double rho = wave_sum(lane < nsp ? y0 : 0.0);
    const double T = bcast(y0, nsp) / rho;
    const double srur = sqrt(UROUND);
    const double fnorm = wrms(fy);
    const double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(h) * UROUND * n * fnorm) : 1.0;
    for (int j = 0; j < nsp; ++j) {
      const double yj = bcast(y0, j);
      const double ewj = bcast(ewt, j);
      const double inc = fmax(srur * fabs(yj), minInc / ewj);
      const double yp = (lane == j) ? yj + inc : y0;
      const double fp = rhs_call_T<NEQ, 2>(tab, S, rc, m.tmid, yp, p_gp, T);
      c_nfe_dq++;
      const double ii = 1.0 / inc;
      fresh(jw)[j * kWave + lane] = ii * fp - ii * fy;
    }
    const double dT = srur * T;
    const double fT = rhs_call_T<NEQ, 0>(tab, S, rc, m.tmid, y0, p_gp, T + dT);
    c_nfe_dq++;
    const double dfdT = act ? (fT - fy) / dT : 0.0;
    const double sy = -T / rho;
    for (int j = 0; j < nsp; ++j) fresh(jw)[j * kWave + lane] += sy * dfdT;
    fresh(jw)[nsp * kWave + lane] = dfdT / rho;
  }

  __device__ __forceinline__ double jac_split(double yv) {
    const double fy = jac_split_base(yv);
    jac_split_cols(yv, fy);
    return fy;
  }

  // Analytic Jacobian at yv into jw (jac_wave below); returns f(yv).
  __device__ __forceinline__ double jac_analytic(double yv) {
    return jac_call<NEQ>(tab, S, jw, m.tmid, act ? yv : 0.0, p_gp);
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
      if constexpr (AJ) jac_analytic(zn[0][lane]);
      else dq_jac(zn[0][lane], ftemp);
    }
    const TGR_GLOBAL double* jr = fresh(jw) + lane;
#pragma unroll
    for (int c = 0; c < NEQ; ++c) {
      double v = -gamma * jr[c * kWave];
      if (c == lane) v += 1.0;
      M[c] = v;
    }
    c_nlu++;
    return getrf() ? 1 : 0;
  }

  // ---- BDF bookkeeping ----------------------------------------------------
  __device__ __forceinline__ void set_bdf() {
    double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = h;
    l[0] = l[1] = 1.0;
#pragma unroll
    for (int i = 2; i <= kQmaxAbs; ++i) l[i] = 0.0;
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < kQmaxAbs; ++j) {
        if (j < q) {
          hsum += tau[j - 1];
          xi_inv = h / hsum;
          alpha0 -= 1.0 / j;
#pragma unroll
          for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
        }
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = kQmaxAbs; i >= 1; --i)
        if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    const double lq = l[q];
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
      hsum += tau[q];
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
    for (int k = 1; k <= kQmaxAbs; ++k)
#pragma unroll
      for (int j = kQmaxAbs; j >= k; --j)
        if (k <= q && j <= q) zn[j - 1][lane] += zn[j][lane];
  }

  __device__ __forceinline__ void restore(double saved_t) {
    tn = saved_t;
#pragma unroll
    for (int k = 1; k <= kQmaxAbs; ++k)
#pragma unroll
      for (int j = kQmaxAbs; j >= k; --j)
        if (k <= q && j <= q) zn[j - 1][lane] -= zn[j][lane];
  }

  __device__ __forceinline__ void rescale() {
    double factor = eta;
#pragma unroll
    for (int j = 1; j <= kQmaxAbs; ++j)
      if (j <= q) {
        zn[j][lane] *= factor;
        factor *= eta;
      }
    h = hscale * eta;
    hscale = h;
  }

  __device__ __forceinline__ void increase_order() {
    double lo[kQmaxAbs + 2];
#pragma unroll
    for (int i = 0; i < kQmaxAbs + 2; ++i) lo[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
    lo[2] = 1.0;
#pragma unroll
    for (int j = 1; j < kQmaxAbs; ++j) {
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
    const double zl = A1 * zn[qmax][lane];
    zn[Lq][lane] = zl;
#pragma unroll
    for (int j = 2; j <= kQmaxAbs; ++j)
      if (j <= q) zn[j][lane] += lo[j] * zl;
  }

  __device__ __forceinline__ void decrease_order() {
    double lo[kQmaxAbs + 2];
#pragma unroll
    for (int i = 0; i < kQmaxAbs + 2; ++i) lo[i] = 0.0;
    lo[2] = 1.0;
    double hsum = 0.0;
#pragma unroll
    for (int j = 1; j <= kQmaxAbs - 2; ++j) {
      if (j <= q - 2) {
        hsum += tau[j];
        const double xi = hsum / hscale;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) lo[i] = lo[i] * xi + lo[i - 1];
      }
    }
    const double zq = zn[q][lane];
#pragma unroll
    for (int j = 2; j < kQmaxAbs; ++j)
      if (j < q) zn[j][lane] -= lo[j] * zq;
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
      const double r = rl1 * zn[1][lane] + acor;
      double b = act ? gamma * ftemp - r : 0.0;
      b = getrs(b);
      if (gamrat != 1.0) b *= 2.0 / (1.0 + gamrat);
      c_nni++;
      const double del = wrms(b);
      acor += b;
      y = zn[0][lane] + acor;
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
      ftemp = f(y);
      c_nfe++;
    }
  }

  __device__ __forceinline__ int nls(int nflag) {
    int convfail = (nflag == NF_FIRST || nflag == NF_PREV_ERR) ? CF_NONE : CF_OTHER;
    bool callSetup = (nflag == NF_PREV_CONV) || (nflag == NF_PREV_ERR) || (nst == 0) || (nst >= nstlp + MSBP) ||
                     (fabs(gamrat - 1.0) > DGMAX);
    for (;;) {
      ftemp = f(zn[0][lane]);
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
      acor = 0.0;
      y = zn[0][lane];
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
    tempv = f(zn[0][lane]);
    c_nfe++;
    zn[1][lane] = act ? h * tempv : 0.0;
    return 1;
  }

  __device__ __forceinline__ void complete_step() {
    nst++;
#pragma unroll
    for (int i = kQmaxAbs; i >= 2; --i)
      if (i <= q) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
#pragma unroll
    for (int j = 0; j <= kQmaxAbs; ++j)
      if (j <= q) zn[j][lane] += l[j] * acor;
    qwait--;
    if (qwait == 1 && q != qmax) {
      zn[qmax][lane] = acor;
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
      const double ddn = wrms(zn[q][lane]) * tq[1];
      etaqm1 = 1.0 / (rpow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    double etaqp1 = 0.0;
    if (q != qmax && saved_tq5 != 0.0) {
      const double cquot = (tq[5] / saved_tq5) * rpow(h / tau[2], (double)Lq);
      const double tv = acor - cquot * zn[qmax][lane];
      const double dup = wrms(tv) * tq[3];
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
      zn[qmax][lane] = acor;
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
    acor *= 1.0 / tq[2];
    return 0;
  }

  __device__ __forceinline__ double initial_h(double tout) {
    const double tdist = fabs(tout - tn);
    const double tround = UROUND * fmax(fabs(tn), fabs(tout));
    const double hlb = HLB_FACTOR * tround;
    const double den = HUB_FACTOR * fabs(zn[0][lane]) + 1.0 / ewt;
    const double hub_inv = (wave_max(act ? fabs(zn[1][lane]) / den : 0.0));
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = sqrt(hlb * hub);
    if (hub < hlb) return hg;
    double hnew = hg;
    bool hnewOK = false;
    for (int it = 1; it <= HIN_ITERS; ++it) {
      const double yt = hg * zn[1][lane] + zn[0][lane];
      double fv = f(yt);
      c_nfe++;
      fv = (fv - zn[1][lane]) * (1.0 / hg);
      const double yddnrm = wrms(fv);
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

  // integrate 0 -> tout; sol: this lane's component, updated in place
  __device__ __forceinline__ int integrate(double tout, double& sol) {
    c_nfe = c_nfe_dq = c_nje = c_nsetups = c_nni = c_netf = c_ncfn = c_nlu = 0;
    nst = 0;
    const bool finite = isfinite(act ? sol : 0.0);
    if (__ballot(!finite) != 0ull) return TGR_ST_BAD_INPUT;
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
    for (int j = 0; j < kQmaxAbs + 2; ++j) tau[j] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) tq[j] = 0.0;
#pragma unroll
    for (int j = 0; j <= kQmaxAbs; ++j) zn[j][lane] = 0.0;
#pragma unroll
    for (int c = 0; c < NEQ; ++c) M[c] = 0.0;
    pivl = 0;
    rdk = 0.0;
    acor = y = tempv = 0.0;
    zn[0][lane] = act ? sol : 0.0;
    set_ewt(zn[0][lane]);
    zn[1][lane] = f(zn[0][lane]);
    c_nfe++;
    const double tstop = tout;
    if (!(fabs(tout - tn) >= 2.0 * UROUND * fabs(tout)) || tout <= 0.0) return TGR_ST_BAD_INPUT;
    h = initial_h(tout);
    if ((tn + h - tstop) * h > 0.0) h = (tstop - tn) * (1.0 - 4.0 * UROUND);
    hscale = h;
    hprime = h;
    zn[1][lane] *= h;
    int nstloc = 0;
    for (;;) {
      if (nst > 0) set_ewt(zn[0][lane]);
      if (nstloc >= max_steps) {
        sol = zn[0][lane];
        return TGR_ST_TOO_MUCH_WORK;
      }
      if (UROUND * wrms(zn[0][lane]) > 1.0) {
        sol = zn[0][lane];
        return TGR_ST_TOO_MUCH_ACC;
      }
      const int k = step(tstop);
      if (k < 0) {
        sol = zn[0][lane];
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
        double v = zn[q][lane];
#pragma unroll
        for (int j = kQmaxAbs - 1; j >= 0; --j)
          if (j <= q - 1) v = zn[j][lane] + s * v;
        sol = v;
        return TGR_ST_OK;
      }
      if ((tn + hprime - tstop) * h > 0.0) {
        hprime = (tstop - tn) * (1.0 - 4.0 * UROUND);
        eta = hprime / h;
      }
    }
  }
  // ---- guided solve (surrogate-guided path) -------------------------------
  // nsub steps of sizes hs[0..nsub) (sum dt) of the three-stage, L-stable,
  // stiffly accurate SDIRK method of order 3 (Alexander 1977; gamma the root
  // of x^3 - 3x^2 + 3x/2 - 1/6 in (1/6, 1/2), c = (gamma, (1+gamma)/2, 1),
  // a21 = (1-gamma)/2, b = (-(6g^2-16g+1)/4, (6g^2-20g+5)/4, gamma)) with one
  // difference-quotient Jacobian at the initial state and one inverse of
  // I - gamma h J per sub-step.  Stage values are solved by simplified Newton
  // starting, with use_guess, from the previous stage value moved by the
  // network's change guess(i, j) - guess(previous stage) (sub-step i, stage j,
  // packed variables), else from the previous stage value.  Error estimate per
  // sub-step: the embedded second-order solution (weights bh = (1 - bh2, bh2,
  // 0), bh2 = (1/2 - gamma) / (c2 - gamma)) compared with the third-order one,
  // filtered by (I - gamma h J)^-1, in units of the acceptance tolerance
  // (wacc = 1 / tolerance of this lane's component).  Returns 0 (accepted,
  // yv = solution), 1 (a Newton solve did not converge in nmax iterations or
  // the matrix is singular), 2 (error estimate above 1).
  static constexpr double kSdG = 0.43586652150845899942;
  static constexpr double kSdC2 = 0.5 * (1.0 + kSdG);
  static constexpr double kSdA21 = 0.5 * (1.0 - kSdG);
  static constexpr double kSdB1 = -(6.0 * kSdG * kSdG - 16.0 * kSdG + 1.0) / 4.0;
  static constexpr double kSdB2 = (6.0 * kSdG * kSdG - 20.0 * kSdG + 5.0) / 4.0;
  static constexpr double kSdBh2 = (0.5 - kSdG) / (kSdC2 - kSdG);
  static constexpr double kSdBh1 = 1.0 - kSdBh2;

  // Simplified Newton for one stage value z = base + gh f(z), the iteration
  // matrix's inverse in M (getri).  The iterate's remaining error is about
  // theta / (1 - theta) times the last correction, theta the contraction rate
  // |d_k| / |d_k-1|, so the iteration stops when max(1, theta / (1 - theta))
  // |d| <= ntol (in acceptance tolerances): never looser than |d| <= ntol, and
  // tighter where a stale Jacobian makes the iteration contract slowly.  theta
  // enters as the previous solve's rate (the first correction has none of its
  // own) and leaves as this solve's last; thmax collects the largest rate
  // measured.  Fails on a non-finite or non-contracting (theta >= 0.99)
  // correction or after nmax iterations.
  __device__ __forceinline__ bool sdirk_newton(double& z, double base, double gh, double wacc, int nmax, double ntol,
                                               int& nit, double& theta, double& thmax) {
    double dprev = 0.0;
    for (int it = 0; it < nmax; ++it) {
      const double fz = f(z);
      c_nfe++;
      const double r = act ? z - base - gh * fz : 0.0;
      const double d = getri_solve(-r);
      nit++;
      if (act) z += d;
      const double dn = wave_max(act ? fabs(d) * wacc : 0.0);
      if (!isfinite(dn)) return false;
      if (it > 0) {
        theta = dn / dprev;
        thmax = fmax(thmax, theta);
        if (theta >= 0.99) return false;
      }
      const double th = fmin(theta, 0.9);
      if (fmax(1.0, th / (1.0 - th)) * dn <= ntol) return true;
      dprev = dn;
    }
    return false;
  }

  // The same method with step-size control: from h = h0 (or, with h0_fac > 0,
  // h0_fac / max_k |f_k(y0)| wacc_k within [dt/1024, dt]), each step's error
  // estimate e (as above) accepts the step when e <= 1 and scales the next
  // step by min(grow, max(0.2, 0.9 e^(-1/3))) -- after an accepted step a
  // factor in [1, keep) keeps h, so that the inverse of I - gamma h J is
  // reused -- a failed step is retried with the smaller h; the last step is
  // cut to end at dt.  One Jacobian at the initial state, one inverse per change
  // of h.  gat(t) is the network's prediction of
  // the state at time t; each stage's Newton starts from the previous stage
  // value moved by the prediction's change between the two stage times;
  // without a network each stage starts from its known part plus gamma h
  // times the previous derivative.  A stage whose Newton iteration does not
  // converge (or a singular iteration matrix) retries the step at h / 4.
  // Returns 0 (accepted), or after max_att attempted steps without reaching
  // dt (or a non-finite error estimate) 1 if the last failed attempt was a
  // Newton or singular-matrix failure, 2 if it failed the error test.
  //
  // Jacobian updates: when a step attempt's Newton iterations contracted
  // slower than jac_theta (or failed) with a Jacobian taken at an earlier
  // state, the next attempt first evaluates the Jacobian at the current state
  // and refactors, up to max_jac Jacobians per cell (1: the initial one only).
  template <class GuessAt>
  __device__ __forceinline__ int sdirk_adaptive(double& yv, double dt, double h0, double h0_fac, const GuessAt& gat,
                                                bool use_guess, double wacc, int nmax, double ntol, int max_att,
                                                double grow, double keep, double jac_theta, int max_jac, int& nit) {
    c_nfe = c_nfe_dq = c_nje = c_nsetups = c_nni = c_netf = c_ncfn = c_nlu = 0;
    nst = 0;
    const double y0 = act ? yv : 0.0;
    set_ewt(y0);
    double f0;
    if constexpr (AJ) {
      f0 = jac_analytic(y0);  // the Jacobian does not depend on the step size
    } else if (rc) {
      f0 = jac_split_base(y0);
    } else {
      f0 = f(y0);
      c_nfe++;
    }
    if (h0_fac > 0.0) {
      // first step from the initial rate: the step that moves the state by
      // h0_fac acceptance tolerances at that rate
      const double r0 = wave_max(act ? fabs(f0) * wacc : 0.0);
      h0 = (r0 > 0.0) ? h0_fac / r0 : dt;
      h0 = fmin(dt, fmax(dt * (1.0 / 1024.0), h0));
    }
    // the difference-quotient increments of either Jacobian follow the
    // first step actually taken
    h = h0;
    if constexpr (!AJ) {
      if (rc) jac_split_cols(y0, f0);
      else dq_jac(y0, f0);
    }
    c_nje++;
    double yi = y0, t = 0.0, hh = fmin(h0, dt);
    double kl = act ? f0 : 0.0;  // derivative at yi (the last stage's, the method being stiffly accurate)
    int why = 2;  // reason of the last failed attempt: 1 Newton / singular, 2 error test
    double h_lu = -1.0;  // step size of the factors held in M (-1: none)
    double theta = 0.0;  // Newton contraction rate, carried from solve to solve
    bool jfresh = true, refresh = false;  // Jacobian taken at yi; update it before the next attempt
    for (int att = 0; att < max_att; ++att) {
      if (dt - t <= 1.0e-12 * dt) {
        yv = yi;
        return 0;
      }
      if (refresh && c_nje < max_jac) {
        h = hh;  // the difference-quotient increments follow the step about to be taken
        if constexpr (AJ) {
          jac_analytic(yi);
        } else if (rc) {
          const double fi = jac_split_base(yi);
          jac_split_cols(yi, fi);
        } else {
          const double fi = f(yi);
          c_nfe++;
          dq_jac(yi, fi);
        }
        c_nje++;
        jfresh = true;
        h_lu = -1.0;
        theta = 0.0;
      }
      refresh = false;
      double thmax = 0.0;
      if (t + hh > dt || dt - (t + hh) < 1.0e-3 * hh) hh = dt - t;
      const double gh = kSdG * hh;
      if (hh != h_lu) {
        const TGR_GLOBAL double* jr = fresh(jw) + lane;
#pragma unroll
        for (int c = 0; c < NEQ; ++c) {
          double v = -gh * jr[c * kWave];
          if (c == lane) v += 1.0;
          M[c] = v;
        }
        c_nlu++;
        h_lu = -1.0;
        if (getri()) {
          hh *= 0.25;  // singular iteration matrix: retry with a smaller step
          why = 1;
          continue;
        }
        h_lu = hh;
      }
      const double rg = 1.0 / gh;
      const double ts[3] = {t + kSdG * hh, t + kSdC2 * hh, t + hh};
      // stage predictors: the network's change between the stage times, or
      // without a network the previous derivative carried over the stage
      double gp = use_guess ? gat(t) : 0.0;
      double gc = 0.0;
      double z = act ? yi + gh * kl : 0.0;
      if (use_guess) {
        gc = gat(ts[0]);
        z = yi + (gc - gp);
        gp = gc;
      }
      if (!sdirk_newton(z, yi, gh, wacc, nmax, ntol, nit, theta, thmax)) {
        hh *= 0.25;  // no convergence: retry with a smaller step
        why = 1;
        refresh = !jfresh;
        continue;
      }
      const double k1 = act ? (z - yi) * rg : 0.0;
      const double base2 = act ? yi + kSdA21 * hh * k1 : 0.0;
      if (use_guess) {
        gc = gat(ts[1]);
        z += gc - gp;
        gp = gc;
      } else {
        z = act ? base2 + gh * k1 : 0.0;
      }
      if (!sdirk_newton(z, base2, gh, wacc, nmax, ntol, nit, theta, thmax)) {
        hh *= 0.25;
        why = 1;
        refresh = !jfresh;
        continue;
      }
      const double k2 = act ? (z - base2) * rg : 0.0;
      const double base3 = act ? yi + hh * (kSdB1 * k1 + kSdB2 * k2) : 0.0;
      if (use_guess) {
        gc = gat(ts[2]);
        z += gc - gp;
      } else {
        z = act ? base3 + gh * k2 : 0.0;
      }
      if (!sdirk_newton(z, base3, gh, wacc, nmax, ntol, nit, theta, thmax)) {
        hh *= 0.25;
        why = 1;
        refresh = !jfresh;
        continue;
      }
      const double k3 = act ? (z - base3) * rg : 0.0;
      const double e =
          getri_solve(act ? hh * ((kSdB1 - kSdBh1) * k1 + (kSdB2 - kSdBh2) * k2 + kSdG * k3) : 0.0);
      const double en = wave_max(act ? fabs(e) * wacc : 0.0);
      if (!isfinite(en)) return 1;
      double fac = fmin(grow, fmax(0.2, 0.9 * rpow(fmax(en, 1.0e-10), -1.0 / 3.0)));
      if (en <= 1.0) {
        nst++;
        t += hh;
        yi = z;
        kl = k3;
        jfresh = false;
        if (fac >= 1.0 && fac < keep) fac = 1.0;  // keep h and reuse the factors
      } else {
        why = 2;
      }
      refresh = thmax > jac_theta && !jfresh;
      hh *= fac;
    }
    if (dt - t <= 1.0e-12 * dt) {
      yv = yi;
      return 0;
    }
    return why;
  }

  template <class Guess>
  __device__ __forceinline__ int sdirk_guided(double& yv, int nsub, const double* hs, const Guess& guess,
                                              bool use_guess, double wacc, int nmax, double ntol, int& nit) {
    c_nfe = c_nfe_dq = c_nje = c_nsetups = c_nni = c_netf = c_ncfn = c_nlu = 0;
    nst = 0;
    h = hs[0];
    const double y0 = act ? yv : 0.0;
    set_ewt(y0);
    if constexpr (AJ) {
      jac_analytic(y0);
    } else if (rc) {
      jac_split(y0);
    } else {
      const double f0 = f(y0);
      c_nfe++;
      dq_jac(y0, f0);
    }
    c_nje++;
    double yi = y0;
    double theta = 0.0, thmax = 0.0;  // Newton contraction rate (carried), largest seen
    // Newton starts from the previous stage's converged value plus the
    // network's predicted change between the two stage times (its bias,
    // smooth in time, cancels); gp: the network's value at the previous stage
    double gp = y0;
    auto start = [&](int i, int j, double zprev) -> double {
      if (!use_guess) return zprev;
      const double gc = guess(i, j);
      const double z0 = zprev + (gc - gp);
      gp = gc;
      return z0;
    };
    double h_lu = -1.0;  // step size of the factors held in M (equal sub-steps share them)
    for (int i = 0; i < nsub; ++i) {
      const double hh = hs[i];
      const double gh = kSdG * hh;
      if (hh != h_lu) {
        const TGR_GLOBAL double* jr = fresh(jw) + lane;
#pragma unroll
        for (int c = 0; c < NEQ; ++c) {
          double v = -gh * jr[c * kWave];
          if (c == lane) v += 1.0;
          M[c] = v;
        }
        c_nlu++;
        if (getri()) return 1;
        h_lu = hh;
      }
      const double rg = 1.0 / gh;
      double z = start(i, 0, yi);
      if (!sdirk_newton(z, yi, gh, wacc, nmax, ntol, nit, theta, thmax)) return 1;
      const double k1 = act ? (z - yi) * rg : 0.0;
      const double base2 = act ? yi + kSdA21 * hh * k1 : 0.0;
      z = start(i, 1, z);
      if (!sdirk_newton(z, base2, gh, wacc, nmax, ntol, nit, theta, thmax)) return 1;
      const double k2 = act ? (z - base2) * rg : 0.0;
      const double base3 = act ? yi + hh * (kSdB1 * k1 + kSdB2 * k2) : 0.0;
      z = start(i, 2, z);
      if (!sdirk_newton(z, base3, gh, wacc, nmax, ntol, nit, theta, thmax)) return 1;
      const double k3 = act ? (z - base3) * rg : 0.0;
      const double e =
          getri_solve(act ? hh * ((kSdB1 - kSdBh1) * k1 + (kSdB2 - kSdBh2) * k2 + kSdG * k3) : 0.0);
      const double en = wave_max(act ? fabs(e) * wacc : 0.0);
      nst++;
      if (!(en <= 1.0)) return 2;
      yi = z;
    }
    yv = yi;
    return 0;
  }
};

// One cell of finitechem_source_chemistry (finitechem.f90:996-1058): pack
// sol = [max(RHO*Y, 0), RHO*T], integrate 0 -> dt, unpack by dividing by
// RHO_, heat release and H2O/CO/CO2 sources.  Returns 0 for a wall cell
// (mask == 1, skipped), else 1 with the integrator counters in cnt.
template <int NEQ, bool AJ = false>
__device__ __forceinline__ int process_cell(const StepParams& p, const DevMech& m, tchar* tab, TGR_LDS RhsLds<NEQ>* rl,
                                            int64_t cell, TGR_LDS double (&zn_lds)[kQmaxAbs + 2][kWave], double* ws,
                                            int* cnt) {
  const int lane = lane_id();
  const int nsp = NEQ - 1;
  const int i = (int)(cell % p.nx);
  const int j = (int)((cell / p.nx) % p.ny);
  const int k = (int)(cell / ((int64_t)p.nx * p.ny));
  bool skip = false;
  if (p.mask) skip = p.mask[p.mask_base + (int64_t)j * p.mask_ld + i] == 1;
  if (skip) return 0;
  {
    const int64_t off = p.base + i * p.sx + j * p.sy + k * p.sz;
    const double rho = p.rho[off];
    double sol = 0.0;
    if (lane < nsp) sol = fmax(rho * p.sc[(int64_t)(p.isc_sc0 + lane) * p.plane + off], 0.0);
    else if (lane == nsp) sol = rho * p.sc[(int64_t)p.isc_T0 * p.plane + off];
    const double sol0 = sol;
    Bdf<NEQ, AJ> b(m, tab, rl, zn_lds, ws, p.p_gp, p.rtol, p.atol, p.qmax, p.max_steps);
    const int st = b.integrate(p.dt, sol);
    // unpack (finitechem.f90:1041-1042) and diagnostics (1047-1058)
    const double v = sol / rho;
    if (lane < nsp) p.sc[(int64_t)(p.isc_sc0 + lane) * p.plane + off] = v;
    else if (lane == nsp) p.sc[(int64_t)p.isc_T0 * p.plane + off] = v;
    const double Tn = bcast(v, nsp);
    double h, cp;
    thermo_lane<NEQ>(tab, m.tmid, Tn, lane, h, cp);
    const double cpm = (wave_sum(lane < nsp ? v * cp : 0.0));
    if (p.hr && lane == nsp) p.hr[off] = cpm * (sol - sol0) / p.dt;
    if (p.chemsrc) {
      const double src = (sol - sol0) / p.dt;
      const int sp3[3] = {p.idx_h2o, p.idx_co, p.idx_co2};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (sp3[q] >= 0 && lane == sp3[q]) p.chemsrc[q * p.plane + off] = src;
        if (sp3[q] < 0 && lane == nsp) p.chemsrc[q * p.plane + off] = 0.0;
      }
    }
    cnt[0] = b.nst;
    cnt[1] = b.c_nfe;
    cnt[2] = b.c_nfe_dq;
    cnt[3] = b.c_nje;
    cnt[4] = b.c_nsetups;
    cnt[5] = b.c_nni;
    cnt[6] = b.c_netf;
    cnt[7] = b.c_ncfn;
    cnt[8] = b.c_nlu;
    cnt[9] = st;
  }
  return 1;
}

}  // namespace tgr

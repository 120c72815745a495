// Device-side mechanism tables and the wave-per-cell chemistry right-hand side
// for CDNA4 (gfx950, wave64).
//
// One wavefront owns one cell.  Lane i < nsp holds species i (rho*Y_i), lane
// nsp holds rho*T; the remaining lanes hold 0.  The production rates of the
// generated PRODRATES (reference combustion/HYDROGENF.f90:4-277,
// grimech30F.f90:4-2580, DME39F.f90:4-1943) are evaluated table-driven:
//   * lane j < n_third forms the third-body sum M_j = sum_k eff_jk C_k as the
//     total concentration plus a short list of (eff_jk - 1) corrections
//     (most efficiencies are 1), reading concentrations broadcast in LDS;
//   * reactions are dealt to lanes round-robin (fall-off reactions first so
//     that only the first round diverges), each lane computing k_f, k_b
//     (Arrhenius / GETLINDRATECOEFF Troe form, HYDROGENF.f90:280-316) and
//     the net rate of progress q = (W_f - W_b) [* M] into LDS;
//   * CDOT is gathered deterministically: the (reaction, nu) entries of each
//     species are split into lane-sized chunks, summed per lane, then summed
//     per species (fixed order, no atomics).
// The right-hand side then follows finitechem_compute_rhs
// (combustion/finitechem.f90:93-175): rho = sum(rho*Y), T = rhoT/rho,
// C = rho*Y/W, rhs_k = CDOT_k*W_k, rhs_T = -sum(h_k rhs_k)/cp_mix.
//
// The tables are one block whose layout is a compile-time function of the
// mechanism shape (MechT), structure-of-arrays with one 64-lane row per
// field, so that every access of a wave is one conflict-free LDS read.  The
// kernels copy the block into LDS once per workgroup.  Records are padded so
// that the reaction loop has no data-dependent branches except the fall-off
// form: unused concentration factors and the third-body factor of reactions
// without one point to an LDS slot holding 1.0, irreversible reactions carry
// A = 0, and padding entries of the gather and correction lists have weight 0.
#pragma once
#include "wave_target.h"
#include <stdint.h>

#include "fp64_math.h"

namespace tgr {

constexpr int kWave = 64;
constexpr int kMaxNeq = 64;
// LDS slots of the concentration / third-body vectors holding 1.0 and 0.0
constexpr int kOne = 63;
constexpr int kZero = 62;

// binary32 values of the reference's unsuffixed literals (Makefile.in:19 sets
// no real promotion): R = 8314.34 and the Troe constants of GETLINDRATECOEFF.
constexpr double kRgasF4 = (double)8314.34f;
constexpr double kTroeN0 = 0.75;
constexpr double kTroeN1 = (double)1.27f;
constexpr double kTroeC0 = (double)0.4f;
constexpr double kTroeC1 = (double)0.67f;
constexpr double kTroeD = (double)0.14f;
constexpr double kInvLn10 = 0.43429448190325182765;

// Compile-time shape of each compiled mechanism, keyed by neq = nsp + 1.  The
// host builds the tables for exactly this shape and refuses a mismatch.
//   gmax: longest CDOT gather chunk per lane; maxe: longest third-body
//   correction list.
template <int NEQ>
struct MechT;
template <>
struct MechT<10> {  // h2 (HYDROGENF)
  static constexpr int nrxn = 25, nfo = 2, nthird = 5, gmax = 2, maxe = 3;
};
template <>
struct MechT<37> {  // gri30 (grimech30F, no N chemistry)
  static constexpr int nrxn = 219, nfo = 26, nthird = 32, gmax = 17, maxe = 8;
};
template <>
struct MechT<40> {  // dme39 (DME39F)
  static constexpr int nrxn = 175, nfo = 11, nthird = 20, gmax = 15, maxe = 6;
};
template <>
struct MechT<46> {  // ethyred (EthyRedF, the mechanism combustion/Makefile:3 links by default)
  static constexpr int nrxn = 193, nfo = 18, nthird = 7, gmax = 18, maxe = 6;
};

// Byte layout of the table block for a mechanism shape (rows of 64 lanes:
// a double row is 512 bytes, an int row 256).
//   RX  per 64-reaction round: double rows A_f b_f e_f A_b b_b e_b, int rows
//       I0 = r0 | r1 << 8 | r2 << 16 | m << 24 (reactant slots, third-body slot)
//       I1 = p0 | p1 << 8 | p2 << 16 | falloff << 24 | rev << 25
//   FO  (round-0 lanes < nfo) double rows A0_f b0_f e0_f A0_b b0_b e0_b c0 c1 c2
//       g0 g1 g2 and an int row kinds | mfo << 8: Fc = sum_i c_i exp(-g_i T) or,
//       where bit i of kinds is set, c_i exp(-g_i / T); mfo: slot of the
//       third-body sum of GETLINDRATECOEFF (kZero: none, conc = P/RT)
//   TH  [2][13] double rows (range, coefficient), species in lanes
//   W, IW  double rows: molar mass and its reciprocal
//   EV, EK  [maxe] double / int rows: third-body corrections (eff - 1, slot)
//   GN, GR  [gmax] double / int rows: CDOT gather (nu, q index)
//   SF, SN  int rows: first gather lane and number of chunks of each species
constexpr size_t kRow = 512, kIRow = 256;
constexpr size_t kRxRound = 6 * kRow + 2 * kIRow;
struct TabLayout {
  int nsp, nrxn, rounds, nfo, nthird, gmax, maxe;
  size_t rx, fo, th, w, iw, ev, ek, gn, gr, sf, sn, bytes;
  __host__ __device__ constexpr TabLayout(int nsp_, int nrxn_, int nfo_, int nthird_, int gmax_, int maxe_)
      : nsp(nsp_), nrxn(nrxn_), rounds((nrxn_ + 63) / 64), nfo(nfo_), nthird(nthird_), gmax(gmax_), maxe(maxe_),
        rx(0), fo(rx + (size_t)((nrxn_ + 63) / 64) * kRxRound), th(fo + 12 * kRow + kIRow), w(th + 26 * kRow),
        iw(w + kRow), ev(iw + kRow), ek(ev + (size_t)maxe_ * kRow), gn(ek + (size_t)maxe_ * kIRow),
        gr(gn + (size_t)gmax_ * kRow), sf(gr + (size_t)gmax_ * kIRow), sn(sf + kIRow), bytes(sn + kIRow) {}
};

template <int NEQ>
__host__ __device__ constexpr TabLayout tab_layout() {
  return TabLayout(NEQ - 1, MechT<NEQ>::nrxn, MechT<NEQ>::nfo, MechT<NEQ>::nthird, MechT<NEQ>::gmax,
                   MechT<NEQ>::maxe);
}

// Device view of a mechanism: the table block (HBM) plus a few sizes.
struct DevMech {
  const char* tab;
  int nsp;
  double tmid;
};

// Table block as the kernels read it (LDS).
using tchar = const TGR_LDS char;

// Per-wave LDS scratch for one right-hand-side evaluation.
template <int NEQ>
struct RhsLds {
  double C[kWave];
  double M[kWave];
  double part[kWave];
  double q[((MechT<NEQ>::nrxn + 63) / 64) * 64];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Intra-wave ordering point for LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP move of a double (both halves), lanes without a source get `old`.
template <int CTRL, int ROWMASK, int BANKMASK, bool BOUND0>
__device__ __forceinline__ double dpp(double old, double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROWMASK, BANKMASK, BOUND0);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROWMASK, BANKMASK, BOUND0);
  return __hiloint2double(hi, lo);
}

// value of v in lane `src` (src wave-uniform), as a wave-uniform scalar
__device__ __forceinline__ double bcast(double v, int src) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes with the row_shr / row_bcast DPP tree (total lands
// in lane 63), returned wave-uniform.  Every level writes all lanes, lanes
// without a source adding 0 (bound_ctrl): lanes other than 63 then hold
// partial sums that differ from a masked tree, but lane 63 sees the same
// operands in the same order, so the result is the masked tree's, bit for
// bit, without the register copies that masked levels need.  Fixed order:
// deterministic.
__device__ __forceinline__ double wave_sum(double v) {
  double s = v + dpp<0x111, 0xf, 0xf, true>(0.0, v);  // row_shr:1
  s = s + dpp<0x112, 0xf, 0xf, true>(0.0, v);         // row_shr:2
  s = s + dpp<0x113, 0xf, 0xf, true>(0.0, v);         // row_shr:3
  s = s + dpp<0x114, 0xf, 0xf, true>(0.0, s);         // row_shr:4
  s = s + dpp<0x118, 0xf, 0xf, true>(0.0, s);         // row_shr:8
  s = s + dpp<0x142, 0xf, 0xf, true>(0.0, s);         // row_bcast:15
  s = s + dpp<0x143, 0xf, 0xf, true>(0.0, s);         // row_bcast:31
  return bcast(s, 63);
}

// Maximum over the 64 lanes of values >= -1 (same tree; lanes without a
// source take 0, which never reaches lane 63).
__device__ __forceinline__ double wave_max(double v) {
  double s = fmax(v, dpp<0x111, 0xf, 0xf, true>(0.0, v));
  s = fmax(s, dpp<0x112, 0xf, 0xf, true>(0.0, v));
  s = fmax(s, dpp<0x113, 0xf, 0xf, true>(0.0, v));
  s = fmax(s, dpp<0x114, 0xf, 0xf, true>(0.0, s));
  s = fmax(s, dpp<0x118, 0xf, 0xf, true>(0.0, s));
  s = fmax(s, dpp<0x142, 0xf, 0xf, true>(0.0, s));
  s = fmax(s, dpp<0x143, 0xf, 0xf, true>(0.0, s));
  return bcast(s, 63);
}

// Returns the pointer through an empty asm statement so that loads through
// it are issued where they are used: without it the compiler hoists every
// loop-invariant load out of the solver loops and keeps hundreds of values
// live in registers across the whole integration.  The result is qualified
// as device memory so that the loads are global (a laundered generic pointer
// would give flat loads).
template <class T>
__device__ __forceinline__ TGR_GLOBAL T* fresh(T* p) {
  asm volatile("" : TGR_OPAQUE_PTR(p));
  return (TGR_GLOBAL T*)p;
}

template <class T>
__device__ __forceinline__ T ld(tchar* base, size_t off) {
  return *reinterpret_cast<const TGR_LDS T*>(base + off);
}

#ifndef TGR_WAVE_EMU
// Copies the table block from HBM into LDS (all threads of the workgroup,
// 16-byte chunks); the caller synchronises the workgroup afterwards.
__device__ __forceinline__ void stage_tables(TGR_LDS char* dst, const char* src, size_t bytes) {
  const TGR_GLOBAL uint64_t* s8 = (const TGR_GLOBAL uint64_t*)src;
  TGR_LDS uint64_t* d8 = (TGR_LDS uint64_t*)dst;
  for (size_t i = threadIdx.x; i < bytes / 16; i += blockDim.x) {
    const uint64_t a = s8[2 * i], b = s8[2 * i + 1];
    d8[2 * i] = a;
    d8[2 * i + 1] = b;
  }
}
#endif

__device__ __forceinline__ double arrhenius(double A, double b, double e, double lt, double irt) {
  return A * fexp(fma(b, lt, e * irt));
}

// NASA-7 enthalpy / heat capacity of the species in this lane
// (COMPTHERMODATA, HYDROGENF.f90:514-673); lanes >= nsp return 0.
template <int NEQ>
__device__ __forceinline__ void thermo_lane(tchar* tab, double tmid, double T, int lane, double& h, double& cp) {
  constexpr TabLayout L = tab_layout<NEQ>();
  tchar* c = tab + L.th + (T > tmid ? 0 : 13 * kRow) + lane * 8;
  const double a = ld<double>(c, 0), c1 = ld<double>(c, kRow), c2 = ld<double>(c, 2 * kRow),
               c3 = ld<double>(c, 3 * kRow), c4 = ld<double>(c, 4 * kRow), c5 = ld<double>(c, 5 * kRow),
               c6 = ld<double>(c, 6 * kRow);
  const double ac = ld<double>(c, 7 * kRow), d1 = ld<double>(c, 8 * kRow), d2 = ld<double>(c, 9 * kRow),
               d3 = ld<double>(c, 10 * kRow), d4 = ld<double>(c, 11 * kRow), d5 = ld<double>(c, 12 * kRow);
  h = a * (T * (c1 + T * (c2 + T * (c3 + T * (c4 + T * c5)))) + c6);
  cp = ac * (d1 + T * (d2 + T * (d3 + T * (d4 + T * d5))));
  if (lane >= NEQ - 1) h = cp = 0.0;
}

// Troe fall-off rate (GETLINDRATECOEFF, HYDROGENF.f90:280-316, ITROE = 1):
// k = k_inf * F * Pr / (1 + Pr), Pr = k0 * conc / max(k_inf, 1e-60),
// log10 F = log10 Fc / (1 + ((log10 Pr + c) / (n - d (log10 Pr + c)))^2).
__device__ __forceinline__ double troe(double kinf, double k0, double conc, double lnfc, double ntmp, double cco) {
  const double pr = k0 * conc / fmax(kinf, 1.0e-60);
  const double lg = fma(flog(pr), kInvLn10, cco);
  const double x = lg / fma(-kTroeD, lg, ntmp);
  const double F = fexp(lnfc / fma(x, x, 1.0));
  return kinf * F * pr / (1.0 + pr);
}

// Temperature-dependent factors of one right-hand side, kept for others at
// the same temperature (the species columns of a difference-quotient
// Jacobian taken at fixed T): rate constants of every reaction lane (the
// high-pressure limits on fall-off lanes), the low-pressure limits and
// ln Fcent of the fall-off lanes, and each species' h and cp.  Every lane
// reads back only what it wrote.
template <int NEQ>
struct RateCache {
  double kf[((MechT<NEQ>::nrxn + 63) / 64) * 64];
  double kb[((MechT<NEQ>::nrxn + 63) / 64) * 64];
  double k0f[kWave], k0b[kWave], lnfc[kWave];
  double h[kWave], cp[kWave];
};

// Production rates CDOT (kmol/m3/s) of this lane's species from the
// concentrations (lane k holds C_k), temperature and pressure.  ctot is the
// total concentration sum_k C_k (wave-uniform).  MODE 1 also saves the
// temperature-dependent factors into rc; MODE 2 takes them from rc instead of
// evaluating them (T must be the temperature they were saved at).
template <int NEQ, int MODE = 0>
__device__ __forceinline__ double prodrates_wave(tchar* tab, TGR_LDS RhsLds<NEQ>& S, double Ck, double ctot,
                                                 double T, double P, TGR_LDS RateCache<NEQ>* rc = nullptr) {
  using MT = MechT<NEQ>;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const double lt = MODE == 2 ? 0.0 : flog(T);
  const double rt = kRgasF4 * T;
  const double irt = 1.0 / rt;
  S.C[lane] = lane == kOne ? 1.0 : Ck;
  wave_sync();
  // third-body sums: M_j = ctot + sum_e (eff_je - 1) C_e (padding: weight 0 on slot kOne)
  {
    double s = ctot;
#pragma unroll
    for (int e = 0; e < MT::maxe; ++e)
      s = fma(ld<double>(tab, L.ev + e * kRow + lane * 8), S.C[ld<int>(tab, L.ek + e * kIRow + lane * 4)], s);
    S.M[lane] = lane == kOne ? 1.0 : (lane == kZero ? 0.0 : s);
  }
  wave_sync();
  // rates of progress, one reaction per lane per round (fall-off in round 0)
#pragma unroll
  for (int rd = 0; rd < L.rounds; ++rd) {
    tchar* rx = tab + L.rx + rd * kRxRound + lane * 8;
    tchar* rxi = tab + L.rx + rd * kRxRound + 6 * kRow + lane * 4;
    const int i0 = ld<int>(rxi, 0), i1 = ld<int>(rxi, kIRow);
    double kf, kb;
    if constexpr (MODE == 2) {
      kf = rc->kf[rd * 64 + lane];
      kb = rc->kb[rd * 64 + lane];
    } else {
      kf = arrhenius(ld<double>(rx, 0), ld<double>(rx, kRow), ld<double>(rx, 2 * kRow), lt, irt);
      kb = arrhenius(ld<double>(rx, 3 * kRow), ld<double>(rx, 4 * kRow), ld<double>(rx, 5 * kRow), lt, irt);
      if constexpr (MODE == 1) {
        rc->kf[rd * 64 + lane] = kf;
        rc->kb[rd * 64 + lane] = kb;
      }
    }
    if (rd == 0 && ((i1 >> 24) & 1)) {
      tchar* fo = tab + L.fo + lane * 8;
      const int fi = ld<int>(tab, L.fo + 12 * kRow + lane * 4);
      const double mv = S.M[(fi >> 8) & 255];
      const double conc = mv > 0.0 ? mv : P / rt;
      if constexpr (MODE == 2) {
        const double lnfc = rc->lnfc[lane];
        const double lfc = lnfc * kInvLn10;
        const double ntmp = fma(-kTroeN1, lfc, kTroeN0);
        const double cco = fma(-kTroeC1, lfc, -kTroeC0);
        kf = troe(kf, rc->k0f[lane], conc, lnfc, ntmp, cco);
        if ((i1 >> 25) & 1) kb = troe(kb, rc->k0b[lane], conc, lnfc, ntmp, cco);
      } else {
        const double invT = 1.0 / T;
        double fc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double x = -(ld<double>(fo, (9 + i) * kRow) * (((fi >> i) & 1) ? invT : T));
          fc = fma(ld<double>(fo, (6 + i) * kRow), fexp(x), fc);
        }
        const double lnfc = flog(fc);
        const double lfc = lnfc * kInvLn10;
        const double ntmp = fma(-kTroeN1, lfc, kTroeN0);
        const double cco = fma(-kTroeC1, lfc, -kTroeC0);
        const double k0f = arrhenius(ld<double>(fo, 0), ld<double>(fo, kRow), ld<double>(fo, 2 * kRow), lt, irt);
        if constexpr (MODE == 1) {
          rc->lnfc[lane] = lnfc;
          rc->k0f[lane] = k0f;
        }
        kf = troe(kf, k0f, conc, lnfc, ntmp, cco);
        if ((i1 >> 25) & 1) {
          const double k0b =
              arrhenius(ld<double>(fo, 3 * kRow), ld<double>(fo, 4 * kRow), ld<double>(fo, 5 * kRow), lt, irt);
          if constexpr (MODE == 1) rc->k0b[lane] = k0b;
          kb = troe(kb, k0b, conc, lnfc, ntmp, cco);
        }
      }
    }
    const double wf = kf * S.C[i0 & 255] * S.C[(i0 >> 8) & 255] * S.C[(i0 >> 16) & 255];
    const double wb = kb * S.C[i1 & 255] * S.C[(i1 >> 8) & 255] * S.C[(i1 >> 16) & 255];
    S.q[rd * 64 + lane] = (wf - wb) * S.M[(i0 >> 24) & 255];
  }
  wave_sync();
  // deterministic gather of CDOT
  {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < MT::gmax; ++i)
      acc = fma(ld<double>(tab, L.gn + i * kRow + lane * 8), S.q[ld<int>(tab, L.gr + i * kIRow + lane * 4)], acc);
    S.part[lane] = acc;
  }
  wave_sync();
  double cdot = 0.0;
  if (lane < NEQ - 1) {
    const int f0 = ld<int>(tab, L.sf + lane * 4), nn = ld<int>(tab, L.sn + lane * 4);
    for (int i = 0; i < nn; ++i) cdot += S.part[f0 + i];
  }
  wave_sync();
  return cdot;
}

// finitechem_compute_rhs (finitechem.f90:93-175), constant-pressure branch.
// y: this lane's component of [rho*Y (nsp), rho*T]; returns the same lane of
// the right-hand side.
template <int NEQ>
__device__ __forceinline__ double rhs_wave(tchar* tab, TGR_LDS RhsLds<NEQ>& S, double tmid, double y, double p_gp) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const bool sp = lane < nsp;
  const double rho = wave_sum(sp ? y : 0.0);
  const double T = bcast(y, nsp) / rho;
  double h, cp;
  thermo_lane<NEQ>(tab, tmid, T, lane, h, cp);
  const double Ck = sp ? y * ld<double>(tab, L.iw + lane * 8) : 0.0;
  const double cpm = wave_sum(sp ? (y / rho) * cp : 0.0);
  const double ctot = wave_sum(Ck);
  const double cdot = prodrates_wave<NEQ>(tab, S, Ck, ctot, T, p_gp);
  const double rk = sp ? cdot * ld<double>(tab, L.w + lane * 8) : 0.0;
  const double tmp = wave_sum(-h * rk);
  return sp ? rk : (lane == nsp ? tmp / cpm : 0.0);
}

// The right-hand side with the temperature given (instead of y[nsp] / rho);
// MODE as in prodrates_wave, MODE 2 also takes h and cp from rc.
template <int NEQ, int MODE>
__device__ __forceinline__ double rhs_wave_T(tchar* tab, TGR_LDS RhsLds<NEQ>& S, TGR_LDS RateCache<NEQ>* rc,
                                             double tmid, double y, double p_gp, double T) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const bool sp = lane < nsp;
  const double rho = wave_sum(sp ? y : 0.0);
  double h, cp;
  if (MODE == 2) {
    h = rc->h[lane];
    cp = rc->cp[lane];
  } else {
    thermo_lane<NEQ>(tab, tmid, T, lane, h, cp);
    if (MODE == 1) {
      rc->h[lane] = h;
      rc->cp[lane] = cp;
    }
  }
  const double Ck = sp ? y * ld<double>(tab, L.iw + lane * 8) : 0.0;
  const double cpm = wave_sum(sp ? (y / rho) * cp : 0.0);
  const double ctot = wave_sum(Ck);
  const double cdot = prodrates_wave<NEQ, MODE>(tab, S, Ck, ctot, T, p_gp, rc);
  const double rk = sp ? cdot * ld<double>(tab, L.w + lane * 8) : 0.0;
  const double tmp = wave_sum(-h * rk);
  return sp ? rk : (lane == nsp ? tmp / cpm : 0.0);
}

// ---- analytic Jacobian ----------------------------------------------------
// The reference has no Jacobian routine (CVODE differences the right-hand
// side, finitechem.f90:1028).  This is the exact derivative of the same
// rate expressions, evaluated from the same tables: per reaction lane and
// round, the coefficients that one column of dCDOT/dC needs are kept in
// registers, so that each species column is a pass over the reaction rounds
// (no exp / log) plus the deterministic CDOT gather.
//   q_r = (k_f P_f - k_b P_b) mm,  P = product of the concentration factors,
//   mm = M_r for a third-body reaction, else 1;
//   dq_r/dC_j = mm (k_f dP_f/dC_j - k_b dP_b/dC_j) + g_r eff(m_r, j)
// with g_r = k_f P_f - k_b P_b for a third-body reaction and, for a fall-off
// reaction whose third-body sum is positive (conc = M_r, GETLINDRATECOEFF
// HYDROGENF.f90:295-299), g_r = dk_f/dconc P_f - dk_b/dconc P_b.

// NASA-7 h, cp and dcp/dT of the species in this lane (COMPTHERMODATA forms);
// lanes >= nsp return 0.
template <int NEQ>
__device__ __forceinline__ void thermo_lane_d(tchar* tab, double tmid, double T, int lane, double& h, double& cp,
                                              double& dcp) {
  constexpr TabLayout L = tab_layout<NEQ>();
  tchar* c = tab + L.th + (T > tmid ? 0 : 13 * kRow) + lane * 8;
  const double a = ld<double>(c, 0), c1 = ld<double>(c, kRow), c2 = ld<double>(c, 2 * kRow),
               c3 = ld<double>(c, 3 * kRow), c4 = ld<double>(c, 4 * kRow), c5 = ld<double>(c, 5 * kRow),
               c6 = ld<double>(c, 6 * kRow);
  const double ac = ld<double>(c, 7 * kRow), d1 = ld<double>(c, 8 * kRow), d2 = ld<double>(c, 9 * kRow),
               d3 = ld<double>(c, 10 * kRow), d4 = ld<double>(c, 11 * kRow), d5 = ld<double>(c, 12 * kRow);
  h = a * (T * (c1 + T * (c2 + T * (c3 + T * (c4 + T * c5)))) + c6);
  cp = ac * (d1 + T * (d2 + T * (d3 + T * (d4 + T * d5))));
  dcp = ac * (d2 + T * (2.0 * d3 + T * (3.0 * d4 + T * (4.0 * d5))));
  if (lane >= NEQ - 1) h = cp = dcp = 0.0;
}

// Troe rate (as troe()) with its derivatives: dk_dT at fixed conc through
// k_inf, k_0, Fc (and conc = P/RT when dlconc = -1/T), dk_dconc at fixed T.
// dl*: logarithmic derivatives d ln(.)/dT.
__device__ __forceinline__ double troe_d(double kinf, double dlkinf, double k0, double dlk0, double conc,
                                         double dlconc, double lnfc, double dlfc, double ntmp, double cco,
                                         double& dk_dT, double& dk_dconc) {
  const double kin = fmax(kinf, 1.0e-60);
  const double pr = k0 * conc / kin;
  const double lg = fma(flog(pr), kInvLn10, cco);
  const double D = fma(-kTroeD, lg, ntmp);
  const double x = lg / D;
  const double x2 = fma(x, x, 1.0);
  const double F = fexp(lnfc / x2);
  const double k = kinf * F * pr / (1.0 + pr);  // as troe(), bit for bit
  const double ix2 = 1.0 / x2;
  const double ipr1 = 1.0 / (1.0 + pr);
  const double dlpr = dlk0 + dlconc - (kinf > 1.0e-60 ? dlkinf : 0.0);
  const double dL = (dlpr - kTroeC1 * dlfc) * kInvLn10;
  const double dn = -kTroeN1 * dlfc * kInvLn10;
  const double iD2 = 1.0 / (D * D);
  const double dx = (ntmp * dL - lg * dn) * iD2;
  const double c2x = lnfc * 2.0 * x * ix2 * ix2;
  dk_dT = k * (dlkinf + dlfc * ix2 - c2x * dx + dlpr * ipr1);
  const double dlFp = -c2x * ntmp * kInvLn10 * iD2;
  dk_dconc = kinf * F * (dlFp * ipr1 + ipr1 * ipr1) * (k0 / kin);
  return k;
}

// What one reaction lane keeps for the Jacobian columns: the products of two
// of its three concentration factors times mm k (forward a**, backward b**),
// the third-body coefficient g and the packed factor slots; md: the slot of
// the third-body sum g multiplies (kOne: none).
struct JacLane {
  double a01, a02, a12, b01, b02, b12, g;
  int i0, i1, md;
};

// Production rates (same values as prodrates_wave) plus dCDOT/dT at fixed C
// in dcdT, filling jl for jac_column.  ctot: total concentration.
template <int NEQ>
__device__ __forceinline__ double prodrates_jac_base(tchar* tab, TGR_LDS RhsLds<NEQ>& S, double Ck, double ctot,
                                                     double T, double P, JacLane (&jl)[(MechT<NEQ>::nrxn + 63) / 64],
                                                     double& dcdT) {
  using MT = MechT<NEQ>;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const double lt = flog(T);
  const double rt = kRgasF4 * T;
  const double irt = 1.0 / rt;
  const double invT = 1.0 / T;
  S.C[lane] = lane == kOne ? 1.0 : Ck;
  wave_sync();
  {
    double s = ctot;
#pragma unroll
    for (int e = 0; e < MT::maxe; ++e)
      s = fma(ld<double>(tab, L.ev + e * kRow + lane * 8), S.C[ld<int>(tab, L.ek + e * kIRow + lane * 4)], s);
    S.M[lane] = lane == kOne ? 1.0 : (lane == kZero ? 0.0 : s);
  }
  wave_sync();
  double qT[L.rounds];
#pragma unroll
  for (int rd = 0; rd < L.rounds; ++rd) {
    tchar* rx = tab + L.rx + rd * kRxRound + lane * 8;
    tchar* rxi = tab + L.rx + rd * kRxRound + 6 * kRow + lane * 4;
    const int i0 = ld<int>(rxi, 0), i1 = ld<int>(rxi, kIRow);
    const double bf = ld<double>(rx, kRow), ef = ld<double>(rx, 2 * kRow);
    const double bb = ld<double>(rx, 4 * kRow), eb = ld<double>(rx, 5 * kRow);
    double kf = arrhenius(ld<double>(rx, 0), bf, ef, lt, irt);
    double kb = arrhenius(ld<double>(rx, 3 * kRow), bb, eb, lt, irt);
    // d ln k / dT of the Arrhenius forms: (b - e/(R T)) / T
    double dkf = kf * (fma(-ef, irt, bf) * invT), dkb = kb * (fma(-eb, irt, bb) * invT);
    double dkfc = 0.0, dkbc = 0.0;
    int md = (i0 >> 24) & 255;
    if (rd == 0 && ((i1 >> 24) & 1)) {
      tchar* fo = tab + L.fo + lane * 8;
      const int fi = ld<int>(tab, L.fo + 12 * kRow + lane * 4);
      const double mv = S.M[(fi >> 8) & 255];
      const bool usem = mv > 0.0;
      const double conc = usem ? mv : P / rt;
      const double dlconc = usem ? 0.0 : -invT;
      double fc = 0.0, dfc = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const bool inv = (fi >> i) & 1;
        const double gi = ld<double>(fo, (9 + i) * kRow);
        const double x = -(gi * (inv ? invT : T));
        const double ci = ld<double>(fo, (6 + i) * kRow), ex = fexp(x);
        fc = fma(ci, ex, fc);
        dfc = fma(ci * ex, inv ? gi * invT * invT : -gi, dfc);
      }
      const double lnfc = flog(fc);
      const double lfc = lnfc * kInvLn10;
      const double ntmp = fma(-kTroeN1, lfc, kTroeN0);
      const double cco = fma(-kTroeC1, lfc, -kTroeC0);
      const double dlfc = dfc / fc;
      const double bf0 = ld<double>(fo, kRow), ef0 = ld<double>(fo, 2 * kRow);
      const double k0f = arrhenius(ld<double>(fo, 0), bf0, ef0, lt, irt);
      const double dlinf = fma(-ef, irt, bf) * invT;
      kf = troe_d(kf, dlinf, k0f, fma(-ef0, irt, bf0) * invT, conc, dlconc, lnfc, dlfc, ntmp, cco, dkf, dkfc);
      if ((i1 >> 25) & 1) {
        const double bb0 = ld<double>(fo, 4 * kRow), eb0 = ld<double>(fo, 5 * kRow);
        const double k0b = arrhenius(ld<double>(fo, 3 * kRow), bb0, eb0, lt, irt);
        kb = troe_d(kb, fma(-eb, irt, bb) * invT, k0b, fma(-eb0, irt, bb0) * invT, conc, dlconc, lnfc, dlfc, ntmp,
                    cco, dkb, dkbc);
      }
      md = usem ? ((fi >> 8) & 255) : kZero;
    }
    const double c0 = S.C[i0 & 255], c1 = S.C[(i0 >> 8) & 255], c2 = S.C[(i0 >> 16) & 255];
    const double p0 = S.C[i1 & 255], p1 = S.C[(i1 >> 8) & 255], p2 = S.C[(i1 >> 16) & 255];
    const double mm = S.M[(i0 >> 24) & 255];
    const double pf = c0 * c1 * c2, pb = p0 * p1 * p2;
    const double wf = kf * pf, wb = kb * pb;
    S.q[rd * 64 + lane] = (wf - wb) * mm;
    qT[rd] = (dkf * pf - dkb * pb) * mm;
    const double af = kf * mm, ab = kb * mm;
    JacLane& J = jl[rd];
    J.a01 = af * (c0 * c1);
    J.a02 = af * (c0 * c2);
    J.a12 = af * (c1 * c2);
    J.b01 = ab * (p0 * p1);
    J.b02 = ab * (p0 * p2);
    J.b12 = ab * (p1 * p2);
    // third-body reactions (mm = M_r): d/dM of (wf - wb) M is wf - wb; fall-off: d/dconc
    J.g = (rd == 0 && ((i1 >> 24) & 1)) ? dkfc * pf - dkbc * pb : wf - wb;
    J.i0 = i0;
    J.i1 = i1;
    J.md = md;
  }
  wave_sync();
  double cdot = 0.0, dT = 0.0;
  // gather CDOT, then dCDOT/dT from the same schedule
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
#pragma unroll
      for (int rd = 0; rd < L.rounds; ++rd) S.q[rd * 64 + lane] = qT[rd];
      wave_sync();
    }
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < MT::gmax; ++i)
      acc = fma(ld<double>(tab, L.gn + i * kRow + lane * 8), S.q[ld<int>(tab, L.gr + i * kIRow + lane * 4)], acc);
    S.part[lane] = acc;
    wave_sync();
    double v = 0.0;
    if (lane < NEQ - 1) {
      const int f0 = ld<int>(tab, L.sf + lane * 4), nn = ld<int>(tab, L.sn + lane * 4);
      for (int i = 0; i < nn; ++i) v += S.part[f0 + i];
    }
    wave_sync();
    if (pass == 0) cdot = v;
    else dT = v;
  }
  dcdT = dT;
  return cdot;
}

// Column j (species, wave-uniform) of dCDOT/dC: lane k returns
// d CDOT_k / d C_j.  S.C must hold the concentrations of the base pass;
// S.M is overwritten with the efficiencies of species j.
template <int NEQ>
__device__ __forceinline__ double jac_column(tchar* tab, TGR_LDS RhsLds<NEQ>& S,
                                             const JacLane (&jl)[(MechT<NEQ>::nrxn + 63) / 64], int j) {
  using MT = MechT<NEQ>;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  {
    // efficiency of species j in third-body sum `lane`: 1 plus its correction
    double e = 1.0;
#pragma unroll
    for (int q = 0; q < MT::maxe; ++q)
      e += ld<int>(tab, L.ek + q * kIRow + lane * 4) == j ? ld<double>(tab, L.ev + q * kRow + lane * 8) : 0.0;
    S.M[lane] = lane < MT::nthird ? e : 0.0;
  }
  wave_sync();
#pragma unroll
  for (int rd = 0; rd < L.rounds; ++rd) {
    const JacLane& J = jl[rd];
    double d = J.g * S.M[J.md];
    d += ((J.i0 & 255) == j) ? J.a12 : 0.0;
    d += (((J.i0 >> 8) & 255) == j) ? J.a02 : 0.0;
    d += (((J.i0 >> 16) & 255) == j) ? J.a01 : 0.0;
    d -= ((J.i1 & 255) == j) ? J.b12 : 0.0;
    d -= (((J.i1 >> 8) & 255) == j) ? J.b02 : 0.0;
    d -= (((J.i1 >> 16) & 255) == j) ? J.b01 : 0.0;
    S.q[rd * 64 + lane] = d;
  }
  wave_sync();
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < MT::gmax; ++i)
    acc = fma(ld<double>(tab, L.gn + i * kRow + lane * 8), S.q[ld<int>(tab, L.gr + i * kIRow + lane * 4)], acc);
  S.part[lane] = acc;
  wave_sync();
  double v = 0.0;
  if (lane < NEQ - 1) {
    const int f0 = ld<int>(tab, L.sf + lane * 4), nn = ld<int>(tab, L.sn + lane * 4);
    for (int i = 0; i < nn; ++i) v += S.part[f0 + i];
  }
  return v;
}

template <int NEQ, int MODE>
__device__ __attribute__((noinline)) double rhs_call_T(tchar* tab, TGR_LDS RhsLds<NEQ>* S, TGR_LDS RateCache<NEQ>* rc,
                                                       double tmid, double y, double p_gp, double T) {
  return rhs_wave_T<NEQ, MODE>(tab, *S, rc, tmid, y, p_gp, T);
}

// Out-of-line right-hand side for the integrator: one copy of the code
// serves every call site (the solver calls it from six places), which keeps
// the integrator's instruction footprint within the instruction cache.
template <int NEQ>
__device__ __attribute__((noinline)) double rhs_call(tchar* tab, TGR_LDS RhsLds<NEQ>* S, double tmid, double y,
                                                     double p_gp) {
  return rhs_wave<NEQ>(tab, *S, tmid, y, p_gp);
}

}  // namespace tgr

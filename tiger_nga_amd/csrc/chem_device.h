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
//     the net rate of progress q = W_f - W_b into LDS;
//   * CDOT is gathered deterministically: the (reaction, nu) entries of each
//     species are split into lane-sized chunks, summed per lane, then summed
//     per species (fixed order, no atomics).
// The right-hand side then follows finitechem_compute_rhs
// (combustion/finitechem.f90:93-175): rho = sum(rho*Y), T = rhoT/rho,
// C = rho*Y/W, rhs_k = CDOT_k*W_k, rhs_T = -sum(h_k rhs_k)/cp_mix.
//
// Tables live in one device allocation whose layout is a compile-time
// function of the mechanism shape (MechT), so that every table load is a
// single base register plus lane offset plus immediate: nothing per-table has
// to stay live in registers across the solver loop.
#pragma once
#include "wave_target.h"
#include <stdint.h>

#include "fp64_math.h"

namespace tgr {

constexpr int kWave = 64;
constexpr int kMaxNeq = 64;

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

// Byte layout of the table block for a mechanism shape.
//   RX  [rounds*64][8] doubles: A_f b_f e_f A_b b_b e_b | meta, reac, prod, 0 (int32)
//   FO  [64][16] doubles: A0f b0f e0f A0b b0b e0b c0 c1 c2 d0 d1 d2 | nfc, kinds, 0, 0 (int32) | pad
//   TH  [2][13][64] doubles (coefficient-major, species in lanes)
//   W, IW [64] doubles
//   EV  [maxe][64] doubles, EK [maxe][64] int32, EN [64] int32
//   GN  [gmax][64] doubles, GR [gmax][64] int32, GC, SF, SN [64] int32
struct TabLayout {
  int nsp, nrxn, rounds, nfo, nthird, gmax, maxe;
  size_t rx, fo, th, w, iw, ev, ek, en, gn, gr, gc, sf, sn, bytes;
  __host__ __device__ constexpr TabLayout(int nsp_, int nrxn_, int nfo_, int nthird_, int gmax_, int maxe_)
      : nsp(nsp_), nrxn(nrxn_), rounds((nrxn_ + 63) / 64), nfo(nfo_), nthird(nthird_), gmax(gmax_), maxe(maxe_),
        rx(0), fo(rx + (size_t)((nrxn_ + 63) / 64) * 64 * 64), th(fo + 64 * 128), w(th + 2 * 13 * 64 * 8),
        iw(w + 64 * 8), ev(iw + 64 * 8), ek(ev + (size_t)maxe_ * 64 * 8), en(ek + (size_t)maxe_ * 64 * 4),
        gn(en + 64 * 4), gr(gn + (size_t)gmax_ * 64 * 8), gc(gr + (size_t)gmax_ * 64 * 4), sf(gc + 64 * 4),
        sn(sf + 64 * 4), bytes(sn + 64 * 4) {}
};

template <int NEQ>
__host__ __device__ constexpr TabLayout tab_layout() {
  return TabLayout(NEQ - 1, MechT<NEQ>::nrxn, MechT<NEQ>::nfo, MechT<NEQ>::nthird, MechT<NEQ>::gmax,
                   MechT<NEQ>::maxe);
}

// meta word: kind (2 bits) | rev (1) | nr (2) | np (2) | m+1 (7) | fo index (8)
__host__ __device__ inline int rm_kind(int w) { return w & 3; }
__host__ __device__ inline int rm_rev(int w) { return (w >> 2) & 1; }
__host__ __device__ inline int rm_nr(int w) { return (w >> 3) & 3; }
__host__ __device__ inline int rm_np(int w) { return (w >> 5) & 3; }
__host__ __device__ inline int rm_m(int w) { return ((w >> 7) & 127) - 1; }
__host__ __device__ inline int rm_fo(int w) { return (w >> 14) & 255; }

// Device view of a mechanism: the table block plus a few sizes.
struct DevMech {
  const char* tab;
  int nsp;
  double tmid;
};

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
// in lane 63), returned wave-uniform.  Fixed order: deterministic.
__device__ __forceinline__ double wave_sum(double v) {
  double s = v + dpp<0x111, 0xf, 0xf, true>(0.0, v);  // row_shr:1
  s = s + dpp<0x112, 0xf, 0xf, true>(0.0, v);         // row_shr:2
  s = s + dpp<0x113, 0xf, 0xf, true>(0.0, v);         // row_shr:3
  s = s + dpp<0x114, 0xf, 0xe, false>(0.0, s);        // row_shr:4, banks 1-3
  s = s + dpp<0x118, 0xf, 0xc, false>(0.0, s);        // row_shr:8, banks 2-3
  s = s + dpp<0x142, 0xa, 0xf, false>(0.0, s);        // row_bcast:15, rows 1,3
  s = s + dpp<0x143, 0xc, 0xf, false>(0.0, s);        // row_bcast:31, rows 2,3
  return bcast(s, 63);
}

// Maximum over the 64 lanes (same tree; lanes without a source keep their value).
__device__ __forceinline__ double wave_max(double v) {
  double s = fmax(v, dpp<0x111, 0xf, 0xf, false>(v, v));
  s = fmax(s, dpp<0x112, 0xf, 0xf, false>(v, v));
  s = fmax(s, dpp<0x113, 0xf, 0xf, false>(v, v));
  s = fmax(s, dpp<0x114, 0xf, 0xe, false>(s, s));
  s = fmax(s, dpp<0x118, 0xf, 0xc, false>(s, s));
  s = fmax(s, dpp<0x142, 0xa, 0xf, false>(s, s));
  s = fmax(s, dpp<0x143, 0xc, 0xf, false>(s, s));
  return bcast(s, 63);
}

// Returns the table base through an empty asm statement so that table loads
// are issued where they are used: without it the compiler hoists every
// loop-invariant table load out of the solver loops and keeps hundreds of
// values live in registers across the whole integration.
// The result is qualified as device memory so that the loads are global
// (a laundered generic pointer would give flat loads).
template <class T>
__device__ __forceinline__ TGR_GLOBAL T* fresh(T* p) {
  asm volatile("" : TGR_OPAQUE_PTR(p));
  return (TGR_GLOBAL T*)p;
}

using gchar = const TGR_GLOBAL char;

template <class T>
__device__ __forceinline__ T ld(gchar* base, size_t off) {
  return *reinterpret_cast<const TGR_GLOBAL T*>(base + off);
}

__device__ __forceinline__ double arrhenius(double A, double b, double e, double lt, double irt) {
  return A * fexp(fma(b, lt, e * irt));
}

// NASA-7 enthalpy / heat capacity of the species in this lane
// (COMPTHERMODATA, HYDROGENF.f90:514-673); lanes >= nsp return 0.
template <int NEQ>
__device__ __forceinline__ void thermo_lane(gchar* tab, double tmid, double T, int lane, double& h, double& cp) {
  constexpr TabLayout L = tab_layout<NEQ>();
  h = 0.0;
  cp = 0.0;
  if (lane < NEQ - 1) {
    gchar* c = tab + L.th + (T > tmid ? 0 : 13 * 64 * 8) + lane * 8;
    const double a = ld<double>(c, 0), c1 = ld<double>(c, 512), c2 = ld<double>(c, 1024),
                 c3 = ld<double>(c, 1536), c4 = ld<double>(c, 2048), c5 = ld<double>(c, 2560),
                 c6 = ld<double>(c, 3072);
    const double ac = ld<double>(c, 3584), d1 = ld<double>(c, 4096), d2 = ld<double>(c, 4608),
                 d3 = ld<double>(c, 5120), d4 = ld<double>(c, 5632), d5 = ld<double>(c, 6144);
    h = a * (T * (c1 + T * (c2 + T * (c3 + T * (c4 + T * c5)))) + c6);
    cp = ac * (d1 + T * (d2 + T * (d3 + T * (d4 + T * d5))));
  }
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

// Production rates CDOT (kmol/m3/s) of this lane's species from the
// concentrations (lane k holds C_k), temperature and pressure.  ctot is the
// total concentration sum_k C_k (wave-uniform).
template <int NEQ>
__device__ __forceinline__ double prodrates_wave(const DevMech& m, RhsLds<NEQ>& S, double Ck, double ctot, double T,
                                                 double P) {
  using MT = MechT<NEQ>;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  gchar* tab = fresh(m.tab);
  const double lt = flog(T);
  const double rt = kRgasF4 * T;
  const double irt = 1.0 / rt;
  S.C[lane] = Ck;
  wave_sync();
  // third-body sums: M_j = ctot + sum_e (eff_je - 1) C_e
  if (lane < MT::nthird) {
    double s = ctot;
    const int ne = ld<int>(tab, L.en + lane * 4);
#pragma unroll
    for (int e = 0; e < MT::maxe; ++e)
      if (e < ne) s = fma(ld<double>(tab, L.ev + (e * 64 + lane) * 8), S.C[ld<int>(tab, L.ek + (e * 64 + lane) * 4)], s);
    S.M[lane] = s;
  }
  wave_sync();
  // rates of progress, one reaction per lane per round (fall-off in round 0)
#pragma unroll
  for (int rd = 0; rd < L.rounds; ++rd) {
    const int r = rd * 64 + lane;
    if (r < MT::nrxn) {
      gchar* rx = tab + L.rx + (size_t)r * 64;
      const int w = ld<int>(rx, 48), rs = ld<int>(rx, 52), ps = ld<int>(rx, 56);
      const int kind = rm_kind(w), rev = rm_rev(w), nr = rm_nr(w), np = rm_np(w), mi = rm_m(w);
      double kf = arrhenius(ld<double>(rx, 0), ld<double>(rx, 8), ld<double>(rx, 16), lt, irt);
      double kb = rev ? arrhenius(ld<double>(rx, 24), ld<double>(rx, 32), ld<double>(rx, 40), lt, irt) : 0.0;
      if (rd == 0 && kind == 2) {
        gchar* fo = tab + L.fo + (size_t)rm_fo(w) * 128;
        const int nfc = ld<int>(fo, 96), kinds = ld<int>(fo, 100);
        double fc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (i < nfc) {
            const double c = ld<double>(fo, 48 + 8 * i), d = ld<double>(fo, 72 + 8 * i);
            const double x = ((kinds >> i) & 1) ? -(d / T) : -(T / d);
            fc = fma(c, fexp(x), fc);
          }
        }
        const double conc = (mi >= 0 && S.M[mi] > 0.0) ? S.M[mi] : P / rt;
        const double lnfc = flog(fc);
        const double lfc = lnfc * kInvLn10;
        const double ntmp = fma(-kTroeN1, lfc, kTroeN0);
        const double cco = fma(-kTroeC1, lfc, -kTroeC0);
        kf = troe(kf, arrhenius(ld<double>(fo, 0), ld<double>(fo, 8), ld<double>(fo, 16), lt, irt), conc, lnfc,
                  ntmp, cco);
        if (rev)
          kb = troe(kb, arrhenius(ld<double>(fo, 24), ld<double>(fo, 32), ld<double>(fo, 40), lt, irt), conc, lnfc,
                    ntmp, cco);
      }
      double wf = kf, wb = kb;
      if (nr > 0) wf *= S.C[rs & 255];
      if (nr > 1) wf *= S.C[(rs >> 8) & 255];
      if (nr > 2) wf *= S.C[(rs >> 16) & 255];
      if (np > 0) wb *= S.C[ps & 255];
      if (np > 1) wb *= S.C[(ps >> 8) & 255];
      if (np > 2) wb *= S.C[(ps >> 16) & 255];
      double q = wf - wb;
      if (kind == 1) q *= S.M[mi];
      S.q[r] = q;
    }
  }
  wave_sync();
  // deterministic gather of CDOT
  {
    double acc = 0.0;
    const int cnt = ld<int>(tab, L.gc + lane * 4);
#pragma unroll
    for (int i = 0; i < MT::gmax; ++i)
      if (i < cnt) acc = fma(ld<double>(tab, L.gn + (i * 64 + lane) * 8), S.q[ld<int>(tab, L.gr + (i * 64 + lane) * 4)], acc);
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
__device__ __forceinline__ double rhs_wave(const DevMech& m, RhsLds<NEQ>& S, double y, double p_gp) {
  constexpr int nsp = NEQ - 1;
  constexpr TabLayout L = tab_layout<NEQ>();
  const int lane = lane_id();
  const bool sp = lane < nsp;
  gchar* tab = fresh(m.tab);
  const double rho = wave_sum(sp ? y : 0.0);
  const double T = bcast(y, nsp) / rho;
  double h, cp;
  thermo_lane<NEQ>(tab, m.tmid, T, lane, h, cp);
  const double Ck = sp ? y * ld<double>(tab, L.iw + lane * 8) : 0.0;
  const double cpm = wave_sum(sp ? (y / rho) * cp : 0.0);
  const double ctot = wave_sum(Ck);
  const double cdot = prodrates_wave<NEQ>(m, S, Ck, ctot, T, p_gp);
  const double rk = sp ? cdot * ld<double>(tab, L.w + lane * 8) : 0.0;
  const double tmp = wave_sum(-h * rk);
  return sp ? rk : (lane == nsp ? tmp / cpm : 0.0);
}

// Out-of-line right-hand side for the integrator: one copy of the code
// serves every call site (the solver calls it from six places), which keeps
// the integrator's instruction footprint within the instruction cache.  Its
// LDS scratch is allocated here, once per workgroup.
template <int NEQ>
__device__ __attribute__((noinline)) double rhs_call(const char* tab, double tmid, double y, double p_gp) {
  __shared__ RhsLds<NEQ> S;
  // arguments arrive in vector registers; the table base is wave-uniform
  const uint64_t a = (uint64_t)(uintptr_t)tab;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
  const DevMech m{(const char*)(uintptr_t)(((uint64_t)hi << 32) | lo), NEQ - 1, tmid};
  return rhs_wave<NEQ>(m, S, y, p_gp);
}

}  // namespace tgr

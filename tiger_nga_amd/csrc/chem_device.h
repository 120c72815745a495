// Device-side mechanism tables and the wave-per-cell chemistry right-hand side
// for CDNA4 (gfx950, wave64).
//
// One wavefront owns one cell.  Lane i < nsp holds species i (rho*Y_i), lane
// nsp holds rho*T; the remaining lanes hold 0.  The production rates of the
// generated PRODRATES (reference combustion/HYDROGENF.f90:4-277,
// grimech30F.f90:4-2580, DME39F.f90:4-1943) are evaluated table-driven:
//   * lane j < n_third forms the third-body sum M_j from the concentrations
//     broadcast through LDS;
//   * reactions are dealt to lanes round-robin (fall-off reactions first so
//     that only the first round diverges), each lane computing k_f, k_b
//     (Arrhenius / GETLINDRATECOEFF Troe form, HYDROGENF.f90:280-316) and
//     the rates of progress W_f, W_b into LDS;
//   * CDOT is gathered deterministically: the (reaction, nu) entries of each
//     species are split into lane-sized chunks, summed per lane, then summed
//     per species (fixed order, no atomics).
// The right-hand side then follows finitechem_compute_rhs
// (combustion/finitechem.f90:93-175): rho = sum(rho*Y), T = rhoT/rho,
// C = rho*Y/W, rhs_k = CDOT_k*W_k, rhs_T = -sum(h_k rhs_k)/cp_mix.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Reaction metadata word: kind (2 bits) | rev (1) | nr (2) | np (2) | m+1 (7) | fo (8)
__host__ __device__ inline int rm_kind(int w) { return w & 3; }
__host__ __device__ inline int rm_rev(int w) { return (w >> 2) & 1; }
__host__ __device__ inline int rm_nr(int w) { return (w >> 3) & 3; }
__host__ __device__ inline int rm_np(int w) { return (w >> 5) & 3; }
__host__ __device__ inline int rm_m(int w) { return ((w >> 7) & 127) - 1; }
__host__ __device__ inline int rm_fo(int w) { return (w >> 14) & 255; }

// All pointers are device pointers into one allocation owned by the context.
struct DevMech {
  int nsp, nrxn, nfo, nthird;
  int rounds;      // ceil(nrxn / 64)
  int gmax;        // longest CDOT gather chunk
  int smax;        // most chunks of one species
  double tmid;
  const double* W;       // [nsp]
  const double* invW;    // [nsp]
  const double* thermo;  // [2][13][64]: coefficient-major, species in lanes
  const double* effT;    // [nsp][64]: efficiency of species k in sum j (lane j)
  const double* arr;     // [6][nrxn]: A_f b_f e_f A_b b_b e_b (fall-off: k_inf)
  const int* meta;       // [nrxn]
  const int* sp;         // [6][nrxn]: reac0..2 prod0..2
  const double* fo;      // [15][nfo]: A0f b0f e0f A0b b0b e0b c0 c1 c2 d0 d1 d2 kind0 kind1 kind2
  const int* fo_nfc;     // [nfo]
  const int* g_start;    // [64] chunk start per lane
  const int* g_cnt;      // [64] chunk length per lane (0 idle)
  const int* g_r;        // [nnz] reaction of entry
  const double* g_nu;    // [nnz] coefficient of entry
  const int* s_first;    // [64] first chunk lane of species k (lane k)
  const int* s_n;        // [64] number of chunks of species k
};

// Per-wave LDS scratch for one right-hand-side evaluation.
struct RhsLds {
  double C[kWave];
  double M[kWave];
  double part[kWave];
  double Wf[256 + 64];
  double Wb[256 + 64];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Intra-wave ordering point for LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// value of v in lane `src` (src wave-uniform), as a wave-uniform scalar
__device__ __forceinline__ double bcast(double v, int src) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// a value that is equal in all lanes, made visibly uniform to the compiler
__device__ __forceinline__ double uni(double v) {
  int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double arrhenius(double A, double b, double e, double lt, double irt) {
  return A * exp(b * lt + e * irt);
}

// NASA-7 enthalpy / heat capacity of the species in this lane
// (COMPTHERMODATA, HYDROGENF.f90:514-673); lanes >= nsp return 0.
__device__ __forceinline__ void thermo_lane(const DevMech& m, double T, int lane, double& h, double& cp) {
  h = 0.0;
  cp = 0.0;
  if (lane < m.nsp) {
    const double* c = m.thermo + (T > m.tmid ? 0 : 13 * kWave) + lane;
    double a = c[0], c1 = c[kWave], c2 = c[2 * kWave], c3 = c[3 * kWave], c4 = c[4 * kWave], c5 = c[5 * kWave],
           c6 = c[6 * kWave];
    double ac = c[7 * kWave], d1 = c[8 * kWave], d2 = c[9 * kWave], d3 = c[10 * kWave], d4 = c[11 * kWave],
           d5 = c[12 * kWave];
    h = a * (T * (c1 + T * (c2 + T * (c3 + T * (c4 + T * c5)))) + c6);
    cp = ac * (d1 + T * (d2 + T * (d3 + T * (d4 + T * d5))));
  }
}

// Production rates CDOT (kmol/m3/s) of this lane's species from the
// concentrations (lane k holds C_k), temperature and pressure.
__device__ inline double prodrates_wave(const DevMech& m, RhsLds& L, double Ck, double T, double P) {
  const int lane = lane_id();
  const double lt = log(T);
  const double rt = kRgasF4 * T;
  const double irt = 1.0 / rt;
  L.C[lane] = (lane < m.nsp) ? Ck : 0.0;
  wave_sync();
  // third-body sums
  if (lane < m.nthird) {
    double s = 0.0;
    const double* e = m.effT + lane;
    for (int k = 0; k < m.nsp; ++k) s += e[k * kWave] * L.C[k];
    L.M[lane] = s;
  }
  wave_sync();
  // rates of progress, one reaction per lane per round
  for (int rd = 0; rd < m.rounds; ++rd) {
    const int r = rd * kWave + lane;
    if (r < m.nrxn) {
      const int w = m.meta[r];
      const int kind = rm_kind(w), rev = rm_rev(w), nr = rm_nr(w), np = rm_np(w), mi = rm_m(w);
      const int n = m.nrxn;
      double kf = arrhenius(m.arr[r], m.arr[n + r], m.arr[2 * n + r], lt, irt);
      double kb = rev ? arrhenius(m.arr[3 * n + r], m.arr[4 * n + r], m.arr[5 * n + r], lt, irt) : 0.0;
      if (kind == 2) {
        // Troe fall-off, GETLINDRATECOEFF (HYDROGENF.f90:280-316)
        const int f = rm_fo(w), nf = m.nfo;
        const double* fo = m.fo + f;
        const int nfc = m.fo_nfc[f];
        double fc = 0.0;
        for (int i = 0; i < nfc; ++i) {
          double c = fo[(6 + i) * nf], d = fo[(9 + i) * nf];
          double x = (fo[(12 + i) * nf] == 0.0) ? -(T / d) : -(d / T);
          fc = fc + c * exp(x);
        }
        const double conc = (mi >= 0 && L.M[mi] > 0.0) ? L.M[mi] : P / (kRgasF4 * T);
        const double lnfc = log(fc);
        const double lfc = lnfc * kInvLn10;
        const double ntmp = kTroeN0 - kTroeN1 * lfc;
        const double cco = -kTroeC0 - kTroeC1 * lfc;
        {
          double k0 = arrhenius(fo[0], fo[nf], fo[2 * nf], lt, irt);
          double pr = k0 * conc / fmax(kf, 1.0e-60);
          double lg = log(pr) * kInvLn10 + cco;
          double x = lg / (ntmp - kTroeD * lg);
          double F = exp(lnfc / (x * x + 1.0));
          kf = kf * F * pr / (1.0 + pr);
        }
        if (rev) {
          double k0 = arrhenius(fo[3 * nf], fo[4 * nf], fo[5 * nf], lt, irt);
          double pr = k0 * conc / fmax(kb, 1.0e-60);
          double lg = log(pr) * kInvLn10 + cco;
          double x = lg / (ntmp - kTroeD * lg);
          double F = exp(lnfc / (x * x + 1.0));
          kb = kb * F * pr / (1.0 + pr);
        }
      }
      double wf = kf, wb = kb;
      const int* s = m.sp + r;
      if (nr > 0) wf *= L.C[s[0]];
      if (nr > 1) wf *= L.C[s[n]];
      if (nr > 2) wf *= L.C[s[2 * n]];
      if (np > 0) wb *= L.C[s[3 * n]];
      if (np > 1) wb *= L.C[s[4 * n]];
      if (np > 2) wb *= L.C[s[5 * n]];
      if (kind == 1) {
        wf *= L.M[mi];
        wb *= L.M[mi];
      }
      L.Wf[r] = wf;
      L.Wb[r] = wb;
    }
  }
  wave_sync();
  // deterministic gather of CDOT
  {
    double acc = 0.0;
    const int st = m.g_start[lane], cnt = m.g_cnt[lane];
    for (int q = 0; q < cnt; ++q) {
      const int r = m.g_r[st + q];
      acc += m.g_nu[st + q] * (L.Wf[r] - L.Wb[r]);
    }
    L.part[lane] = acc;
  }
  wave_sync();
  double cdot = 0.0;
  if (lane < m.nsp) {
    const int f0 = m.s_first[lane], nn = m.s_n[lane];
    for (int q = 0; q < nn; ++q) cdot += L.part[f0 + q];
  }
  wave_sync();
  return cdot;
}

// finitechem_compute_rhs (finitechem.f90:93-175), constant-pressure branch.
// y: this lane's component of [rho*Y (nsp), rho*T]; returns the same lane of
// the right-hand side.
__device__ inline double rhs_wave(const DevMech& m, RhsLds& L, double y, double p_gp) {
  const int lane = lane_id();
  const int nsp = m.nsp;
  const double rho = uni(wave_sum(lane < nsp ? y : 0.0));
  const double T = bcast(y, nsp) / rho;
  double h, cp;
  thermo_lane(m, T, lane, h, cp);
  const double cpm = uni(wave_sum(lane < nsp ? (y / rho) * cp : 0.0));
  const double Ck = (lane < nsp) ? y * m.invW[lane < nsp ? lane : 0] : 0.0;
  const double cdot = prodrates_wave(m, L, Ck, T, p_gp);
  const double rk = (lane < nsp) ? cdot * m.W[lane < nsp ? lane : 0] : 0.0;
  const double tmp = uni(wave_sum(-h * rk));
  return (lane < nsp) ? rk : (lane == nsp ? tmp / cpm : 0.0);
}

}  // namespace tgr

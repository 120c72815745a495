// CPU build of the wave-per-cell chemistry solver (test infrastructure).
//
// Runs tgr::process_cell -- the same source the HIP step kernel executes --
// with 64 fibers standing in for the lanes of one wavefront, over a
// packed batch of cells (SoA: sc[i * ncell + c], species then rho*T / T as in
// the Fortran block with one cell row).  Used by tests/test_wave_emu.py to
// check the GPU solver logic against the C oracle without a GPU, and with
// sanitizers for bounds checks of the device code.
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "bdf_lane.h"
#include "bdf_wave.h"
#include "lane_rhs_dme39.h"
#include "lane_rhs_ethyred.h"
#include "lane_rhs_gri30.h"
#include "lane_rhs_h2.h"
#include "mech_dme39.h"
#include "mech_ethyred.h"
#include "mech_gri30.h"
#include "mech_h2.h"
#include "tables_host.h"

struct tgr_emu_jac_args {
  long n;
  double p_gp, rtol, atol, h;
  const double* sol;
  double *jdq, *jsplit, *f;
  double *jan, *fan;  // analytic Jacobian and its base evaluation (may be null)
};

namespace {

const tgr_mech_tables* find_mech(const char* name) {
  if (!std::strcmp(name, "h2")) return &h2_tables;
  if (!std::strcmp(name, "gri30")) return &gri30_tables;
  if (!std::strcmp(name, "dme39")) return &dme39_tables;
  if (!std::strcmp(name, "ethyred")) return &ethyred_tables;
  return nullptr;
}

int species_index(const tgr_mech_tables* t, const char* name) {
  for (int k = 0; k < t->nsp; ++k)
    if (!std::strcmp(t->species[k], name)) return k;
  return -1;
}

template <int NEQ>
struct Run {
  const tgr::StepParams* p;
  int* counters;
  tgr::RhsLds<NEQ> rl;  // shared by the lanes, as the wave's LDS is
  // per-lane copy of the LDS history block: every lane reads and writes only
  // its own column, plus one wave-uniform row that all lanes update with the
  // same values -- one update in SIMT, but sequential fibers sharing it would
  // apply it 64 times
  double zn[64][tgr::kQmaxAbs + 2][tgr::kWave];
  std::vector<double> ws;
  tgr_emu::Wave wave;
  int finished = 0;

  static void lane_main(unsigned hi, unsigned lo, int l) {
    Run* r = (Run*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
    for (int64_t c = 0; c < r->p->ncell; ++c) {
      int cnt[TGR_NCOUNT] = {0};
      const int done = r->p->jac ? tgr::process_cell<NEQ, true>(*r->p, r->p->m, r->p->m.tab, &r->rl, c, r->zn[l],
                                                                 r->ws.data(), cnt)
                                 : tgr::process_cell<NEQ, false>(*r->p, r->p->m, r->p->m.tab, &r->rl, c, r->zn[l],
                                                                  r->ws.data(), cnt);
      if (l == 0 && r->counters)
        for (int q = 0; q < TGR_NCOUNT; ++q) r->counters[c * TGR_NCOUNT + q] = done ? cnt[q] : 0;
    }
    // lanes leave in order; the last one returns to the caller
    r->finished++;
    if (r->finished == 64) {
      setcontext(&r->wave.home);
    } else {
      const int nx = (l + 1) & 63;
      tgr_emu::tl_lane = nx;
      setcontext(&r->wave.ctx[nx]);
    }
  }
};

template <int NEQ>
void run(const tgr::StepParams& p, int* counters) {
  Run<NEQ>* r = new Run<NEQ>();
  r->p = &p;
  r->counters = counters;
  r->ws.assign((size_t)NEQ * tgr::kWave, 0.0);
  const size_t stack = 1 << 20;
  std::vector<char> stacks(stack * 64);
  tgr_emu::tw = &r->wave;
  for (int l = 0; l < 64; ++l) {
    getcontext(&r->wave.ctx[l]);
    r->wave.ctx[l].uc_stack.ss_sp = stacks.data() + stack * l;
    r->wave.ctx[l].uc_stack.ss_size = stack;
    r->wave.ctx[l].uc_link = nullptr;
    const uintptr_t a = (uintptr_t)r;
    makecontext(&r->wave.ctx[l], (void (*)())Run<NEQ>::lane_main, 3, (unsigned)(a >> 32), (unsigned)(a & 0xffffffffu), l);
  }
  tgr_emu::tl_lane = 0;
  swapcontext(&r->wave.home, &r->wave.ctx[0]);
  delete r;
}

// The two Jacobians of the one-step solves on packed states: Bdf::dq_jac
// (every column a full difference quotient) and Bdf::jac_split (rate
// constants evaluated once), J[c][row][col].
template <int NEQ>
struct JacRun {
  const tgr_emu_jac_args* a;
  tgr::DevMech m;
  tgr::RhsLds<NEQ> rl;
  tgr::RateCache<NEQ> rc;
  std::vector<double> ws;
  tgr_emu::Wave wave;
  int finished = 0;

  static void lane_main(unsigned hi, unsigned lo, int l) {
    JacRun* r = (JacRun*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
    const tgr_emu_jac_args& a = *r->a;
    for (long c = 0; c < a.n; ++c) {
      const double yv = l < NEQ ? a.sol[c * NEQ + l] : 0.0;
      tgr::Bdf<NEQ> b(r->m, r->m.tab, &r->rl, r->ws.data(), a.p_gp, a.rtol, a.atol);
      b.set_ewt(yv);
      b.h = a.h;
      const double f0 = b.f(yv);
      b.dq_jac(yv, f0);
      if (l < NEQ)
        for (int k = 0; k < NEQ; ++k) a.jdq[(c * NEQ + l) * NEQ + k] = r->ws[k * tgr::kWave + l];
      b.rc = &r->rc;
      const double fs = b.jac_split(yv);
      if (l < NEQ) {
        for (int k = 0; k < NEQ; ++k) a.jsplit[(c * NEQ + l) * NEQ + k] = r->ws[k * tgr::kWave + l];
        a.f[c * NEQ + l] = fs - f0;  // the base evaluations agree
      }
      if (a.jan) {
        tgr::Bdf<NEQ, true> ba(r->m, r->m.tab, &r->rl, r->ws.data(), a.p_gp, a.rtol, a.atol);
        const double fa = ba.jac_analytic(yv);
        if (l < NEQ) {
          for (int k = 0; k < NEQ; ++k) a.jan[(c * NEQ + l) * NEQ + k] = r->ws[k * tgr::kWave + l];
          a.fan[c * NEQ + l] = fa;
        }
      }
    }
    r->finished++;
    if (r->finished == 64) {
      setcontext(&r->wave.home);
    } else {
      const int nx = (l + 1) & 63;
      tgr_emu::tl_lane = nx;
      setcontext(&r->wave.ctx[nx]);
    }
  }
};

template <int NEQ>
void run_jac(const tgr_emu_jac_args& a, const tgr::DevMech& m) {
  JacRun<NEQ>* r = new JacRun<NEQ>();
  r->a = &a;
  r->m = m;
  r->ws.assign((size_t)NEQ * tgr::kWave, 0.0);
  const size_t stack = 1 << 20;
  std::vector<char> stacks(stack * 64);
  tgr_emu::tw = &r->wave;
  for (int l = 0; l < 64; ++l) {
    getcontext(&r->wave.ctx[l]);
    r->wave.ctx[l].uc_stack.ss_sp = stacks.data() + stack * l;
    r->wave.ctx[l].uc_stack.ss_size = stack;
    r->wave.ctx[l].uc_link = nullptr;
    const uintptr_t u = (uintptr_t)r;
    makecontext(&r->wave.ctx[l], (void (*)())JacRun<NEQ>::lane_main, 3, (unsigned)(u >> 32),
                (unsigned)(u & 0xffffffffu), l);
  }
  tgr_emu::tl_lane = 0;
  swapcontext(&r->wave.home, &r->wave.ctx[0]);
  delete r;
}

}  // namespace

// The Jacobians (dq_jac, jac_split and, with jan, jac_analytic) on n packed states sol[n][neq]
// ([rho Y_k, rho T]) with step size h for the increments: jdq, jsplit
// [n][neq][neq] (row-major), f[n][neq] the difference of the two base
// right-hand sides (zero when the fixed-temperature evaluation reproduces
// the plain one).
extern "C" int emu_jac(const char* mech, long n, double p_gp, double rtol, double atol, double h, const double* sol,
                       double* jdq, double* jsplit, double* f, double* jan, double* fan) {
  const tgr_mech_tables* t = find_mech(mech);
  if (!t) return -1;
  std::vector<char> blob;
  tgr_mech_info info;
  std::string err;
  if (tgr::build_table_blob(t, blob, &info, &err)) return -2;
  tgr::DevMech m;
  m.tab = blob.data();
  m.nsp = t->nsp;
  m.tmid = t->tmid;
  tgr_emu_jac_args a{n, p_gp, rtol, atol, h, sol, jdq, jsplit, f, jan, fan};
  switch (t->nsp + 1) {
    case 10: run_jac<10>(a, m); break;
    case 37: run_jac<37>(a, m); break;
    case 40: run_jac<40>(a, m); break;
    case 46: run_jac<46>(a, m); break;
    default: return -3;
  }
  return 0;
}

// One chemistry step over ncell packed cells: rho[ncell], sc[(nsp+1) * ncell]
// (mass fractions then T, overwritten), hr[ncell], chemsrc[3 * ncell],
// counters[ncell * 10] (nst nfe nfe_dq nje nsetups nni netf ncfn nlu status).
extern "C" int emu_step(const char* mech, long ncell, double* rho, double* sc, double dt, double p_gp, double rtol,
                        double atol, int max_steps, int* counters, double* hr, double* chemsrc, int jac) {
  const tgr_mech_tables* t = find_mech(mech);
  if (!t) return -1;
  std::vector<char> blob;
  tgr_mech_info info;
  std::string err;
  if (tgr::build_table_blob(t, blob, &info, &err)) return -2;
  tgr::StepParams p{};
  p.m.tab = blob.data();
  p.m.nsp = t->nsp;
  p.m.tmid = t->tmid;
  p.dt = dt;
  p.p_gp = p_gp;
  p.rtol = rtol;
  p.atol = atol;
  p.max_steps = max_steps;
  p.qmax = 5;
  p.jac = jac;
  p.nx = (int)ncell;
  p.ny = p.nz = 1;
  p.sx = 1;
  p.sy = p.sz = ncell;
  p.base = 0;
  p.plane = ncell;
  p.isc_sc0 = 0;
  p.isc_T0 = t->nsp;
  p.rho = rho;
  p.sc = sc;
  p.hr = hr;
  p.chemsrc = chemsrc;
  p.idx_h2o = species_index(t, "H2O");
  p.idx_co = species_index(t, "CO");
  p.idx_co2 = species_index(t, "CO2");
  p.ncell = ncell;
  switch (t->nsp + 1) {
    case 10: run<10>(p, counters); break;
    case 37: run<37>(p, counters); break;
    case 40: run<40>(p, counters); break;
    case 46: run<46>(p, counters); break;
    default: return -3;
  }
  return 0;
}

// The generated lane-per-state right-hand side on n packed states
// sol[n][neq] -> rhs[n][neq] (plain C++, one state at a time).
extern "C" int emu_lane_rhs(const char* mech, long n, double p_gp, const double* sol, double* rhs) {
  const tgr_mech_tables* t = find_mech(mech);
  if (!t) return -1;
  auto run_n = [&](auto tag) {
    constexpr int NEQ = decltype(tag)::value;
    for (long c = 0; c < n; ++c) {
      double s[NEQ], f[NEQ];
      for (int k = 0; k < NEQ; ++k) s[k] = sol[c * NEQ + k];
      tgr::LaneRhs<NEQ>::eval(s, p_gp, f);
      for (int k = 0; k < NEQ; ++k) rhs[c * NEQ + k] = f[k];
    }
  };
  switch (t->nsp + 1) {
    case 10: run_n(std::integral_constant<int, 10>()); break;
    case 37: run_n(std::integral_constant<int, 37>()); break;
    case 40: run_n(std::integral_constant<int, 40>()); break;
    case 46: run_n(std::integral_constant<int, 46>()); break;
    default: return -3;
  }
  return 0;
}

// The thread-per-cell solver (LaneBdf, bdf_lane.h) on ncell packed cells, one
// cell after another on this host thread (it has no cross-lane operations):
// sol[ncell][neq] integrated in place, counters[ncell * 10] as emu_step.
extern "C" int emu_lane_step(const char* mech, long ncell, double* sol, double dt, double p_gp, double rtol,
                             double atol, int max_steps, int* counters) {
  const tgr_mech_tables* t = find_mech(mech);
  if (!t || t->nsp + 1 != 10) return -1;
  constexpr int NEQ = 10;
  std::vector<double> ws((size_t)NEQ * NEQ * tgr::kWave), ml((size_t)NEQ * NEQ * tgr::kWave);
  for (long c = 0; c < ncell; ++c) {
    tgr::LaneBdf<NEQ> b;
    b.p_gp = p_gp;
    b.rtol = rtol;
    b.atol = atol;
    b.qmax = tgr::kQmaxAbs;
    b.max_steps = max_steps;
    b.jw = ws.data();
    b.Ml = ml.data();
    double s[NEQ];
    for (int k = 0; k < NEQ; ++k) s[k] = sol[c * NEQ + k];
    const int st = b.integrate(dt, s);
    for (int k = 0; k < NEQ; ++k) sol[c * NEQ + k] = s[k];
    const int cnt[TGR_NCOUNT] = {b.nst, b.c_nfe, b.c_nfe_dq, b.c_nje, b.c_nsetups, b.c_nni, b.c_netf, b.c_ncfn,
                                 b.c_nlu, st};
    for (int q = 0; q < TGR_NCOUNT; ++q) counters[c * TGR_NCOUNT + q] = cnt[q];
  }
  return 0;
}

namespace {

// Dense solves of the wave solver on n matrices a[n][NEQ][NEQ] (row-major)
// with right-hand sides b[n][NEQ]: Bdf::getrf + getrs into xlu, Bdf::getri +
// getri_solve into xinv; info[n][2] the two factorisations' return codes.
template <int NEQ>
struct LinRun {
  long n;
  const double *a, *b;
  double *xlu, *xinv;
  int* info;
  tgr::DevMech m{};
  tgr::RhsLds<NEQ> rl;
  std::vector<double> ws;
  tgr_emu::Wave wave;
  int finished = 0;

  static void lane_main(unsigned hi, unsigned lo, int l) {
    LinRun* r = (LinRun*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
    for (long c = 0; c < r->n; ++c) {
      tgr::Bdf<NEQ> s(r->m, r->m.tab, &r->rl, r->ws.data(), 1.0e5, 1.0e-6, 1.0e-12);
      const double bv = l < NEQ ? r->b[c * NEQ + l] : 0.0;
      for (int k = 0; k < NEQ; ++k) s.M[k] = l < NEQ ? r->a[(c * NEQ + l) * NEQ + k] : 0.0;
      const int i1 = s.getrf();
      const double x1 = i1 ? 0.0 : s.getrs(bv);
      for (int k = 0; k < NEQ; ++k) s.M[k] = l < NEQ ? r->a[(c * NEQ + l) * NEQ + k] : 0.0;
      const int i2 = s.getri();
      // two right-hand sides on one inverse, as the Newton iterations use it
      const double x2 = i2 ? 0.0 : s.getri_solve(bv);
      const double x3 = i2 ? 0.0 : s.getri_solve(bv);
      if (l < NEQ) {
        r->xlu[c * NEQ + l] = x1;
        r->xinv[c * NEQ + l] = (x2 == x3) ? x2 : NAN;
      }
      if (l == 0) {
        r->info[2 * c] = i1;
        r->info[2 * c + 1] = i2;
      }
    }
    r->finished++;
    if (r->finished == 64) {
      setcontext(&r->wave.home);
    } else {
      const int nx = (l + 1) & 63;
      tgr_emu::tl_lane = nx;
      setcontext(&r->wave.ctx[nx]);
    }
  }
};

template <int NEQ>
void run_lin(long n, const double* a, const double* b, double* xlu, double* xinv, int* info) {
  LinRun<NEQ>* r = new LinRun<NEQ>();
  r->n = n;
  r->a = a;
  r->b = b;
  r->xlu = xlu;
  r->xinv = xinv;
  r->info = info;
  r->ws.assign((size_t)NEQ * tgr::kWave, 0.0);
  const size_t stack = 1 << 20;
  std::vector<char> stacks(stack * 64);
  tgr_emu::tw = &r->wave;
  for (int l = 0; l < 64; ++l) {
    getcontext(&r->wave.ctx[l]);
    r->wave.ctx[l].uc_stack.ss_sp = stacks.data() + stack * l;
    r->wave.ctx[l].uc_stack.ss_size = stack;
    r->wave.ctx[l].uc_link = nullptr;
    const uintptr_t u = (uintptr_t)r;
    makecontext(&r->wave.ctx[l], (void (*)())LinRun<NEQ>::lane_main, 3, (unsigned)(u >> 32),
                (unsigned)(u & 0xffffffffu), l);
  }
  tgr_emu::tl_lane = 0;
  swapcontext(&r->wave.home, &r->wave.ctx[0]);
  delete r;
}

}  // namespace

// The wave solver's two dense solves (LU and Gauss-Jordan inverse) for neq
// 10, 37, 40 or 46 (the shipped mechanisms' sizes).
extern "C" int emu_linsolve(int neq, long n, const double* a, const double* b, double* xlu, double* xinv,
                            int* info) {
  switch (neq) {
    case 10: run_lin<10>(n, a, b, xlu, xinv, info); return 0;
    case 37: run_lin<37>(n, a, b, xlu, xinv, info); return 0;
    case 40: run_lin<40>(n, a, b, xlu, xinv, info); return 0;
    case 46: run_lin<46>(n, a, b, xlu, xinv, info); return 0;
    default: return -1;
  }
}

// C-ABI layer of the chemistry step: context, device tables, launches.
// Entry points are declared in include/tiger_chem.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tiger_chem.h"
#include "../../include/tiger_mlp.h"
#include "chem_kernels.h"
#include "tables_host.h"
#include "mech_dme39.h"
#include "mech_ethyred.h"
#include "mech_gri30.h"
#include "mech_h2.h"

using tgr::DevMech;
using tgr::StepParams;
using tgr::SurrParams;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(TGR_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

const tgr_mech_tables* find_mech(const char* name) {
  if (!name) return nullptr;
  if (!strcmp(name, "h2")) return &h2_tables;
  if (!strcmp(name, "gri30")) return &gri30_tables;
  if (!strcmp(name, "dme39")) return &dme39_tables;
  if (!strcmp(name, "ethyred")) return &ethyred_tables;
  return nullptr;
}

// Device buffer that grows on demand.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(TGR_E_NOMEM, "hipMalloc failed");
    cap = bytes;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace

struct tgr_chem {
  const tgr_mech_tables* t = nullptr;
  int nsp = 0, neq = 0;
  double rtol = 1e-6, atol = 1e-12;
  int max_steps = 10000, qmax = 5;
  int jac = 0;  // 0: difference-quotient Jacobian (the reference's CVODE FCVDENSE); 1: analytic
  int step_mapping = 0;  // 0: a wave per cell; 1: a thread per cell (small mechanisms, bdf_lane.h)
  int step_waves = 2;    // wave per cell: resident waves per SIMD (2, or 4 where the kernel supports it)
  int rhs_lanes = 1;  // batched right-hand side: one state per lane (1) or per wave (0)
  int64_t grid_override = 0;
  int device = 0;
  int n_cu = 0;
  tgr_mech_info info{};
  DevMech dm{};
  void* d_tables = nullptr;
  void* d_props = nullptr;  // transport table (properties kernel)
  unsigned int* d_queue = nullptr;
  unsigned long long* d_totals = nullptr;
  DevBuf b_ws;  // per-block solver workspace
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // staging for the host-array entry point
  DevBuf b_rho, b_sc, b_mask, b_hr, b_chem;
  // surrogate-guided step: network (borrowed), feature / output scaling,
  // network inputs and outputs, rejected-cell list and counters
  tgr_mlp* net = nullptr;
  tgr_mlp* own_net = nullptr;  // a network loaded by tgr_chem_load_surrogate (destroyed with the context)
  int net_in_pad = 0, net_out_pad = 0;
  double* d_elem = nullptr;  // [TGR_NELEM][64] element atoms / W (surrogate projection)
  float* d_in_shift = nullptr;
  float* d_in_scale = nullptr;
  double* d_out_scale = nullptr;
  std::vector<float> h_in_shift, h_in_scale;
  DevBuf b_x, b_guess, b_rej;
  DevBuf b_diff;  // diffusion workspace: W_gp, cp_mix, corr_vel x / y / z (5 planes)
  unsigned int* d_rej_count = nullptr;
  unsigned long long* d_stotals = nullptr;
  int surr_nsub = 8, surr_guess = 0, surr_newton_max = 8;
  // adaptive use: while the accepted fraction of the last guided call is
  // below surr_min_accept, calls go to the plain step, with a guided probe
  // every surr_probe calls (0: always guided)
  int surr_probe = 8, surr_skipped = 0;
  int surr_adaptive = 1, surr_max_att = 48, surr_split_jac = 1;
  int props_lane = 1;  // properties kernel: 1 thread per cell, 0 wave per cell
  int surr_mode = 0;        // 0: guided solve; 1: standalone surrogate with an FP64 residual check
  double surr_sa_tol = 1.0;  // standalone: largest Newton correction accepted, in acceptance tolerances
  double surr_h0_fac = 0.0, surr_grow = 10.0, surr_keep = 1.5, surr_jac_theta = 0.3;
  int surr_max_jac = 4;
  double surr_min_accept = 0.15, surr_last_accept = 1.0;
  double surr_newton_tol = 0.03, surr_tol_y = 2.5e-7, surr_tol_t = 1.0e-2, surr_grade = 2.0;
  hipEvent_t ev2 = nullptr;
};

namespace {

// Builds the device table block on the host and uploads it.
int build_tables(tgr_chem* c) {
  std::vector<char> blob;
  std::string err;
  if (tgr::build_table_blob(c->t, blob, &c->info, &err)) return fail(TGR_E_MECH, err);
  HIP_TRY(hipMalloc(&c->d_tables, blob.size()));
  HIP_TRY(hipMemcpy(c->d_tables, blob.data(), blob.size(), hipMemcpyHostToDevice));
  c->dm.tab = (const char*)c->d_tables;
  c->dm.nsp = c->t->nsp;
  c->dm.tmid = c->t->tmid;
  // transport table: mucoeff, kOverEps, W (64 lanes each), then the Wilke
  // pair factors A(k,j) = (1/sqrt 8) (1 + W_k/W_j)^-1/2 and
  // B(k,j) = (W_j/W_k)^1/4 at [j][k] (finitechem.f90:376-380); then, for the
  // thread-per-cell kernel, (sqrt A(k,j), sqrt A(k,j) B(k,j)) at [k][j], zero
  // for j = k (the reference's sum skips it)
  const int n = c->t->nsp;
  const size_t pair = 192 + 2 * (size_t)n * 64;
  std::vector<double> pt(pair + 2 * (size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) {
    pt[k] = c->t->mucoeff[k];
    pt[64 + k] = c->t->koveps[k];
    pt[128 + k] = c->t->W[k];
  }
  for (int k = n; k < 64; ++k) pt[128 + k] = 1.0;
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      pt[192 + (size_t)j * 64 + k] = (1.0 / std::sqrt(8.0)) * std::pow(1.0 + (c->t->W[k] / c->t->W[j]), -0.5);
      pt[192 + (size_t)n * 64 + (size_t)j * 64 + k] = std::pow(c->t->W[j] / c->t->W[k], 0.25);
      if (j != k) {
        const double sa = std::sqrt(pt[192 + (size_t)j * 64 + k]);
        pt[pair + 2 * ((size_t)k * n + j)] = sa;
        pt[pair + 2 * ((size_t)k * n + j) + 1] = sa * pt[192 + (size_t)n * 64 + (size_t)j * 64 + k];
      }
    }
  HIP_TRY(hipMalloc(&c->d_props, pt.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(c->d_props, pt.data(), pt.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> el((size_t)TGR_NELEM * 64, 0.0);
  for (int e = 0; e < TGR_NELEM; ++e)
    for (int k = 0; k < n; ++k) el[(size_t)e * 64 + k] = c->t->elem[(size_t)e * n + k] / c->t->W[k];
  HIP_TRY(hipMalloc(&c->d_elem, el.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(c->d_elem, el.data(), el.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int species_index(const tgr_mech_tables* t, const char* name) {
  for (int k = 0; k < t->nsp; ++k)
    if (!strcmp(t->species[k], name)) return k;
  return -1;
}

int check_grid(const tgr_grid* g, int nsp) {
  if (!g) return fail(TGR_E_ARG, "grid descriptor is NULL");
  if (g->imin_ < g->imino_ || g->imax_ > g->imaxo_ || g->jmin_ < g->jmino_ || g->jmax_ > g->jmaxo_ ||
      g->kmin_ < g->kmino_ || g->kmax_ > g->kmaxo_)
    return fail(TGR_E_ARG, "interior extents exceed array extents");
  if (g->isc_sc < 1 || g->isc_sc + nsp - 1 > g->nscalar || g->isc_T < 1 || g->isc_T > g->nscalar)
    return fail(TGR_E_ARG, "scalar indices out of range");
  if (g->isc_T >= g->isc_sc && g->isc_T <= g->isc_sc + nsp - 1)
    return fail(TGR_E_ARG, "isc_T overlaps the species block");
  return 0;
}

// step parameters of a block (no queue / workspace yet)
// waves per SIMD of the wave-per-cell step kernel in force: the four-wave
// build has the difference-quotient Jacobian only
int step_waves_eff(const tgr_chem* c) { return c->jac ? 2 : c->step_waves; }

StepParams step_params(const tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g,
                       double* rho, double* sc, int isc_sc0, int isc_T0, const int* mask, double* hr, double* chem) {
  const int nx = g->imax_ - g->imin_ + 1, ny = g->jmax_ - g->jmin_ + 1, nz = g->kmax_ - g->kmin_ + 1;
  const int64_t nxo = g->imaxo_ - g->imino_ + 1, nyo = g->jmaxo_ - g->jmino_ + 1, nzo = g->kmaxo_ - g->kmino_ + 1;
  StepParams p{};
  p.m = c->dm;
  p.dt = dt;
  p.p_gp = 0.5 * (p_thermo + p_thermo_old);
  p.rtol = c->rtol;
  p.atol = c->atol;
  p.max_steps = c->max_steps;
  p.qmax = c->qmax;
  p.jac = c->jac;
  p.waves = step_waves_eff(c);
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.sx = 1;
  p.sy = nxo;
  p.sz = nxo * nyo;
  p.base = (int64_t)(g->imin_ - g->imino_) + (int64_t)(g->jmin_ - g->jmino_) * nxo +
           (int64_t)(g->kmin_ - g->kmino_) * nxo * nyo;
  p.plane = nxo * nyo * nzo;
  p.isc_sc0 = isc_sc0;
  p.isc_T0 = isc_T0;
  p.mask = mask;
  p.mask_ld = g->imaxo - g->imino + 1;
  p.mask_base = (int64_t)(g->imin_ - g->imino) + (int64_t)(g->jmin_ - g->jmino) * p.mask_ld;
  p.rho = rho;
  p.sc = sc;
  p.hr = hr;
  p.chemsrc = chem;
  p.idx_h2o = species_index(c->t, "H2O");
  p.idx_co = species_index(c->t, "CO");
  p.idx_co2 = species_index(c->t, "CO2");
  p.ncell = (int64_t)nx * ny * nz;
  p.totals = c->d_totals;
  p.cell_stats = nullptr;
  return p;
}

// active (mask /= 1) interior cells of a block
int count_cells(const tgr_grid* g, const int* mask, int64_t mask_ld, int64_t ncell, int64_t* cells) {
  if (!mask) {
    *cells = ncell;
    return 0;
  }
  std::vector<int> hm((size_t)(g->imaxo - g->imino + 1) * (g->jmaxo - g->jmino + 1));
  HIP_TRY(hipMemcpy(hm.data(), mask, hm.size() * sizeof(int), hipMemcpyDefault));
  int64_t n = 0;
  for (int j = g->jmin_; j <= g->jmax_; ++j)
    for (int i = g->imin_; i <= g->imax_; ++i) n += hm[(size_t)(j - g->jmino) * mask_ld + (i - g->imino)] != 1;
  *cells = n * (g->kmax_ - g->kmin_ + 1);
  return 0;
}

// persistent grid of a solver kernel: every wave of every resident workgroup pulls cells
int persistent_grid(const tgr_chem* c, int occupancy, int64_t ncell, int wpb) {
  const int per_cu = (c->grid_override <= 0 && occupancy > 0) ? occupancy : 1;
  const int64_t gs = c->grid_override > 0 ? c->grid_override : (int64_t)c->n_cu * per_cu;
  return (int)std::max<int64_t>(1, std::min<int64_t>(gs, (ncell + wpb - 1) / wpb));
}

// Launch shape of the plain step over ncell cells: the grid, whether the
// thread-per-cell kernel runs, and the solver workspace it needs (bytes).  A
// thread per cell: one 64-thread workgroup per wave slot, each thread taking
// cells one by one (persistent grid of the resident capacity); a wave per
// cell: the persistent grid of the kernel's occupancy, neq x 64 doubles per wave.
size_t step_shape(const tgr_chem* c, int64_t ncell, int* grid, bool* lanes) {
  *lanes = c->step_mapping == 1 && tgr_step_lane_supported(c->neq) && !c->jac;
  if (*lanes) {
    *grid = (int)std::max<int64_t>(
        1, std::min<int64_t>(c->grid_override > 0 ? c->grid_override
                                                   : (int64_t)c->n_cu * std::max(1, tgr_step_lane_occupancy(c->neq)),
                             (ncell + 63) / 64));
    return (size_t)*grid * tgr_step_lane_ws(c->neq) * sizeof(double);
  }
  *grid = persistent_grid(c, tgr_step_occupancy_w(c->neq, step_waves_eff(c)), ncell, tgr_step_block_waves());
  return (size_t)*grid * tgr_step_block_waves() * c->neq * 64 * sizeof(double);
}

// Layout of an external scratch block (tgr_chem_source_chemistry_device_noalloc):
// the work counter at byte 0, the counter totals at byte 64, the per-wave
// Jacobian workspace from byte 256.
constexpr size_t kScrTotals = 64, kScrWs = 256;

// the plain step; with list / nlist only the listed cells (device list and
// count); with scratch (and stats == NULL) the counters and workspace come from
// that block of scratch_bytes instead of the context's buffers
int run_step(tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g, double* rho,
             double* sc, int isc_sc0, int isc_T0, const int* mask, double* hr, double* chem, hipStream_t s,
             tgr_stats* stats, const int* list = nullptr, const unsigned int* nlist = nullptr,
             void* scratch = nullptr, size_t scratch_bytes = 0) {
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(TGR_E_ARG, "dt must be positive and finite");
  const int nx = g->imax_ - g->imin_ + 1, ny = g->jmax_ - g->jmin_ + 1, nz = g->kmax_ - g->kmin_ + 1;
  if (stats) memset(stats, 0, sizeof(*stats));
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  StepParams p = step_params(c, dt, p_thermo, p_thermo_old, g, rho, sc, isc_sc0, isc_T0, mask, hr, chem);
  p.cell_list = list;
  p.nlist = nlist;
  bool lanes = false;
  int grid = 0;
  const size_t ws_bytes = step_shape(c, p.ncell, &grid, &lanes);
  if (scratch) {
    if (ws_bytes > scratch_bytes - kScrWs)
      return fail(TGR_E_ARG, "the step's workspace (" + std::to_string(ws_bytes) + " bytes for a grid of " +
                                 std::to_string(grid) + " workgroups) exceeds the " +
                                 std::to_string(scratch_bytes - kScrWs) +
                                 "-byte device scratch of the non-allocating entry point; set a smaller \"grid\"");
    p.queue = (unsigned int*)scratch;
    p.totals = (unsigned long long*)((char*)scratch + kScrTotals);
    p.ws = (double*)((char*)scratch + kScrWs);
  } else {
    int rc0 = c->b_ws.ensure(ws_bytes);
    if (rc0) return rc0;
    p.queue = c->d_queue;
    p.ws = (double*)c->b_ws.p;
  }
  // with stats == NULL the call enqueues two kernels and nothing else (no
  // memset, no event), so that it can be captured in a HIP graph
  if (stats) HIP_TRY(hipEventRecord(c->ev0, s));
  int rc = lanes ? tgr_launch_step_lane(&p, c->neq, grid, s) : tgr_launch_step(&p, c->neq, grid, s);
  if (rc) return fail(TGR_E_HIP, "step kernel launch failed");
  if (stats) HIP_TRY(hipEventRecord(c->ev1, s));
  if (stats) {
    unsigned long long tot[TGR_NCOUNT];
    HIP_TRY(hipMemcpyAsync(tot, c->d_totals, sizeof(tot), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    int64_t cells = 0;
    // active cells, recomputed on the host from the mask
    if ((rc = count_cells(g, mask, p.mask_ld, p.ncell, &cells))) return rc;
    stats->cells = cells;
    stats->nst = tot[0];
    stats->nfe = tot[1];
    stats->nfe_dq = tot[2];
    stats->nje = tot[3];
    stats->nsetups = tot[4];
    stats->nni = tot[5];
    stats->netf = tot[6];
    stats->ncfn = tot[7];
    stats->nlu = tot[8];
    stats->failed = tot[9];
    stats->kernel_ms = ms;
  }
  return 0;
}

}  // namespace

extern "C" {

const char* tgr_chem_last_error(void) { return g_err.c_str(); }

int tgr_chem_create(const char* mech, double rtol, double atol, int max_steps, tgr_chem** out) {
  if (!out) return fail(TGR_E_ARG, "out is NULL");
  *out = nullptr;
  const tgr_mech_tables* t = find_mech(mech);
  if (!t) return fail(TGR_E_MECH, std::string("unknown mechanism: ") + (mech ? mech : "(null)"));
  if (!(rtol > 0.0) || !(atol > 0.0)) return fail(TGR_E_ARG, "tolerances must be positive");
  const int neq = t->nsp + 1;
  if (neq != 10 && neq != 37 && neq != 40 && neq != 46) return fail(TGR_E_MECH, "no kernel instance for this size");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(TGR_E_HIP, "no HIP device");
  tgr_chem* c = new tgr_chem();
  c->t = t;
  c->nsp = t->nsp;
  c->neq = neq;
  c->rtol = rtol;
  c->atol = atol;
  c->max_steps = max_steps > 0 ? max_steps : 10000;
  int rc = 0;
  if (hipGetDevice(&c->device) != hipSuccess) rc = fail(TGR_E_HIP, "hipGetDevice failed");
  hipDeviceProp_t prop;
  if (!rc && hipGetDeviceProperties(&prop, c->device) != hipSuccess) rc = fail(TGR_E_HIP, "device query failed");
  if (!rc) c->n_cu = prop.multiProcessorCount;
  if (!rc) rc = build_tables(c);
  if (!rc && hipMalloc((void**)&c->d_queue, 256) != hipSuccess) rc = fail(TGR_E_NOMEM, "queue alloc");
  if (!rc && hipMalloc((void**)&c->d_totals, 256) != hipSuccess) rc = fail(TGR_E_NOMEM, "totals alloc");
  if (!rc && (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess))
    rc = fail(TGR_E_HIP, "event create");
  // device address of this GPU's scratch for the non-allocating entry point
  // (resolved here, with the context's device current, once per device)
  void* scr = nullptr;
  size_t scr_bytes = 0;
  if (!rc && tgr_scratch_resolve(c->device, &scr, &scr_bytes) != 0)
    rc = fail(TGR_E_HIP, "hipGetSymbolAddress of the device scratch failed");
  if (rc) {
    tgr_chem_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int tgr_chem_destroy(tgr_chem* c) {
  if (!c) return 0;
  if (c->d_tables) (void)hipFree(c->d_tables);
  if (c->d_props) (void)hipFree(c->d_props);
  if (c->d_elem) (void)hipFree(c->d_elem);
  if (c->d_queue) (void)hipFree(c->d_queue);
  if (c->d_totals) (void)hipFree(c->d_totals);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev2) (void)hipEventDestroy(c->ev2);
  for (void* q : {(void*)c->d_in_shift, (void*)c->d_in_scale, (void*)c->d_out_scale, (void*)c->d_rej_count,
                  (void*)c->d_stotals})
    if (q) (void)hipFree(q);
  if (c->own_net) tgr_mlp_destroy(c->own_net);
  c->b_x.release();
  c->b_guess.release();
  c->b_rej.release();
  c->b_diff.release();
  c->b_ws.release();
  c->b_rho.release();
  c->b_sc.release();
  c->b_mask.release();
  c->b_hr.release();
  c->b_chem.release();
  delete c;
  return 0;
}

int tgr_chem_mech_info(const tgr_chem* c, tgr_mech_info* info) {
  if (!c || !info) return fail(TGR_E_ARG, "NULL argument");
  *info = c->info;
  return 0;
}

int tgr_chem_species_name(const tgr_chem* c, int k, char* buf, int len) {
  if (!c || !buf || len <= 0 || k < 0 || k >= c->nsp) return fail(TGR_E_ARG, "bad species query");
  snprintf(buf, (size_t)len, "%s", c->t->species[k]);
  return 0;
}

int tgr_chem_molar_mass(const tgr_chem* c, double* W) {
  if (!c || !W) return fail(TGR_E_ARG, "NULL argument");
  memcpy(W, c->t->W, sizeof(double) * c->nsp);
  return 0;
}

int tgr_chem_set_option(tgr_chem* c, const char* key, int64_t v) {
  if (!c || !key) return fail(TGR_E_ARG, "NULL argument");
  if (!strcmp(key, "grid")) c->grid_override = v;
  else if (!strcmp(key, "rhs_mapping")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "rhs_mapping must be 0 (wave per state) or 1 (lane per state)");
    c->rhs_lanes = (int)v;
  } else if (!strcmp(key, "step_waves")) {
    if (!tgr_step_waves_supported(c->neq, (int)v))
      return fail(TGR_E_ARG, "step_waves must be 2, or 4 for a 10-equation mechanism (H2)");
    c->step_waves = (int)v;
  } else if (!strcmp(key, "step_mapping")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "step_mapping must be 0 (wave per cell) or 1 (thread per cell)");
    c->step_mapping = (int)v;
  } else if (!strcmp(key, "props_mapping")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "props_mapping must be 0 (wave per cell) or 1 (thread per cell)");
    c->props_lane = (int)v;
  } else if (!strcmp(key, "jacobian")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "jacobian must be 0 (difference quotients) or 1 (analytic)");
    c->jac = (int)v;
  } else if (!strcmp(key, "max_order")) {
    if (v < 1 || v > 5) return fail(TGR_E_ARG, "max_order must be 1..5");
    c->qmax = (int)v;
  } else if (!strcmp(key, "max_steps")) {
    if (v < 1) return fail(TGR_E_ARG, "max_steps must be positive");
    c->max_steps = (int)v;
  } else if (!strcmp(key, "surr_nsub")) {
    if (v < 1 || v > tgr::kSurrMaxSub) return fail(TGR_E_ARG, "surr_nsub must be 1..8");
    c->surr_nsub = (int)v;
  } else if (!strcmp(key, "surr_guess")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "surr_guess must be 0 or 1");
    c->surr_guess = (int)v;
  } else if (!strcmp(key, "surr_split_jac")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "surr_split_jac must be 0 or 1");
    c->surr_split_jac = (int)v;
  } else if (!strcmp(key, "surr_adaptive")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "surr_adaptive must be 0 or 1");
    c->surr_adaptive = (int)v;
  } else if (!strcmp(key, "surr_max_att")) {
    if (v < 1 || v > 1000) return fail(TGR_E_ARG, "surr_max_att must be 1..1000");
    c->surr_max_att = (int)v;
  } else if (!strcmp(key, "surr_probe")) {
    if (v < 0 || v > 1000000) return fail(TGR_E_ARG, "surr_probe must be >= 0");
    c->surr_probe = (int)v;
    c->surr_skipped = 0;
    c->surr_last_accept = 1.0;
  } else if (!strcmp(key, "surr_max_jac")) {
    if (v < 1 || v > 64) return fail(TGR_E_ARG, "surr_max_jac must be 1..64");
    c->surr_max_jac = (int)v;
  } else if (!strcmp(key, "surr_mode")) {
    if (v != 0 && v != 1) return fail(TGR_E_ARG, "surr_mode must be 0 (guided solve) or 1 (standalone surrogate)");
    c->surr_mode = (int)v;
  } else if (!strcmp(key, "surr_newton_max")) {
    if (v < 1 || v > 50) return fail(TGR_E_ARG, "surr_newton_max must be 1..50");
    c->surr_newton_max = (int)v;
  } else {
    return fail(TGR_E_ARG, std::string("unknown option: ") + key);
  }
  return 0;
}

int tgr_chem_source_chemistry_device(tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g,
                                     double* RHO_, double* SC_, const int* mask, double* HR_gp, double* chemsrc_gp,
                                     void* stream, tgr_stats* stats) {
  if (!c || !RHO_ || !SC_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  return run_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, g->isc_sc - 1, g->isc_T - 1, mask, HR_gp,
                  chemsrc_gp, (hipStream_t)stream, stats);
}

int tgr_chem_source_chemistry_device_noalloc(tgr_chem* c, double dt, double p_thermo, double p_thermo_old,
                                             const tgr_grid* g, double* RHO_, double* SC_, const int* mask,
                                             double* HR_gp, double* chemsrc_gp, void* stream) {
  if (!c || !RHO_ || !SC_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  void* scr = nullptr;
  size_t scr_bytes = 0;
  if (tgr_scratch_resolve(c->device, &scr, &scr_bytes) != 0)
    return fail(TGR_E_HIP, "device scratch not resolved for the context's device");
  return run_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, g->isc_sc - 1, g->isc_T - 1, mask, HR_gp,
                  chemsrc_gp, (hipStream_t)stream, nullptr, nullptr, nullptr, scr, scr_bytes);
}

int tgr_chem_set_surrogate(tgr_chem* c, tgr_mlp* net, const float* in_shift, const float* in_scale,
                           const double* out_scale) {
  if (!c) return fail(TGR_E_ARG, "NULL context");
  c->net = nullptr;
  if (!net) return 0;
  if (!in_shift || !in_scale || !out_scale) return fail(TGR_E_ARG, "NULL scaling array");
  int nin = 0, nout = 0;
  if (tgr_mlp_shape(net, &nin, &nout, nullptr, nullptr)) return fail(TGR_E_ARG, "bad network handle");
  if (nin != c->nsp + 2 || nout != c->neq)
    return fail(TGR_E_ARG, "network shape must be nsp + 2 inputs (log10 Y, T, log10 dt) and nsp + 1 outputs");
  int64_t np = 0;
  if (tgr_mlp_layout(net, &np, &c->net_in_pad, &c->net_out_pad)) return fail(TGR_E_ARG, "bad network handle");
  for (int k = 0; k < nin; ++k)
    if (!std::isfinite(in_shift[k]) || !std::isfinite(in_scale[k])) return fail(TGR_E_ARG, "non-finite input scaling");
  for (int k = 0; k < nout; ++k)
    if (!std::isfinite(out_scale[k])) return fail(TGR_E_ARG, "non-finite output scaling");
  c->h_in_shift.assign(c->net_in_pad, 0.0f);
  c->h_in_scale.assign(c->net_in_pad, 0.0f);
  for (int k = 0; k < nin; ++k) {
    c->h_in_shift[k] = in_shift[k];
    c->h_in_scale[k] = in_scale[k];
  }
  if (!c->d_in_shift) HIP_TRY(hipMalloc(&c->d_in_shift, 64 * 4 * 4));
  if (!c->d_in_scale) HIP_TRY(hipMalloc(&c->d_in_scale, 64 * 4 * 4));
  if (!c->d_out_scale) HIP_TRY(hipMalloc(&c->d_out_scale, 64 * 8));
  if (!c->d_rej_count) HIP_TRY(hipMalloc((void**)&c->d_rej_count, 256));
  if (!c->d_stotals) HIP_TRY(hipMalloc((void**)&c->d_stotals, 256));
  if (!c->ev2) HIP_TRY(hipEventCreate(&c->ev2));
  if (c->net_in_pad > 256) return fail(TGR_E_ARG, "network input too wide");
  HIP_TRY(hipMemcpy(c->d_in_shift, c->h_in_shift.data(), c->net_in_pad * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_in_scale, c->h_in_scale.data(), c->net_in_pad * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_out_scale, out_scale, c->neq * 8, hipMemcpyHostToDevice));
  c->net = net;
  c->surr_last_accept = 1.0;
  c->surr_skipped = 0;
  return 0;
}

int tgr_chem_set_option_real(tgr_chem* c, const char* key, double v) {
  if (!c || !key) return fail(TGR_E_ARG, "NULL argument");
  if (!std::isfinite(v)) return fail(TGR_E_ARG, "value must be finite");
  if (!strcmp(key, "surr_h0_fac")) {
    if (v < 0.0) return fail(TGR_E_ARG, "surr_h0_fac must be >= 0 (0: first sub-step of the schedule)");
    c->surr_h0_fac = v;
    return 0;
  }
  if (!(v > 0.0)) return fail(TGR_E_ARG, "value must be positive and finite");
  if (!strcmp(key, "surr_grow")) {
    if (v < 1.5 || v > 100.0) return fail(TGR_E_ARG, "surr_grow must be in [1.5, 100]");
    c->surr_grow = v;
  } else if (!strcmp(key, "surr_keep")) {
    if (v < 1.0 || v > 4.0) return fail(TGR_E_ARG, "surr_keep must be in [1, 4]");
    c->surr_keep = v;
  } else if (!strcmp(key, "surr_jac_theta")) {
    if (!(v > 0.0) || v >= 1.0) return fail(TGR_E_ARG, "surr_jac_theta must be in (0, 1)");
    c->surr_jac_theta = v;
  } else if (!strcmp(key, "surr_tol_y")) c->surr_tol_y = v;
  else if (!strcmp(key, "surr_tol_t")) c->surr_tol_t = v;
  else if (!strcmp(key, "surr_newton_tol")) c->surr_newton_tol = v;
  else if (!strcmp(key, "surr_sa_tol")) {
    if (!(v > 0.0)) return fail(TGR_E_ARG, "surr_sa_tol must be positive");
    c->surr_sa_tol = v;
  } else if (!strcmp(key, "surr_min_accept")) {
    if (v > 1.0) return fail(TGR_E_ARG, "surr_min_accept must be in (0, 1]");
    c->surr_min_accept = v;
    c->surr_skipped = 0;
    c->surr_last_accept = 1.0;
  } else if (!strcmp(key, "surr_grade")) {
    if (v < 1.0 || v > 16.0) return fail(TGR_E_ARG, "surr_grade must be in [1, 16]");
    c->surr_grade = v;
  }
  else return fail(TGR_E_ARG, std::string("unknown option: ") + key);
  return 0;
}

int tgr_chem_surrogate_features_device(tgr_chem* c, double dt, const tgr_grid* g, const double* SC_, uint16_t* X,
                                       int64_t rows, void* stream) {
  if (!c || !SC_ || !X) return fail(TGR_E_ARG, "NULL argument");
  if (!c->net) return fail(TGR_E_ARG, "no network set (tgr_chem_set_surrogate)");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(TGR_E_ARG, "dt must be positive and finite");
  StepParams p = step_params(c, dt, 0.0, 0.0, g, nullptr, (double*)SC_, g->isc_sc - 1, g->isc_T - 1, nullptr, nullptr,
                             nullptr);
  if (p.ncell <= 0) return 0;
  if (rows < p.ncell || rows % 64) return fail(TGR_E_ARG, "rows must be >= the interior cells and a multiple of 64");
  if (tgr_launch_surr_features(&p, c->neq, rows, c->net_in_pad, c->d_in_shift, c->d_in_scale, (float)std::log10(dt),
                               X, (hipStream_t)stream))
    return fail(TGR_E_HIP, "feature kernel launch failed");
  return 0;
}

}  // extern "C"

namespace {

// The surrogate-guided step on device fields whose species block starts at
// scalar plane isc_sc0 and whose T is plane isc_T0 (0-based).
int run_surrogate(tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g, double* RHO_,
                  double* SC_, int isc_sc0, int isc_T0, const int* mask, double* HR_gp, double* chemsrc_gp,
                  hipStream_t s, tgr_stats* stats, tgr_surrogate_stats* sstats) {
  int rc = 0;
  if (!c->net) return fail(TGR_E_ARG, "no network set (tgr_chem_set_surrogate)");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(TGR_E_ARG, "dt must be positive and finite");
  if (stats) memset(stats, 0, sizeof(*stats));
  if (sstats) memset(sstats, 0, sizeof(*sstats));
  StepParams p = step_params(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, isc_sc0, isc_T0, mask, HR_gp, chemsrc_gp);
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return 0;
  if (c->surr_probe > 0 && c->surr_last_accept < c->surr_min_accept && c->surr_skipped + 1 < c->surr_probe) {
    // the guided solve has not been paying: plain step only
    c->surr_skipped++;
    if (sstats) sstats->guided = 0;
    rc = run_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, isc_sc0, isc_T0, mask, HR_gp, chemsrc_gp, s, stats);
    if (!rc && sstats && stats) {
      sstats->rejected = stats->cells;
      sstats->fallback_ms = stats->kernel_ms;
    }
    return rc;
  }
  c->surr_skipped = 0;
  if (sstats) sstats->guided = 1;
  // standalone surrogate (surr_mode 1): the fixed schedule with one FP64
  // Newton correction per stage from the network's stage values, the stage
  // accepted when that correction -- the distance of the prediction from the
  // stage equation's solution -- is within surr_sa_tol acceptance tolerances
  const bool standalone = c->surr_mode == 1;
  const int adaptive = standalone ? 0 : c->surr_adaptive;
  const int64_t rows = (p.ncell + 63) / 64 * 64;
  // network output times: every stage of the fixed schedule, or the end of
  // each sub-step of the schedule for the adaptive solve (which interpolates)
  const int nst = adaptive ? c->surr_nsub : tgr::kSurrStages * c->surr_nsub;
  // sub-step sizes: geometric with ratio surr_grade (1: equal), summing to dt
  double hsub[tgr::kSurrMaxSub] = {0.0};
  {
    const int n = c->surr_nsub;
    const double r = c->surr_grade;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
      hsub[i] = (r == 1.0) ? dt / n : dt * std::pow(r, (double)i) * (r - 1.0) / (std::pow(r, (double)n) - 1.0);
      acc += hsub[i];
    }
    hsub[n - 1] += dt - acc;  // exact sum
  }
  const int grid_s = persistent_grid(c, tgr_surr_occupancy(c->neq, adaptive), p.ncell, tgr_surr_block_waves());
  int grid_p = 0;
  bool lanes_p = false;
  // the workspace serves both kernels: each wave of the guided kernel owns
  // neq x 64 doubles, the plain step on the rejected cells its own shape
  const size_t ws_bytes = std::max((size_t)grid_s * tgr_surr_block_waves() * c->neq * 64 * sizeof(double),
                                   step_shape(c, p.ncell, &grid_p, &lanes_p));
  // every buffer sized before the first launch (a resize frees memory)
  if ((rc = c->b_ws.ensure(ws_bytes)) ||
      (rc = c->b_x.ensure((size_t)rows * c->net_in_pad * 2)) ||
      (rc = c->b_guess.ensure((size_t)nst * rows * c->net_out_pad * 4)) ||
      (rc = c->b_rej.ensure((size_t)p.ncell * sizeof(int))))
    return rc;
  const bool timing = stats || sstats || c->surr_probe > 0;
  if (timing) HIP_TRY(hipEventRecord(c->ev0, s));
  uint16_t* X = (uint16_t*)c->b_x.p;
  float* guess = (float*)c->b_guess.p;
  if (c->surr_guess) {
    if (tgr_launch_surr_features(&p, c->neq, rows, c->net_in_pad, c->d_in_shift, c->d_in_scale,
                                 (float)std::log10(dt), X, s))
      return fail(TGR_E_HIP, "feature kernel launch failed");
    const int col = c->nsp + 1;
    static const double cst[tgr::kSurrStages] = {0.43586652150845899942, 0.5 * (1.0 + 0.43586652150845899942), 1.0};
    double t0 = 0.0;
    for (int k = 0; k < nst; ++k) {
      const int i = adaptive ? k : k / tgr::kSurrStages;
      const int j = adaptive ? tgr::kSurrStages - 1 : k % tgr::kSurrStages;
      if ((adaptive || j == 0) && i > 0) t0 += hsub[i - 1];
      const double tau = t0 + cst[j] * hsub[i];
      const float v = (float)((std::log10(tau) - c->h_in_shift[col]) * c->h_in_scale[col]);
      if (tgr_launch_surr_set_column(X, rows, c->net_in_pad, col, v, s)) return fail(TGR_E_HIP, "column launch");
      if (tgr_mlp_forward(c->net, X, rows, guess + (size_t)k * rows * c->net_out_pad, s))
        return fail(TGR_E_HIP, std::string("network forward failed: ") + tgr_mlp_last_error());
    }
  }
  SurrParams P{};
  P.s = p;
  P.s.queue = c->d_queue;
  P.s.ws = (double*)c->b_ws.p;
  P.adaptive = adaptive;
  P.max_att = c->surr_max_att;
  P.h0_fac = c->surr_h0_fac;
  P.ntau = nst;
  P.grow = c->surr_grow;
  P.split_jac = c->surr_split_jac;
  P.keep = c->surr_keep;
  P.jac_theta = c->surr_jac_theta;
  P.max_jac = c->surr_max_jac;
  {
    // times of the network outputs, in the order they were computed
    static const double cs[tgr::kSurrStages] = {0.43586652150845899942, 0.5 * (1.0 + 0.43586652150845899942), 1.0};
    double t0 = 0.0;
    for (int k = 0; k < nst; ++k) {
      const int i = adaptive ? k : k / tgr::kSurrStages;
      const int j = adaptive ? tgr::kSurrStages - 1 : k % tgr::kSurrStages;
      if ((adaptive || j == 0) && i > 0) t0 += hsub[i - 1];
      P.tau[k] = t0 + cs[j] * hsub[i];
    }
  }
  P.guess = guess;
  P.guess_stride = rows * c->net_out_pad;
  P.guess_ld = c->net_out_pad;
  for (int i = 0; i < tgr::kSurrMaxSub; ++i) P.hsub[i] = hsub[i];
  P.out_scale = c->d_out_scale;
  P.elem_w = c->d_elem;
  P.nsub = c->surr_nsub;
  P.use_guess = c->surr_guess;
  P.newton_max = standalone ? 1 : c->surr_newton_max;
  P.newton_tol = standalone ? c->surr_sa_tol : c->surr_newton_tol;
  P.tol_y = c->surr_tol_y;
  P.tol_t = c->surr_tol_t;
  P.rej_count = c->d_rej_count;
  P.rej_list = (int*)c->b_rej.p;
  P.stotals = c->d_stotals;
  if (tgr_launch_surr(&P, c->neq, grid_s, s)) return fail(TGR_E_HIP, "surrogate kernel launch failed");
  if (timing) HIP_TRY(hipEventRecord(c->ev2, s));
  // the plain integrator on the rejected cells
  if ((rc = run_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, isc_sc0, isc_T0, mask, HR_gp, chemsrc_gp, s,
                     nullptr, P.rej_list, c->d_rej_count)))
    return rc;
  // the accepted fraction steers the adaptive use: read back after every
  // guided call (a small copy; the call is then synchronous)
  const bool need_host = stats || sstats || c->surr_probe > 0;
  if (need_host) {
    HIP_TRY(hipEventRecord(c->ev1, s));
    unsigned long long tot[TGR_NCOUNT], st[tgr::kSurrNCount];
    HIP_TRY(hipMemcpyAsync(tot, c->d_totals, sizeof(tot), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(st, c->d_stotals, sizeof(st), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms_s = 0.f, ms_p = 0.f;
    if (stats || sstats) {
      HIP_TRY(hipEventElapsedTime(&ms_s, c->ev0, c->ev2));
      HIP_TRY(hipEventElapsedTime(&ms_p, c->ev2, c->ev1));
    }
    c->surr_last_accept = (double)st[0] / (double)std::max<unsigned long long>(1ull, st[0] + st[1]);
    if (stats) {
      int64_t cells = 0;
      if ((rc = count_cells(g, mask, p.mask_ld, p.ncell, &cells))) return rc;
      stats->cells = cells;
      stats->nst = tot[0];
      stats->nfe = tot[1];
      stats->nfe_dq = tot[2];
      stats->nje = tot[3];
      stats->nsetups = tot[4];
      stats->nni = tot[5];
      stats->netf = tot[6];
      stats->ncfn = tot[7];
      stats->nlu = tot[8];
      stats->failed = tot[9];
      stats->kernel_ms = ms_s + ms_p;
    }
    if (sstats) {
      sstats->accepted = st[0];
      sstats->rejected = st[1];
      sstats->newton_iters = st[2];
      sstats->rhs_evals = st[3];
      sstats->rej_newton = st[4];
      sstats->rej_error = st[5];
      sstats->rej_enthalpy = st[6];
      sstats->substeps = st[8];
      sstats->factorizations = st[7];
      sstats->jacobians = st[9];
      sstats->surrogate_ms = ms_s;
      sstats->fallback_ms = ms_p;
    }
  }
  return 0;
}

// host arrays: stage the block's density, species and T (and mask,
// diagnostics) in device buffers, run the plain or the surrogate-guided step,
// copy the results back; synchronous
int host_step(tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g, double* RHO_,
              double* SC_, const int* mask, double* HR_gp, double* chemsrc_gp, tgr_stats* stats, bool surrogate,
              tgr_surrogate_stats* sstats) {
  if (!c || !RHO_ || !SC_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  const int nsp = c->nsp;
  const size_t plane = (size_t)(g->imaxo_ - g->imino_ + 1) * (g->jmaxo_ - g->jmino_ + 1) * (g->kmaxo_ - g->kmino_ + 1);
  const size_t pb = plane * sizeof(double);
  const size_t mbytes = (size_t)(g->imaxo - g->imino + 1) * (g->jmaxo - g->jmino + 1) * sizeof(int);
  if ((rc = c->b_rho.ensure(pb)) || (rc = c->b_sc.ensure(pb * (nsp + 1)))) return rc;
  if (mask && (rc = c->b_mask.ensure(mbytes))) return rc;
  if (HR_gp && (rc = c->b_hr.ensure(pb))) return rc;
  if (chemsrc_gp && (rc = c->b_chem.ensure(3 * pb))) return rc;
  double* d_rho = (double*)c->b_rho.p;
  double* d_sc = (double*)c->b_sc.p;
  hipStream_t s = nullptr;
  // species block -> planes 0..nsp-1, T -> plane nsp
  HIP_TRY(hipMemcpyAsync(d_rho, RHO_, pb, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_sc, SC_ + (size_t)(g->isc_sc - 1) * plane, pb * nsp, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_sc + (size_t)nsp * plane, SC_ + (size_t)(g->isc_T - 1) * plane, pb,
                         hipMemcpyHostToDevice, s));
  if (mask) HIP_TRY(hipMemcpyAsync(c->b_mask.p, mask, mbytes, hipMemcpyHostToDevice, s));
  if (HR_gp) HIP_TRY(hipMemcpyAsync(c->b_hr.p, HR_gp, pb, hipMemcpyHostToDevice, s));
  if (chemsrc_gp) HIP_TRY(hipMemcpyAsync(c->b_chem.p, chemsrc_gp, 3 * pb, hipMemcpyHostToDevice, s));
  const int* dm = mask ? (const int*)c->b_mask.p : nullptr;
  double* dh = HR_gp ? (double*)c->b_hr.p : nullptr;
  double* dc = chemsrc_gp ? (double*)c->b_chem.p : nullptr;
  tgr_stats local;
  if (surrogate) rc = run_surrogate(c, dt, p_thermo, p_thermo_old, g, d_rho, d_sc, 0, nsp, dm, dh, dc, s,
                                    stats ? stats : &local, sstats);
  else rc = run_step(c, dt, p_thermo, p_thermo_old, g, d_rho, d_sc, 0, nsp, dm, dh, dc, s, stats ? stats : &local);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(SC_ + (size_t)(g->isc_sc - 1) * plane, d_sc, pb * nsp, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(SC_ + (size_t)(g->isc_T - 1) * plane, d_sc + (size_t)nsp * plane, pb,
                         hipMemcpyDeviceToHost, s));
  if (HR_gp) HIP_TRY(hipMemcpyAsync(HR_gp, c->b_hr.p, pb, hipMemcpyDeviceToHost, s));
  if (chemsrc_gp) HIP_TRY(hipMemcpyAsync(chemsrc_gp, c->b_chem.p, 3 * pb, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

extern "C" {

int tgr_chem_source_chemistry_surrogate_device(tgr_chem* c, double dt, double p_thermo, double p_thermo_old,
                                               const tgr_grid* g, double* RHO_, double* SC_, const int* mask,
                                               double* HR_gp, double* chemsrc_gp, void* stream, tgr_stats* stats,
                                               tgr_surrogate_stats* sstats) {
  if (!c || !RHO_ || !SC_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  return run_surrogate(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, g->isc_sc - 1, g->isc_T - 1, mask, HR_gp,
                       chemsrc_gp, (hipStream_t)stream, stats, sstats);
}

int tgr_chem_source_chemistry_surrogate(tgr_chem* c, double dt, double p_thermo, double p_thermo_old,
                                        const tgr_grid* g, double* RHO_, double* SC_, const int* mask, double* HR_gp,
                                        double* chemsrc_gp, tgr_stats* stats, tgr_surrogate_stats* sstats) {
  return host_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, mask, HR_gp, chemsrc_gp, stats, true, sstats);
}

int tgr_chem_load_surrogate(tgr_chem* c, const char* path) {
  if (!c || !path) return fail(TGR_E_ARG, "NULL argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TGR_E_ARG, std::string("cannot open ") + path);
  std::vector<char> raw;
  {
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) raw.insert(raw.end(), buf, buf + n);
    fclose(f);
  }
  if (raw.size() < 32 || memcmp(raw.data(), "TGRMLP01", 8) != 0)
    return fail(TGR_E_ARG, std::string(path) + ": not a network file");
  int32_t shp[4];
  int64_t nparam = 0;
  memcpy(shp, raw.data() + 8, sizeof(shp));
  memcpy(&nparam, raw.data() + 24, sizeof(nparam));
  const int nin = shp[0], nout = shp[1];
  if (nin != c->nsp + 2 || nout != c->neq)
    return fail(TGR_E_ARG, std::string(path) + ": network shape does not match the mechanism");
  if (nparam <= 0 || raw.size() != (size_t)32 + 8 * (size_t)nin + 8 * (size_t)nout + 4 * (size_t)nparam)
    return fail(TGR_E_ARG, std::string(path) + ": size does not match its header");
  std::vector<float> shift(nin), scale(nin), params((size_t)nparam);
  std::vector<double> oscale(nout);
  size_t o = 32;
  memcpy(shift.data(), raw.data() + o, 4 * (size_t)nin);
  memcpy(scale.data(), raw.data() + o + 4 * (size_t)nin, 4 * (size_t)nin);
  o += 8 * (size_t)nin;
  memcpy(oscale.data(), raw.data() + o, 8 * (size_t)nout);
  o += 8 * (size_t)nout;
  memcpy(params.data(), raw.data() + o, 4 * (size_t)nparam);
  tgr_mlp* net = nullptr;
  if (tgr_mlp_create(nin, nout, shp[2], shp[3], &net)) return fail(TGR_E_ARG, tgr_mlp_last_error());
  int64_t np = 0;
  int rc = 0;
  if (tgr_mlp_layout(net, &np, nullptr, nullptr) || np != nparam)
    rc = fail(TGR_E_ARG, std::string(path) + ": parameter count does not match the network shape");
  if (!rc && tgr_mlp_set_params(net, params.data())) rc = fail(TGR_E_ARG, tgr_mlp_last_error());
  if (!rc) rc = tgr_chem_set_surrogate(c, net, shift.data(), scale.data(), oscale.data());
  if (rc) {
    tgr_mlp_destroy(net);
    return rc;
  }
  if (c->own_net) tgr_mlp_destroy(c->own_net);
  c->own_net = net;
  return 0;
}

int tgr_chem_source_chemistry(tgr_chem* c, double dt, double p_thermo, double p_thermo_old, const tgr_grid* g,
                              double* RHO_, double* SC_, const int* mask, double* HR_gp, double* chemsrc_gp,
                              tgr_stats* stats) {
  return host_step(c, dt, p_thermo, p_thermo_old, g, RHO_, SC_, mask, HR_gp, chemsrc_gp, stats, false, nullptr);
}

int tgr_chem_prodrates_device(tgr_chem* c, int64_t n, const double* C, const double* T, const double* P,
                              double* cdot, void* stream) {
  if (!c || (n > 0 && (!C || !T || !P || !cdot))) return fail(TGR_E_ARG, "NULL argument");
  if (n < 0 || n > 0x7fffffff) return fail(TGR_E_ARG, "bad n");
  if (tgr_launch_prodrates(&c->dm, n, C, T, P, cdot, (hipStream_t)stream)) return fail(TGR_E_HIP, "launch");
  return 0;
}

int tgr_chem_properties_device(tgr_chem* c, double p_thermo, const tgr_grid* g, const double* SC_, const int* mask,
                               const double* Lewis, int isc_zmix, double* RHO_, double* VISC_, double* DIFF_,
                               void* stream) {
  if (!c || !SC_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  if (isc_zmix < 0 || isc_zmix > g->nscalar) return fail(TGR_E_ARG, "isc_zmix out of range");
  if (mask && (g->imino > g->imino_ || g->imaxo < g->imaxo_ || g->jmino > g->jmino_ || g->jmaxo < g->jmaxo_))
    return fail(TGR_E_ARG, "mask does not cover the block");
  tgr::PropsParams p{};
  p.ptab = (const double*)c->d_props;
  p.mtab = c->dm.tab;
  p.tmid = c->dm.tmid;
  p.p_thermo = p_thermo;
  p.nx = g->imaxo_ - g->imino_ + 1;
  p.ny = g->jmaxo_ - g->jmino_ + 1;
  p.nz = g->kmaxo_ - g->kmino_ + 1;
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return 0;
  p.plane = (int64_t)p.nx * p.ny * p.nz;
  p.isc_sc0 = g->isc_sc - 1;
  p.isc_T0 = g->isc_T - 1;
  p.isc_z0 = isc_zmix - 1;
  p.mask = mask;
  p.mask_ld = g->imaxo - g->imino + 1;
  p.mask_base = (int64_t)(g->imino_ - g->imino) + (int64_t)(g->jmino_ - g->jmino) * p.mask_ld;
  p.lewis = Lewis;
  p.sc = SC_;
  p.rho = RHO_;
  p.visc = VISC_;
  p.diff = DIFF_;
  if (tgr_launch_props(&p, c->neq, c->props_lane, (hipStream_t)stream)) return fail(TGR_E_HIP, "launch");
  return 0;
}

int tgr_chem_diffusion_device(tgr_chem* c, double dt, const tgr_grid* g, const tgr_metric* m, int flags,
                              const double* RHO_, const double* SC_, const double* DIFF_, const int* mask,
                              double* diff_corr, double* diff_gp, double* src, void* stream) {
  if (!c || !m || !RHO_ || !SC_ || !DIFF_) return fail(TGR_E_ARG, "NULL argument");
  int rc = check_grid(g, c->nsp);
  if (rc) return rc;
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(TGR_E_ARG, "dt must be positive and finite");
  if (m->st1 < 0 || m->st2 < 1 || m->st1 + m->st2 > 8) return fail(TGR_E_ARG, "stencil widths out of range");
  const int w = m->st1 + m->st2;
  if (g->imin_ - w < g->imino_ || g->imax_ + w > g->imaxo_ || g->jmin_ - w < g->jmino_ || g->jmax_ + w > g->jmaxo_ ||
      g->kmin_ - w < g->kmino_ || g->kmax_ + w > g->kmaxo_)
    return fail(TGR_E_ARG, "the block needs st1 + st2 ghost cells on every side");
  if (mask && (g->imino > g->imino_ || g->imaxo < g->imaxo_ || g->jmino > g->jmino_ || g->jmaxo < g->jmaxo_))
    return fail(TGR_E_ARG, "mask must cover the array extents");
  static_assert(sizeof(tgr_metric) == sizeof(tgr::tgr_metric_dev) &&
                    offsetof(tgr_metric, grad_zm) == offsetof(tgr::tgr_metric_dev, grad_zm),
                "tgr_metric and its device view must share a layout");
  const tgr::tgr_metric_dev* md = (const tgr::tgr_metric_dev*)m;
  const double* const* ops = &md->interp_sc_x;
  for (int q = 0; q < 15; ++q)
    if (!ops[q]) return fail(TGR_E_ARG, "a metric operator is NULL");
  tgr::DiffParams p{};
  tgr::DiffGeom& gm = p.gm;
  gm.imin_ = g->imin_, gm.imax_ = g->imax_, gm.jmin_ = g->jmin_, gm.jmax_ = g->jmax_, gm.kmin_ = g->kmin_,
  gm.kmax_ = g->kmax_;
  gm.imino_ = g->imino_, gm.jmino_ = g->jmino_, gm.kmino_ = g->kmino_;
  gm.nxo = g->imaxo_ - g->imino_ + 1, gm.nyo = g->jmaxo_ - g->jmino_ + 1, gm.nzo = g->kmaxo_ - g->kmino_ + 1;
  gm.imino = g->imino, gm.jmino = g->jmino, gm.mx = g->imaxo - g->imino + 1;
  gm.plane = (int64_t)gm.nxo * gm.nyo * gm.nzo;
  gm.isc_sc0 = g->isc_sc - 1, gm.isc_T0 = g->isc_T - 1;
  p.m = *md;
  p.mech = c->dm;
  p.dt = dt;
  p.molar = (flags & TGR_DIFF_MOLAR) != 0;
  p.gasrad = (flags & TGR_DIFF_GASRAD) != 0;
  p.cv = (flags & TGR_DIFF_CV) != 0;
  p.rho = RHO_, p.sc = SC_, p.diff = DIFF_, p.mask = mask;
  p.corr = diff_corr, p.dgp = diff_gp, p.src = src;
  const char* names[4] = {"H2O", "CO", "CO2", "CH4"};
  int* idx[4] = {&p.i_h2o, &p.i_co, &p.i_co2, &p.i_ch4};
  double* wk[4] = {&p.w_h2o, &p.w_co, &p.w_co2, &p.w_ch4};
  for (int q = 0; q < 4; ++q) {
    *idx[q] = species_index(c->t, names[q]);
    *wk[q] = *idx[q] >= 0 ? c->t->W[*idx[q]] : 1.0;
  }
  if ((rc = c->b_diff.ensure((size_t)gm.plane * 5 * sizeof(double)))) return rc;
  p.wg = (double*)c->b_diff.p;
  p.cpm = p.wg + gm.plane;
  p.cvel = p.cpm + gm.plane;
  if (tgr_launch_diffusion(&p, c->neq, (hipStream_t)stream)) return fail(TGR_E_HIP, "diffusion kernel launch failed");
  return 0;
}

int tgr_chem_temperature_device(tgr_chem* c, int64_t n, const double* Y, const double* hmix, const double* T0,
                                double* T, int* iters, void* stream) {
  if (!c || (n > 0 && (!Y || !hmix || !T))) return fail(TGR_E_ARG, "NULL argument");
  if (n < 0 || n > 0x7fffffff) return fail(TGR_E_ARG, "bad n");
  if (tgr_launch_temperature(&c->dm, n, Y, hmix, T0, T, iters, (hipStream_t)stream)) return fail(TGR_E_HIP, "launch");
  return 0;
}

int tgr_chem_rhs_device(tgr_chem* c, int64_t n, double p_gp, const double* sol, double* rhs, void* stream) {
  if (!c || (n > 0 && (!sol || !rhs))) return fail(TGR_E_ARG, "NULL argument");
  if (n < 0 || n > 0x7fffffff) return fail(TGR_E_ARG, "bad n");
  const int rc = c->rhs_lanes ? tgr_launch_rhs_lanes(c->neq, n, p_gp, sol, rhs, (hipStream_t)stream)
                              : tgr_launch_rhs(&c->dm, n, p_gp, sol, rhs, (hipStream_t)stream);
  if (rc) return fail(TGR_E_HIP, "launch");
  return 0;
}

}  // extern "C"

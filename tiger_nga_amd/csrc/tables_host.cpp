// Host-side construction of the device table block of a mechanism (layout
// tgr::TabLayout).  Shared by the C-ABI layer and the CPU wave emulator.
#include "tables_host.h"

#include <algorithm>
#include <utility>
#include <vector>

#include "chem_device.h"

namespace tgr {

static int seterr(std::string* err, const char* msg) {
  if (err) *err = msg;
  return -1;
}

// Builds the table block (layout tgr::TabLayout for the mechanism's
// compile-time shape): reactions re-ordered fall-off first (so only the first
// 64-reaction round of a wave diverges), species padded to 64 lanes, the
// third-body correction lists and the CDOT gather schedule.
int build_table_blob(const tgr_mech_tables* t, std::vector<char>& blob, tgr_mech_info* info, std::string* err) {
  const int nsp = t->nsp, nr = t->nrxn;
  tgr::TabLayout L(0, 0, 0, 0, 0, 0);
  switch (nsp + 1) {
    case 10: L = tgr::tab_layout<10>(); break;
    case 37: L = tgr::tab_layout<37>(); break;
    case 40: L = tgr::tab_layout<40>(); break;
    default: return seterr(err, "no kernel instance for this mechanism size");
  }
  std::vector<int> order;
  for (int r = 0; r < nr; ++r)
    if (t->rxn[r].kind == 2) order.push_back(r);
  const int nfo = (int)order.size();
  for (int r = 0; r < nr; ++r)
    if (t->rxn[r].kind != 2) order.push_back(r);
  // third-body sums as total concentration plus (eff - 1) corrections
  std::vector<std::vector<std::pair<int, double>>> corr(t->n_third);
  int maxe = 0;
  for (int j = 0; j < t->n_third; ++j) {
    for (int k = 0; k < nsp; ++k) {
      const double e = t->eff[(size_t)j * nsp + k];
      if (e != 1.0) corr[j].push_back({k, e - 1.0});
    }
    maxe = std::max(maxe, (int)corr[j].size());
  }
  // CDOT gather schedule: smallest chunk S with sum_k ceil(n_k/S) <= 64
  std::vector<std::vector<std::pair<int, double>>> bysp(nsp);
  for (int g = 0; g < nr; ++g) {
    const tgr_rxn& x = t->rxn[order[g]];
    for (int u = 0; u < x.nnu; ++u) bysp[t->nu[x.nu0 + u].sp].push_back({g, t->nu[x.nu0 + u].nu});
  }
  int S = 1;
  for (;; ++S) {
    int lanes = 0;
    for (int k = 0; k < nsp; ++k) lanes += ((int)bysp[k].size() + S - 1) / S;
    if (lanes <= 64) break;
  }
  if (nr != L.nrxn || nfo != L.nfo || t->n_third != L.nthird || S != L.gmax || maxe != L.maxe)
    return seterr(err, "mechanism tables do not match the compiled kernel shape");

  blob.assign(L.bytes, 0);
  auto D = [&](size_t off) -> double& { return *reinterpret_cast<double*>(&blob[off]); };
  auto I = [&](size_t off) -> int& { return *reinterpret_cast<int*>(&blob[off]); };
  for (int g = 0; g < nr; ++g) {
    const tgr_rxn& x = t->rxn[order[g]];
    const size_t rx = L.rx + (size_t)g * 64;
    for (int q = 0; q < 3; ++q) {
      D(rx + 8 * q) = x.f[q];
      D(rx + 24 + 8 * q) = x.b[q];
    }
    if (x.m + 1 > 127 || x.m + 1 < 0) return seterr(err, "third-body index out of range");
    const int fo = (x.kind == 2) ? g : 0;
    I(rx + 48) = (x.kind & 3) | ((x.rev & 1) << 2) | ((x.nr & 3) << 3) | ((x.np & 3) << 5) |
                 (((x.m + 1) & 127) << 7) | ((fo & 255) << 14);
    int rs = 0, ps = 0;
    for (int q = 0; q < 3; ++q) {
      rs |= (x.reac[q] < 0 ? 0 : x.reac[q]) << (8 * q);
      ps |= (x.prod[q] < 0 ? 0 : x.prod[q]) << (8 * q);
    }
    I(rx + 52) = rs;
    I(rx + 56) = ps;
    if (x.kind == 2) {
      const size_t f = L.fo + (size_t)g * 128;
      int kinds = 0;
      for (int q = 0; q < 3; ++q) {
        D(f + 8 * q) = x.f0[q];
        D(f + 24 + 8 * q) = x.b0[q];
        D(f + 48 + 8 * q) = x.fc_c[q];
        D(f + 72 + 8 * q) = x.fc_d[q];
        kinds |= (x.fc_kind[q] & 1) << q;
      }
      I(f + 96) = x.nfc;
      I(f + 100) = kinds;
    }
  }
  for (int rg = 0; rg < 2; ++rg)
    for (int k = 0; k < nsp; ++k)
      for (int q = 0; q < 13; ++q) D(L.th + ((size_t)(rg * 13 + q) * 64 + k) * 8) = t->thermo[(size_t)(rg * nsp + k) * 13 + q];
  for (int k = 0; k < nsp; ++k) {
    D(L.w + 8 * k) = t->W[k];
    D(L.iw + 8 * k) = 1.0 / t->W[k];
  }
  for (int j = 0; j < t->n_third; ++j) {
    I(L.en + 4 * j) = (int)corr[j].size();
    for (size_t e = 0; e < corr[j].size(); ++e) {
      I(L.ek + (e * 64 + j) * 4) = corr[j][e].first;
      D(L.ev + (e * 64 + j) * 8) = corr[j][e].second;
    }
  }
  int lane = 0;
  for (int k = 0; k < nsp; ++k) {
    const int nk = (int)bysp[k].size();
    const int chunks = (nk + S - 1) / S;
    I(L.sf + 4 * k) = lane;
    I(L.sn + 4 * k) = chunks;
    for (int ch = 0; ch < chunks; ++ch, ++lane) {
      const int lo = ch * S, hi = std::min(nk, lo + S);
      I(L.gc + 4 * lane) = hi - lo;
      for (int e = lo; e < hi; ++e) {
        I(L.gr + ((size_t)(e - lo) * 64 + lane) * 4) = bysp[k][e].first;
        D(L.gn + ((size_t)(e - lo) * 64 + lane) * 8) = bysp[k][e].second;
      }
    }
  }
  tgr_mech_info& in = *info;
  in.nsp = nsp;
  in.nrxn = nr;
  in.nslot = t->nslot;
  in.n_third = t->n_third;
  in.n_falloff = nfo;
  in.n_nu = t->n_nu;
  in.n_third_body_rxn = 0;
  in.n_reversible = 0;
  in.n_fc_terms = 0;
  in.n_conc_factors = 0;
  in.n_rev_falloff = 0;
  in.n_eff_corr = 0;
  for (int j = 0; j < t->n_third; ++j) in.n_eff_corr += (int)corr[j].size();
  for (int r = 0; r < nr; ++r) {
    const tgr_rxn& x = t->rxn[r];
    in.n_third_body_rxn += x.kind == 1;
    in.n_reversible += x.rev;
    in.n_conc_factors += x.nr + (x.rev ? x.np : 0);
    if (x.kind == 2) {
      in.n_fc_terms += x.nfc;
      in.n_rev_falloff += x.rev;
    }
  }
  return 0;
}

}  // namespace tgr

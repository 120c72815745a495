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

static int seterr(std::string* err, const std::string& msg) { return seterr(err, msg.c_str()); }

// Builds the table block (layout tgr::TabLayout for the mechanism's
// compile-time shape): reactions re-ordered fall-off first (so only the first
// 64-reaction round of a wave diverges), species padded to 64 lanes, the
// third-body correction lists and the CDOT gather schedule, every list padded
// to its row length with entries that leave the result unchanged.
int build_table_blob(const tgr_mech_tables* t, std::vector<char>& blob, tgr_mech_info* info, std::string* err) {
  const int nsp = t->nsp, nr = t->nrxn;
  TabLayout L(0, 0, 0, 0, 0, 0);
  switch (nsp + 1) {
    case 10: L = tgr::tab_layout<10>(); break;
    case 37: L = tgr::tab_layout<37>(); break;
    case 40: L = tgr::tab_layout<40>(); break;
    case 46: L = tgr::tab_layout<46>(); break;
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
    return seterr(err, "mechanism tables do not match the compiled kernel shape (nrxn " + std::to_string(nr) +
                           ", nfo " + std::to_string(nfo) + ", nthird " + std::to_string(t->n_third) + ", gmax " +
                           std::to_string(S) + ", maxe " + std::to_string(maxe) + ")");
  if (nsp > kZero || t->n_third > kZero) return seterr(err, "too many species or third-body sums for 64 lanes");
  if (nr >= L.rounds * 64) return seterr(err, "no padding slot in the last reaction round");
  const int qzero = L.rounds * 64 - 1;  // rate-of-progress slot of a padding reaction (always 0)

  blob.assign(L.bytes, 0);
  auto D = [&](size_t off) -> double& { return *reinterpret_cast<double*>(&blob[off]); };
  auto I = [&](size_t off) -> int& { return *reinterpret_cast<int*>(&blob[off]); };
  // reactions; padding lanes: A = 0 and every factor on the 1.0 slot
  for (int g = 0; g < L.rounds * 64; ++g) {
    const size_t rx = L.rx + (size_t)(g / 64) * kRxRound, dl = (size_t)(g % 64) * 8, il = (size_t)(g % 64) * 4;
    int i0 = kOne | kOne << 8 | kOne << 16 | kOne << 24, i1 = kOne | kOne << 8 | kOne << 16;
    if (g < nr) {
      const tgr_rxn& x = t->rxn[order[g]];
      if (x.nr > 3 || x.np > 3) return seterr(err, "more than three concentration factors");
      for (int q = 0; q < 3; ++q) {
        D(rx + q * kRow + dl) = x.f[q];
        D(rx + (3 + q) * kRow + dl) = x.rev ? x.b[q] : 0.0;
      }
      auto put = [](int& w, int q, int v) { w = (w & ~(255 << (8 * q))) | (v << (8 * q)); };
      for (int q = 0; q < x.nr; ++q) put(i0, q, x.reac[q]);
      for (int q = 0; q < x.np; ++q) put(i1, q, x.prod[q]);
      if (x.kind == 1) {
        if (x.m < 0 || x.m >= t->n_third) return seterr(err, "third-body index out of range");
        put(i0, 3, x.m);
      }
      if (x.kind == 2) {
        if (g >= 64) return seterr(err, "more than 64 fall-off reactions");
        i1 |= 1 << 24;
        const size_t f = L.fo + dl;
        int kinds = 0;
        for (int q = 0; q < 3; ++q) {
          D(f + q * kRow) = x.f0[q];
          D(f + (3 + q) * kRow) = x.rev ? x.b0[q] : 0.0;
          if (q < x.nfc) {
            D(f + (6 + q) * kRow) = x.fc_c[q];
            // exp(-T / d) as exp(-(1/d) T); exp(-d / T) as exp(-d (1/T))
            D(f + (9 + q) * kRow) = x.fc_kind[q] ? x.fc_d[q] : 1.0 / x.fc_d[q];
            kinds |= (x.fc_kind[q] & 1) << q;
          }
        }
        if (x.m >= t->n_third) return seterr(err, "third-body index out of range");
        I(L.fo + 12 * kRow + il) = kinds | (x.m >= 0 ? x.m : kZero) << 8;
      }
      if (x.rev) i1 |= 1 << 25;
    }
    I(rx + 6 * kRow + il) = i0;
    I(rx + 6 * kRow + kIRow + il) = i1;
  }
  for (int rg = 0; rg < 2; ++rg)
    for (int k = 0; k < nsp; ++k)
      for (int q = 0; q < 13; ++q) D(L.th + (size_t)(rg * 13 + q) * kRow + k * 8) = t->thermo[(size_t)(rg * nsp + k) * 13 + q];
  for (int k = 0; k < nsp; ++k) {
    D(L.w + 8 * k) = t->W[k];
    D(L.iw + 8 * k) = 1.0 / t->W[k];
  }
  // third-body corrections; padding: weight 0 on the 1.0 slot
  for (int j = 0; j < 64; ++j)
    for (int e = 0; e < maxe; ++e) {
      const bool on = j < t->n_third && e < (int)corr[j].size();
      I(L.ek + e * kIRow + j * 4) = on ? corr[j][e].first : kOne;
      D(L.ev + e * kRow + j * 8) = on ? corr[j][e].second : 0.0;
    }
  // CDOT gather; padding: weight 0 on a padding reaction's slot
  for (int ln = 0; ln < 64; ++ln)
    for (int e = 0; e < S; ++e) I(L.gr + e * kIRow + ln * 4) = qzero;
  int lane = 0;
  for (int k = 0; k < nsp; ++k) {
    const int nk = (int)bysp[k].size();
    const int chunks = (nk + S - 1) / S;
    I(L.sf + 4 * k) = lane;
    I(L.sn + 4 * k) = chunks;
    for (int ch = 0; ch < chunks; ++ch, ++lane) {
      const int lo = ch * S, hi = std::min(nk, lo + S);
      for (int e = lo; e < hi; ++e) {
        I(L.gr + (size_t)(e - lo) * kIRow + lane * 4) = bysp[k][e].first;
        D(L.gn + (size_t)(e - lo) * kRow + lane * 8) = bysp[k][e].second;
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

// CPU wave emulator for the chemistry solver's device code (test build only).
//
// Compiling tiger_nga_amd/csrc/*.h with -DTGR_WAVE_EMU -I tests/emu turns the
// wave-per-cell solver into plain C++ in which every lane of the 64-wide
// wavefront is a fiber (ucontext) on one host thread, switched round-robin.
// The cross-lane operations the solver uses are emulated with a barrier per
// operation (a full round of the 64 fibers):
//   v_readlane / v_readfirstlane, ballot, lane shuffle (ds_bpermute), the DPP
//   row_shr:1-15 / row_bcast:15 / row_bcast:31 moves of the reduction trees,
//   and the intra-wave LDS ordering point (wave_barrier).
// Device code must call these in wave-uniform control flow (it has to on the
// GPU as well for readlane/DPP to see defined inputs); a divergent call
// deadlocks here, which is how this build catches it.
#pragma once
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
// function-scope LDS: one object shared by the lanes (fibers) of the wave
#define __shared__ static

namespace tgr_emu {

struct Wave {
  ucontext_t ctx[64];
  ucontext_t home;
  uint64_t slot[64];
  int done = 0;
};

inline Wave* tw = nullptr;
inline int tl_lane = 0;

// barrier: hand over to the next lane; lane 63 hands back to lane 0, so a
// full round separates what every lane did before the call from what any
// lane does after it
inline void wsync() {
  const int me = tl_lane, nx = (me + 1) & 63;
  tl_lane = nx;
  swapcontext(&tw->ctx[me], &tw->ctx[nx]);
  tl_lane = me;
}

inline uint64_t fetch(uint64_t v, int src) {
  tw->slot[tl_lane] = v;
  wsync();
  const uint64_t r = tw->slot[src & 63];
  wsync();
  return r;
}

}  // namespace tgr_emu

inline int __lane_id() { return tgr_emu::tl_lane; }

inline int __builtin_amdgcn_readlane(int v, int src) { return (int)(uint32_t)tgr_emu::fetch((uint32_t)v, src); }
inline int __builtin_amdgcn_readfirstlane(int v) { return __builtin_amdgcn_readlane(v, 0); }

inline unsigned long long __ballot(int pred) {
  using namespace tgr_emu;
  tw->slot[tl_lane] = pred ? 1u : 0u;
  wsync();
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i)
    if (tw->slot[i]) m |= 1ull << i;
  wsync();
  return m;
}

inline double __shfl(double v, int src, int width = 64) {
  (void)width;
  uint64_t b;
  std::memcpy(&b, &v, 8);
  b = tgr_emu::fetch(b, src);
  double r;
  std::memcpy(&r, &b, 8);
  return r;
}

// DPP move on 32-bit values: row_shr:n (0x111-0x11F), row_bcast:15 (0x142),
// row_bcast:31 (0x143); lanes of disabled rows/banks keep `old`, lanes whose
// source is out of range get 0 with bound_ctrl, else keep `old`.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  using namespace tgr_emu;
  const int lane = tl_lane;
  tw->slot[lane] = (uint32_t)src;
  wsync();
  const int row = lane >> 4, lir = lane & 15, bank = lir >> 2;
  const bool enabled = ((row_mask >> row) & 1) && ((bank_mask >> bank) & 1);
  int res = old;
  if (enabled) {
    int sl = -1;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {
      const int n = ctrl - 0x110;
      if (lir - n >= 0) sl = lane - n;
    } else if (ctrl == 0x142) {
      if (row > 0) sl = (row - 1) * 16 + 15;
    } else if (ctrl == 0x143) {
      if (row >= 2) sl = 31;
    }
    if (sl >= 0) res = (int)(uint32_t)tw->slot[sl];
    else res = bound_ctrl ? 0 : old;
  }
  wsync();
  return res;
}

inline void __builtin_amdgcn_wave_barrier() { tgr_emu::wsync(); }
#define __builtin_amdgcn_fence(...) ((void)0)

inline int __double2loint(double d) {
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return (int)(uint32_t)b;
}
inline int __double2hiint(double d) {
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return (int)(uint32_t)(b >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const uint64_t b = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::frexp;
using std::isfinite;
using std::ldexp;
using std::rint;
using std::sqrt;

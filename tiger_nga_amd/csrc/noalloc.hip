// Device scratch of the chemistry step's entry point that allocates nothing
// (tgr_chem_source_chemistry_device_noalloc, include/tiger_chem.h).
//
// The plain step's kernels need memory beyond the caller's fields: the work
// counter the persistent waves pull cells from, the counter totals, and one
// saved-Jacobian workspace per wave (neq x 64 doubles).  The ordinary entry
// points take these from context buffers that grow with hipMalloc; callers
// that may not allocate (code under HIP graph capture, tight serving loops)
// get them from this file-scope array instead.  Each GPU the library runs on
// holds its own copy of the array, at its own address, until the process
// ends.  The array's name is a host-side stand-in in host code: every use
// goes through the device address that hipGetSymbolAddress returns for the
// current device (resolved once per device, at context creation).
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <mutex>

namespace {

// 48 MiB: the persistent grid at full occupancy on MI355X (256 CUs x 8 waves)
// with the largest compiled mechanism (EthyRedF, neq 46: 23 KiB per wave) needs
// 46 MiB of workspace, plus the 256-byte counter header.
constexpr size_t kScratchBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];
std::mutex g_mu;

}  // namespace

// Device address and size of this process's scratch on `device`.  The first
// call for a device must be made with that device current (it resolves the
// address with hipGetSymbolAddress); later calls return the cached address.
// Returns 0, or -1 (device id out of range), -2 (first call with another
// device current), -3 (hipGetSymbolAddress failed).
extern "C" int tgr_scratch_resolve(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices || !ptr || !bytes) return -1;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return -2;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -3;
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  *bytes = kScratchBytes;
  return 0;
}

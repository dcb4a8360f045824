// Device scratch of the no-allocation filter entry point.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the block
// Clenshaw filter through smalgo_apply_poly_block_noalloc.  The recurrence's
// second ping-pong block lives in `scratch` below, a file-scope device array
// that every GPU the library runs on holds until the process ends.  Its
// device address is looked up once per device, with that device current, when
// a context is created (hipGetSymbolAddress); host code then only ever uses
// that address, never the array's name (which on the host is a stand-in, not
// a device pointer).  Every call on one GPU shares the array: calls that
// overlap on one GPU (two contexts, two threads) race.  The recurrence writes
// every scratch element before reading it, so nothing is cleared between
// calls.
#include <mutex>
#include <string>

#include "smalgo_solver.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_addr[kMaxDevices] = {};
std::string g_fail[kMaxDevices];

}  // namespace

namespace smalgo {

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_addr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    g_fail[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return;
  }
  g_addr[device] = static_cast<double*>(p);
  g_fail[device].clear();
}

double* noalloc_scratch(int device, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices) throw Error(-3, "device id beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[device])
    throw Error(-2, g_fail[device].empty() ? std::string("no-allocation scratch not set up on "
                                                         "this device")
                                           : g_fail[device]);
  if (bytes) *bytes = kScratchBytes;
  return g_addr[device];
}

}  // namespace smalgo

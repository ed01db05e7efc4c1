// An entry point that allocates nothing: the stable (key, value) radix sort of
// bk_sort_pairs with its scratch in a file-scope device array instead of a
// workspace, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
#include "bk_workspace.h"

#include <mutex>
#include <string>
#include <unordered_map>

namespace {

thread_local std::string g_noalloc_error;

int noalloc_fail(const std::string& what) {
  g_noalloc_error = what;
  return -1;
}

// The scratch: two key and two value ping-pong arrays of kNoAllocPairs entries,
// the per-workgroup digit histograms and the 256 digit totals (16.5 MiB), held
// by every GPU the library runs on until the process ends.  In host code
// `scratch` names a host-side stand-in, so host code uses only the address
// hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocPairs = 1 << 20;
constexpr int kNoAllocWgs = (kNoAllocPairs + kRadixTile - 1) / kRadixTile;
constexpr size_t kPairBytes = size_t(4) * kNoAllocPairs;
constexpr size_t kHistBytes = size_t(4) * 256 * kNoAllocWgs;
constexpr size_t kNoAllocBytes = 4 * kPairBytes + kHistBytes + size_t(4) * 256;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mutex;
std::unordered_map<int, uint8_t*> g_scratch_ptr;  // device ordinal -> device address of `scratch` there

// The device address of `scratch` on the current device, looked up on the first call on that device.
uint8_t* scratch_on_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    noalloc_fail(std::string("bk_sort_pairs_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  uint8_t* ptr = nullptr;
  e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    noalloc_fail(std::string("bk_sort_pairs_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    return nullptr;
  }
  g_scratch_ptr[dev] = ptr;
  return ptr;
}

}  // namespace

extern "C" {

const char* bk_noalloc_last_error(void) { return g_noalloc_error.c_str(); }

int32_t bk_sort_pairs_noalloc_capacity(void) { return kNoAllocPairs; }

int bk_sort_pairs_noalloc(const uint32_t* keys_in, const int32_t* values_in, uint32_t* keys_out, int32_t* values_out, int32_t n,
                          int32_t key_bits, void* stream) {
  if (n > kNoAllocPairs)
    return noalloc_fail("bk_sort_pairs_noalloc: " + std::to_string(n) + " pairs exceed the static scratch's " +
                        std::to_string(kNoAllocPairs) + "; use bk_sort_pairs with a workspace");
  if (n < 0 || key_bits < 1 || key_bits > 32 || (n > 0 && (!keys_in || !keys_out || !values_out)))
    return noalloc_fail("bk_sort_pairs_noalloc: bad argument");
  if (n == 0) return 0;
  uint8_t* base = scratch_on_current_device();
  if (!base) return -2;
  bk_workspace view;  // only the scratch fields the radix sort reads, pointing into this device's array
  view.keys_s[0] = reinterpret_cast<uint32_t*>(base);
  view.keys_s[1] = reinterpret_cast<uint32_t*>(base + kPairBytes);
  view.vals_s[0] = reinterpret_cast<int32_t*>(base + 2 * kPairBytes);
  view.vals_s[1] = reinterpret_cast<int32_t*>(base + 3 * kPairBytes);
  view.hist = reinterpret_cast<int32_t*>(base + 4 * kPairBytes);
  view.digit_total = reinterpret_cast<int32_t*>(base + 4 * kPairBytes + kHistBytes);
  int rc = bk_internal_radix_sort(&view, keys_in, values_in, keys_out, values_out, n, key_bits, (hipStream_t)stream);
  if (rc) noalloc_fail("bk_sort_pairs_noalloc: kernel launch failed");
  return rc;
}

}  // extern "C"

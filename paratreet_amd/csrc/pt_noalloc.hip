// pt_sort_noalloc: the radix sort of pt_sort over caller-owned device arrays,
// with every temporary (digit-count table, scan partials, ping-pong key and
// index buffers) in one file-scope device array instead of memory the library
// allocates.  For callers that cannot allocate device memory: code under HIP
// graph capture, tight loops, harnesses that forbid hipMalloc.
//
// The array is per GPU (each device holds its own copy at its own address,
// resolved once per device with hipGetSymbolAddress) and lives until the
// process ends.  Every call on one GPU shares it, so calls that overlap on one
// GPU (two streams, two host threads) race; serialise them.  Every byte a call
// reads from it is written earlier in the same call (count tables are fully
// rewritten per pass, the scan is reduce-then-scan with no flags), so nothing
// needs clearing between calls.
#include <mutex>
#include <string>

#include "../../include/pt_gravity.h"
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int64_t kNoAllocMaxN = PT_SORT_NOALLOC_MAX_N;
// pt_sort_scratch_bytes(kNoAllocMaxN) = 26,216,576 bytes, plus up to 3 x 255
// bytes of the 256-byte rounding pt_radix_sort applies to its pieces.
constexpr size_t kNoAllocBytes = (size_t)26 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};
thread_local std::string g_err;

pt_status noalloc_fail(pt_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

}  // namespace

// pt_last_error(NULL) returns this thread's last pt_sort_noalloc message.
const char* pt_noalloc_last_error() { return g_err.empty() ? nullptr : g_err.c_str(); }

extern "C" pt_status pt_sort_noalloc(const uint64_t* keys, int64_t n, uint64_t* sorted_keys,
                                     uint32_t* perm, void* stream) {
  g_err.clear();
  if (n < 0) return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: negative n");
  if (n > kNoAllocMaxN)
    return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: n = " + std::to_string(n) +
                                        " exceeds the maximum of " + std::to_string(kNoAllocMaxN) +
                                        " keys (PT_SORT_NOALLOC_MAX_N); use pt_sort");
  if (n == 0) return PT_OK;
  if (!keys || !sorted_keys || !perm)
    return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: null array");
  const size_t need = (size_t)pt_sort_scratch_bytes(n) + 3 * 256;
  if (need > kNoAllocBytes)
    return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: scratch needs " + std::to_string(need) +
                                        " bytes, the array holds " +
                                        std::to_string(kNoAllocBytes));
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return noalloc_fail(PT_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: device id out of range");
  hipStream_t s = (hipStream_t)stream;
  if (s) {
    hipDevice_t sdev = -1;
    e = hipStreamGetDevice(s, &sdev);
    if (e != hipSuccess)
      return noalloc_fail(PT_ERR_HIP, std::string("hipStreamGetDevice: ") + hipGetErrorString(e));
    if (sdev != dev)
      return noalloc_fail(PT_ERR_ARG, "pt_sort_noalloc: the stream belongs to device " +
                                          std::to_string(sdev) + ", the current device is " +
                                          std::to_string(dev));
  }
  void* ptr = nullptr;
  {
    // first call on this device: resolve this device's copy of the array
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[dev]) {
      e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return noalloc_fail(PT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
      g_scratch_ptr[dev] = ptr;
    }
    ptr = g_scratch_ptr[dev];
  }
  if (pt_radix_sort(keys, sorted_keys, perm, n, 0, 64, ptr, s) != 0) {
    e = hipGetLastError();
    return noalloc_fail(PT_ERR_HIP, std::string("pt_sort_noalloc: launch failed: ") +
                                        hipGetErrorString(e == hipSuccess ? hipErrorUnknown : e));
  }
  return PT_OK;
}

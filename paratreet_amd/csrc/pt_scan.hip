// Device-wide exclusive prefix sums (uint32 and int64), used by the radix sort
// digit tables, the octree level compaction and the bucket/list offsets.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename TI, typename TO>
__global__ void __launch_bounds__(kScanThreads)
scan_tile_kernel(const TI* __restrict__ in, TO* __restrict__ out, int64_t n,
                 TO* __restrict__ tile_sums) {
  __shared__ TO s_wave[kScanThreads / PT_WAVE];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  TO v[kScanItems];
  TO run = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    const int64_t i = base + j;
    const TO a = i < n ? (TO)in[i] : (TO)0;
    v[j] = run;
    run += a;
  }
  // inclusive scan of per-thread totals across the block
  TO t = run;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int wave = threadIdx.x / PT_WAVE;
#pragma unroll
  for (int d = 1; d < PT_WAVE; d <<= 1) {
    const TO o = __shfl_up(t, d, PT_WAVE);
    if (lane >= d) t += o;
  }
  if (lane == PT_WAVE - 1) s_wave[wave] = t;
  __syncthreads();
  TO wave_off = 0;
  for (int w = 0; w < wave; w++) wave_off += s_wave[w];
  const TO excl = wave_off + t - run;
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    const int64_t i = base + j;
    if (i < n) out[i] = excl + v[j];
  }
  if (tile_sums && threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = excl + run;
}

template <typename TO>
__global__ void __launch_bounds__(kScanThreads)
scan_add_kernel(TO* __restrict__ out, int64_t n, const TO* __restrict__ tile_offs) {
  const TO add = tile_offs[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int j = threadIdx.x; j < kScanTile; j += kScanThreads) {
    const int64_t i = base + j;
    if (i < n) out[i] += add;
  }
}

template <typename TO>
__global__ void zero_kernel(TO* p) {
  if (threadIdx.x == 0) *p = 0;
}

template <typename TO>
int64_t scratch_elems(int64_t n) {
  int64_t total = 0;
  while (n > kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    total += n;
  }
  return total;
}

template <typename TI, typename TO>
int scan_impl(const TI* in, TO* out, int64_t n, TO* scratch, TO* total, hipStream_t s) {
  if (n <= 0) {
    if (total) zero_kernel<TO><<<1, 64, 0, s>>>(total);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  if (nt == 1) {
    scan_tile_kernel<TI, TO><<<1, kScanThreads, 0, s>>>(in, out, n, total);
  } else {
    TO* sums = scratch;
    scan_tile_kernel<TI, TO><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums);
    if (scan_impl<TO, TO>(sums, sums, nt, scratch + nt, total, s) != 0) return -1;
    scan_add_kernel<TO><<<(unsigned)nt, kScanThreads, 0, s>>>(out, n, sums);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int64_t pt_scan_scratch_bytes(int64_t n) {
  return (scratch_elems<int64_t>(n) + 16) * (int64_t)sizeof(int64_t);
}

int pt_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* scratch,
                          uint32_t* total, hipStream_t s) {
  return scan_impl<uint32_t, uint32_t>(in, out, n, (uint32_t*)scratch, total, s);
}

int pt_exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, void* scratch,
                                 int64_t* total, hipStream_t s) {
  return scan_impl<int32_t, int64_t>(in, out, n, (int64_t*)scratch, total, s);
}

// Stable LSD radix sort of 64-bit keys with 32-bit payloads.
//
// Replaces the std::sort of Subtree::buildTree (src/Subtree.h:281) and
// Reader::localSort (src/Reader.C:237-241).  The reference's std::sort is
// unstable; this sort is stable, so equal keys keep ascending original index,
// which is the order the oracle defines.
//
// Per 8-bit digit pass:
//   1. count_kernel   -- LDS histogram of each 2048-key tile, written
//                        digit-major so one exclusive scan yields every
//                        (digit, tile) output offset;
//   2. device-wide exclusive scan of the 256 x tiles table;
//   3. scatter_kernel -- each wave ranks its 512-key sub-tile in 8 rounds of
//                        64 with an 8-ballot digit match and per-wave LDS
//                        counters; ranks become tile-local positions, keys are
//                        staged through LDS and written out in digit runs so
//                        the HBM writes are contiguous.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / PT_WAVE;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per block
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortThreads)
count_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift, int ntiles,
             uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[kRadix];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; j++) {
    const int64_t i = base + (int64_t)j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&hist[(unsigned)(keys[i] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// FIRST: the payload is the key's input index (vals_in unused).
template <bool FIRST>
__global__ void __launch_bounds__(kSortThreads)
scatter_kernel(const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
               int64_t n, int shift, int ntiles, const uint32_t* __restrict__ offsets,
               uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t wave_run[kSortWaves][kRadix];
  __shared__ uint32_t tile_off[kRadix];
  __shared__ uint64_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];

  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int wave = threadIdx.x / PT_WAVE;
  for (int i = threadIdx.x; i < kSortWaves * kRadix; i += kSortThreads)
    (&wave_run[0][0])[i] = 0;
  __syncthreads();

  const int64_t tile_base = (int64_t)blockIdx.x * kSortTile;
  const int64_t wave_base = tile_base + (int64_t)wave * (kSortItems * PT_WAVE);
  uint64_t k[kSortItems];
  uint32_t v[kSortItems];
  uint32_t rank[kSortItems];
  const unsigned long long lt_mask = (1ull << lane) - 1ull;

#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    const int64_t i = wave_base + (int64_t)r * PT_WAVE + lane;
    const bool valid = i < n;
    k[r] = valid ? keys_in[i] : ~0ull;
    v[r] = valid ? (FIRST ? (uint32_t)i : vals_in[i]) : 0u;
    const uint32_t d = (uint32_t)(k[r] >> shift) & (kRadix - 1);
    // lanes holding the same digit (invalid lanes never match a valid one:
    // they are excluded from the initial mask)
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t before = wave_run[wave][d];
    const uint32_t rk = (uint32_t)__popcll(peers & lt_mask);
    rank[r] = before + rk;
    // the lowest peer lane publishes the new running count for this digit
    if (valid && rk == 0) wave_run[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // per-digit prefix over waves and tile-local digit offsets
  {
    const int d = threadIdx.x;  // kSortThreads == kRadix
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
      const uint32_t c = wave_run[w][d];
      wave_run[w][d] = acc;
      acc += c;
    }
    // exclusive scan of the tile's digit totals
    uint32_t t = acc;
#pragma unroll
    for (int dd = 1; dd < PT_WAVE; dd <<= 1) {
      const uint32_t o = __shfl_up(t, dd, PT_WAVE);
      if (lane >= dd) t += o;
    }
    __shared__ uint32_t s_wsum[kSortWaves];
    if (lane == PT_WAVE - 1) s_wsum[wave] = t;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wave; w++) woff += s_wsum[w];
    tile_off[d] = woff + t - acc;
  }
  __syncthreads();

#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    const int64_t i = wave_base + (int64_t)r * PT_WAVE + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(k[r] >> shift) & (kRadix - 1);
      const uint32_t pos = tile_off[d] + wave_run[wave][d] + rank[r];
      s_keys[pos] = k[r];
      s_vals[pos] = v[r];
    }
  }
  __syncthreads();

  const int64_t tile_n = (n - tile_base) < kSortTile ? (n - tile_base) : kSortTile;
  for (int j = threadIdx.x; j < tile_n; j += kSortThreads) {
    const uint64_t kk = s_keys[j];
    const uint32_t d = (uint32_t)(kk >> shift) & (kRadix - 1);
    const int64_t dst = (int64_t)offsets[(int64_t)d * ntiles + blockIdx.x] + (j - tile_off[d]);
    keys_out[dst] = kk;
    vals_out[dst] = s_vals[j];
  }
}

__global__ void __launch_bounds__(256)
iota_kernel(uint32_t* __restrict__ v, int64_t n) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

}  // namespace

int64_t pt_sort_scratch_bytes(int64_t n) {
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const int64_t table = ntiles * kRadix;
  // counts table + scan scratch + ping-pong key/value buffers
  return table * 4 + pt_scan_scratch_bytes(table) + n * 8 + n * 4 + 1024;
}

// Sorts src (n keys, read only) into (keys, vals): vals = input index of
// each sorted key.  Scratch: pt_sort_scratch_bytes(n) bytes; begin_bit /
// end_bit bound the digit passes.
int pt_radix_sort(const uint64_t* src, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                  int end_bit, void* scratch, hipStream_t s) {
  if (n <= 0) return 0;
  if (n == 1 || end_bit <= begin_bit) {
    if (n > 1) return -1;
    if (hipMemcpyAsync(keys, src, 8, hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
    iota_kernel<<<1, 256, 0, s>>>(vals, 1);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  if (n > 0x7fffffffll) return -1;
  const int ntiles = (int)((n + kSortTile - 1) / kSortTile);
  const int64_t table = (int64_t)ntiles * kRadix;
  char* p = (char*)scratch;
  uint32_t* counts = (uint32_t*)p;
  p += ((table * 4 + 255) / 256) * 256;
  void* scan_scratch = p;
  p += ((pt_scan_scratch_bytes(table) + 255) / 256) * 256;
  uint64_t* k2 = (uint64_t*)p;
  p += ((n * 8 + 255) / 256) * 256;
  uint32_t* v2 = (uint32_t*)p;

  // pass p writes (keys, vals) when an even number of passes follow it, the
  // scratch pair otherwise, so the last pass lands in (keys, vals)
  const int passes = (end_bit - begin_bit + 7) / 8;
  const uint64_t* kin = src;
  const uint32_t* vin = nullptr;
  for (int p = 0; p < passes; p++) {
    const int shift = begin_bit + 8 * p;
    const bool to_final = ((passes - 1 - p) % 2) == 0;
    uint64_t* kout = to_final ? keys : k2;
    uint32_t* vout = to_final ? vals : v2;
    count_kernel<<<ntiles, kSortThreads, 0, s>>>(kin, n, shift, ntiles, counts);
    if (pt_exclusive_scan_u32(counts, counts, table, scan_scratch, nullptr, s) != 0) return -1;
    if (p == 0)
      scatter_kernel<true><<<ntiles, kSortThreads, 0, s>>>(kin, vin, n, shift, ntiles, counts, kout,
                                                           vout);
    else
      scatter_kernel<false><<<ntiles, kSortThreads, 0, s>>>(kin, vin, n, shift, ntiles, counts,
                                                            kout, vout);
    kin = kout;
    vin = vout;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_iota(uint32_t* v, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  iota_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(v, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Universe box and Morton keys.
//
// Reference: Driver::remakeUniverse (src/Driver.h:67-80) over the particle
// bounding box (src/BoundingBox.C:29-44), and Decomposition::assignKeys
// (src/Decomposition.C:58-64) = SFC::generateKey + placeholder bit 63.
// The key definition (from the unvendored utility SFC.h) used here:
// u = (x - lesser) / (greater - lesser) in FP64, q = clamp(floor(u * 2^21)),
// x the most significant bit of each triple.  Compiled with
// -ffp-contract=off so the universe and the quantisation are bit-exact with a
// separately rounded CPU evaluation.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kBoxThreads = 256;

__global__ void __launch_bounds__(kBoxThreads)
bbox_partial_kernel(const double* __restrict__ x, const double* __restrict__ y,
                    const double* __restrict__ z, int64_t n, double* __restrict__ part) {
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)kBoxThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBoxThreads) {
    const double p[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = p[k] < lo[k] ? p[k] : lo[k];
      hi[k] = p[k] > hi[k] ? p[k] : hi[k];
    }
  }
  __shared__ double s[6][kBoxThreads];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = lo[k];
    s[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double a = s[k][threadIdx.x + w];
        if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
        const double b = s[3 + k][threadIdx.x + w];
        if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// Reduces the per-block boxes and applies remakeUniverse: a cube centred on
// the bbox centre with half-side fEps * 0.5 * (largest extent).
__global__ void __launch_bounds__(kBoxThreads)
universe_kernel(const double* __restrict__ part, int nparts, double* __restrict__ bbox_out,
                double* __restrict__ box) {
  __shared__ double s[6][kBoxThreads];
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nparts; b += kBoxThreads) {
    for (int k = 0; k < 3; k++) {
      const double a = part[b * 6 + k];
      const double c = part[b * 6 + 3 + k];
      lo[k] = a < lo[k] ? a : lo[k];
      hi[k] = c > hi[k] ? c : hi[k];
    }
  }
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = lo[k];
    s[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int k = 0; k < 3; k++) {
        const double a = s[k][threadIdx.x + w];
        if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
        const double c = s[3 + k][threadIdx.x + w];
        if (c > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = c;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double bs[3], bc[3];
  for (int k = 0; k < 3; k++) {
    const double l = s[k][0], h = s[3 + k][0];
    bbox_out[k] = l;
    bbox_out[3 + k] = h;
    bs[k] = h - l;
    bc[k] = (l + h) / 2.0;
  }
  double mx = (bs[0] > bs[1]) ? bs[0] : bs[1];
  mx = (mx > bs[2]) ? mx : bs[2];
  const double fEps = 1.0 + 1.91e-6;  // src/Driver.h:72-73
  const double half = fEps * 0.5 * mx;
  for (int k = 0; k < 3; k++) {
    box[k] = bc[k] - half;
    box[3 + k] = bc[k] + half;
  }
}

__device__ __forceinline__ uint64_t quantise(double v, double lo, double hi) {
  const double u = (v - lo) / (hi - lo);
  const double s = floor(u * 2097152.0);
  if (!(s > 0.0)) return 0;
  if (s >= 2097151.0) return 2097151u;
  return (uint64_t)s;
}

// Spreads the low 21 bits of v to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x001f00000000ffffull;
  v = (v | (v << 16)) & 0x001f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__global__ void __launch_bounds__(256)
keys_kernel(const double* __restrict__ x, const double* __restrict__ y,
            const double* __restrict__ z, const double* __restrict__ m,
            const double* __restrict__ e, int64_t n, const double* __restrict__ box,
            uint64_t* __restrict__ keys, ParticleRec* __restrict__ rec) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const double px = x[i], py = y[i], pz = z[i];
  const uint64_t qx = quantise(px, box[0], box[3]);
  const uint64_t qy = quantise(py, box[1], box[4]);
  const uint64_t qz = quantise(pz, box[2], box[5]);
  const uint64_t k = (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
  keys[i] = k | (1ull << (PT_KEY_BITS - 1));
  ParticleRec r;
  r.x = px;
  r.y = py;
  r.z = pz;
  r.m = m[i];
  r.e = e[i];
  r.pad[0] = r.pad[1] = r.pad[2] = 0.0;
  rec[i] = r;
}

}  // namespace

int pt_launch_universe(const double* x, const double* y, const double* z, int64_t n,
                       double* scratch_parts, int max_parts, double* bbox_out, double* box,
                       hipStream_t s) {
  int nb = (int)((n + kBoxThreads - 1) / kBoxThreads);
  if (nb > max_parts) nb = max_parts;
  if (nb < 1) nb = 1;
  bbox_partial_kernel<<<nb, kBoxThreads, 0, s>>>(x, y, z, n, scratch_parts);
  universe_kernel<<<1, kBoxThreads, 0, s>>>(scratch_parts, nb, bbox_out, box);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_keys(const double* x, const double* y, const double* z, const double* m,
                   const double* e, int64_t n, const double* box, uint64_t* keys,
                   ParticleRec* rec, hipStream_t s) {
  if (n <= 0) return 0;
  const int64_t nb = (n + 255) / 256;
  keys_kernel<<<(unsigned)nb, 256, 0, s>>>(x, y, z, m, e, n, box, keys, rec);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Particle permutation, leapfrog, direct summation and fill kernels.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// Sorted copies of the particle arrays (the reference sorts whole Particle
// structs, src/Subtree.h:281; here only the fields the path reads move).
__global__ void __launch_bounds__(256)
gather_kernel(const uint32_t* __restrict__ perm, int64_t n, const ParticleRec* __restrict__ rec,
              double* __restrict__ xs, double* __restrict__ ys, double* __restrict__ zs,
              double* __restrict__ ms, double* __restrict__ es) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const ParticleRec r = rec[perm[i]];
  xs[i] = r.x;
  ys[i] = r.y;
  zs[i] = r.z;
  ms[i] = r.m;
  es[i] = r.e;
}

__global__ void __launch_bounds__(256)
unsort_kernel(const uint32_t* __restrict__ perm, int64_t n, const double* __restrict__ a0,
              const double* __restrict__ a1, const double* __restrict__ a2,
              const double* __restrict__ a3, double* __restrict__ o0, double* __restrict__ o1,
              double* __restrict__ o2, double* __restrict__ o3) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  o0[j] = a0[i];
  o1[j] = a1[i];
  o2[j] = a2[i];
  o3[j] = a3[i];
}

// Particle::kick then Particle::perturb (src/Particle.C:7-21): v += a dt/2
// twice, then x += v dt.
__global__ void __launch_bounds__(256)
kick_drift_kernel(int64_t n, double dt, double* __restrict__ x, double* __restrict__ y,
                  double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
                  double* __restrict__ vz, const double* __restrict__ ax,
                  const double* __restrict__ ay, const double* __restrict__ az) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  double v0 = vx[i], v1 = vy[i], v2 = vz[i];
  const double a0 = ax[i], a1 = ay[i], a2 = az[i];
  v0 += a0 * dt / 2;
  v1 += a1 * dt / 2;
  v2 += a2 * dt / 2;
  v0 += (a0 * dt / 2);
  v1 += (a1 * dt / 2);
  v2 += (a2 * dt / 2);
  x[i] += (v0 * dt);
  y[i] += (v1 * dt);
  z[i] += (v2 * dt);
  vx[i] = v0;
  vy[i] = v1;
  vz[i] = v2;
}

__device__ __forceinline__ double spline_b(double r2, double twoh) {
  const double r = sqrt(r2);
  if (r < twoh) {
    const double dih = 2.0 / twoh;
    const double u = r * dih;
    if (u < 1.0) {
      return dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
    }
    const double dir = 1.0 / r;
    return -1.0 / 15.0 * dir * dir * dir +
           dih * dih * dih * (8.0 / 3.0 - 3.0 * u + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
  }
  const double a = 1.0 / r;
  return a * a * a;
}

// FP64 direct sum with the leaf() kernel: one block per target.
__global__ void __launch_bounds__(256)
direct_kernel(const int64_t* __restrict__ targets, int64_t nt, const double* __restrict__ x,
              const double* __restrict__ y, const double* __restrict__ z,
              const double* __restrict__ m, const double* __restrict__ e, int64_t n,
              double* __restrict__ out) {
  __shared__ double red[3][256 / PT_WAVE];
  const int64_t ti = blockIdx.x;
  if (ti >= nt) return;
  const int64_t i = targets[ti];
  const double tx = x[i], ty = y[i], tz = z[i], te = e[i];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    const double dx = x[j] - tx, dy = y[j] - ty, dz = z[j] - tz;
    const double rsq = dx * dx + dy * dy + dz * dz;
    if (rsq != 0) {
      const double f = spline_b(rsq, e[j] + te) * m[j];
      sx += dx * f;
      sy += dy * f;
      sz += dz * f;
    }
  }
  sx = pt_wave_sum(sx);
  sy = pt_wave_sum(sy);
  sz = pt_wave_sum(sz);
  const int wave = threadIdx.x / PT_WAVE;
  if ((threadIdx.x & (PT_WAVE - 1)) == 0) {
    red[0][wave] = sx;
    red[1][wave] = sy;
    red[2][wave] = sz;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int w = 0; w < 256 / PT_WAVE; w++) s += red[threadIdx.x][w];
    out[ti * 3 + threadIdx.x] = s;
  }
}

__global__ void __launch_bounds__(256)
fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256)
fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

int pt_launch_gather(const uint32_t* perm, int64_t n, const ParticleRec* rec, double* xs,
                     double* ys, double* zs, double* ms, double* es, hipStream_t s) {
  if (n <= 0) return 0;
  gather_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, rec, xs, ys, zs, ms, es);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_unsort(const uint32_t* perm, int64_t n, const double* a0, const double* a1,
                     const double* a2, const double* a3, double* o0, double* o1, double* o2,
                     double* o3, hipStream_t s) {
  if (n <= 0) return 0;
  unsort_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, a0, a1, a2, a3, o0, o1, o2, o3);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_kick_drift(int64_t n, double dt, double* x, double* y, double* z, double* vx,
                         double* vy, double* vz, const double* ax, const double* ay,
                         const double* az, hipStream_t s) {
  if (n <= 0) return 0;
  kick_drift_kernel<<<grid_for(n), 256, 0, s>>>(n, dt, x, y, z, vx, vy, vz, ax, ay, az);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_direct(const int64_t* targets, int64_t nt, const double* x, const double* y,
                     const double* z, const double* m, const double* e, int64_t n, double* out,
                     hipStream_t s) {
  if (nt <= 0) return 0;
  direct_kernel<<<(unsigned)nt, 256, 0, s>>>(targets, nt, x, y, z, m, e, n, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return 0;
  fill_u32_kernel<<<grid_for(n), 256, 0, s>>>(p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n <= 0) return 0;
  fill_u64_kernel<<<grid_for(n), 256, 0, s>>>(p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

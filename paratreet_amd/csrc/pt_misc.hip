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

// Sorted-order results to the caller's order in two passes: one 32-byte
// record per particle scattered to its original slot (a whole sector per
// write instead of four 8-byte pieces of four lines), then the records split
// into the four arrays with coalesced reads and writes.
struct __align__(32) Res4 {
  double a0, a1, a2, a3;
};

__global__ void __launch_bounds__(256)
unsort_kernel(const uint32_t* __restrict__ perm, int64_t n, const double* __restrict__ a0,
              const double* __restrict__ a1, const double* __restrict__ a2,
              const double* __restrict__ a3, Res4* __restrict__ stage) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  stage[perm[i]] = Res4{a0[i], a1[i], a2[i], a3[i]};
}

__global__ void __launch_bounds__(256)
split4_kernel(const Res4* __restrict__ stage, int64_t n, double* __restrict__ o0,
              double* __restrict__ o1, double* __restrict__ o2, double* __restrict__ o3) {
  const int64_t j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= n) return;
  const Res4 r = stage[j];
  o0[j] = r.a0;
  o1[j] = r.a1;
  o2[j] = r.a2;
  o3[j] = r.a3;
}

// Particle::kick and Particle::perturb (src/Particle.C:7-21): the kick adds
// a dt/2 to the velocity, the perturb adds a dt/2 again and moves the
// position by v dt.  mode: PT_STEP_KICK_DRIFT both in sequence, PT_STEP_KICK
// the kick alone, PT_STEP_PERTURB the perturb alone (the first two in turn
// give the same bits as the combined step).
__global__ void __launch_bounds__(256)
kick_drift_kernel(int64_t n, double dt, int mode, double* __restrict__ x,
                  double* __restrict__ y, double* __restrict__ z, double* __restrict__ vx,
                  double* __restrict__ vy, double* __restrict__ vz,
                  const double* __restrict__ ax, const double* __restrict__ ay,
                  const double* __restrict__ az) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  double v0 = vx[i], v1 = vy[i], v2 = vz[i];
  const double a0 = ax[i], a1 = ay[i], a2 = az[i];
  if (mode != PT_STEP_PERTURB) {
    v0 += a0 * dt / 2;
    v1 += a1 * dt / 2;
    v2 += a2 * dt / 2;
  }
  if (mode != PT_STEP_KICK) {
    v0 += (a0 * dt / 2);
    v1 += (a1 * dt / 2);
    v2 += (a2 * dt / 2);
    x[i] += (v0 * dt);
    y[i] += (v1 * dt);
    z[i] += (v2 * dt);
  }
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

// Smallest value of v[0..n): per-block minima, then one block over them.
__global__ void __launch_bounds__(256)
min_partial_kernel(const double* __restrict__ v, int64_t n, double* __restrict__ part) {
  __shared__ double s[256];
  double lo = INFINITY;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    lo = v[i] < lo ? v[i] : lo;
  s[threadIdx.x] = lo;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && s[threadIdx.x + w] < s[threadIdx.x])
      s[threadIdx.x] = s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256)
min_final_kernel(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  __shared__ double s[256];
  double lo = INFINITY;
  for (int b = threadIdx.x; b < nparts; b += 256) lo = part[b] < lo ? part[b] : lo;
  s[threadIdx.x] = lo;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && s[threadIdx.x + w] < s[threadIdx.x])
      s[threadIdx.x] = s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ void __launch_bounds__(256)
fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Ewald sum of the periodic build, quadrupole form: LeafEwaldFn with
// EwaldData::EwaldInit (examples/Ewald.h, examples/Ewald.C, non-HEXADECAPOLE
// branches).  After the replica traversals every particle gets the rest of
// the lattice of the root's monopole + quadrupole (mass M, centre cm, raw
// second moments q): erfc-damped real-space terms over the lattice vectors
// within 2.6 L (the explicit replicas -- the hole -- contribute -erf, which
// removes the bare 1/r part the tree summed; a two-term series replaces the
// closed forms below r^2 = 1.2e-3 L^2) and the Fourier terms of the integer
// vectors 0 < |h| <= 2.8, with alpha = 2 / L, L = dxPeriod.  Same
// operations, in the same order, as the oracle's restatement.
// ---------------------------------------------------------------------------

constexpr double kEwAlphaFac = 2.0;  // alpha = 2 / L
constexpr double kEwCut = 2.6;       // real-space radius, in L
constexpr double kEwHCut = 2.8;      // Fourier radius, in units of 2 pi / L
constexpr double kPi = 3.14159265358979323846;

// One thread: the Fourier table {2 pi h / L, C(h)} in the lattice order of the
// loops (h_x outermost); C(h) = -(e^{-k4 h^2}/(pi h^2 L)) (M - (2 pi/L)^2 h.q.h/2)
// (the sine coefficients vanish at quadrupole order).
__global__ void ewald_table_kernel(const double* __restrict__ nd, int64_t nn, double L,
                                   double* __restrict__ table, int32_t* __restrict__ count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double mass = nd[PT_CD_MASS * nn];
  double q[6];
  for (int k = 0; k < 6; k++) q[k] = nd[(PT_CD_Q + k) * nn];
  const double alpha = kEwAlphaFac / L, alpha2 = alpha * alpha;
  const double k4 = kPi * kPi / (alpha2 * L * L);
  const int hr = (int)ceil(kEwHCut);
  int nh = 0;
  for (int hx = -hr; hx <= hr; hx++)
    for (int hy = -hr; hy <= hr; hy++)
      for (int hz = -hr; hz <= hr; hz++) {
        const int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > kEwHCut * kEwHCut) continue;
        const double g0 = exp(-k4 * h2) / (kPi * h2 * L);
        const double w = 2.0 * kPi / L;
        const double hqh = 0.5 * (hx * (q[0] * hx + q[1] * hy + q[2] * hz) +
                                  hy * (q[1] * hx + q[3] * hy + q[4] * hz) +
                                  hz * (q[2] * hx + q[4] * hy + q[5] * hz));
        table[4 * nh + 0] = w * hx;
        table[4 * nh + 1] = w * hy;
        table[4 * nh + 2] = w * hz;
        table[4 * nh + 3] = -(g0 * mass - w * w * g0 * hqh);
        nh++;
      }
  *count = nh;
}

// One thread per particle (sorted order); adds to the accelerations and
// potentials the replica traversals left.
__global__ void __launch_bounds__(256)
ewald_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
             const double* __restrict__ z, const double* __restrict__ nd, int64_t nn, double L,
             int n_replicas, const double* __restrict__ table, const int32_t* __restrict__ count,
             double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
             double* __restrict__ pot) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const double mass = nd[PT_CD_MASS * nn];
  const double cm[3] = {nd[PT_CD_CM * nn], nd[(PT_CD_CM + 1) * nn], nd[(PT_CD_CM + 2) * nn]};
  double q[6];
  for (int k = 0; k < 6; k++) q[k] = nd[(PT_CD_Q + k) * nn];
  const double alpha = kEwAlphaFac / L, alpha2 = alpha * alpha;
  const double k1 = kPi / (alpha2 * L * L * L);
  const double ka = 2.0 * alpha / sqrt(kPi);
  const double q2 = 0.5 * (q[0] + q[3] + q[5]);
  int nrep = (int)ceil(kEwCut);
  if (n_replicas > nrep) nrep = n_replicas;
  const double cut2 = kEwCut * kEwCut * L * L, inner2 = 1.2e-3 * L * L;
  double p = mass * k1, fx = 0.0, fy = 0.0, fz = 0.0;
  const double dx = x[i] - cm[0], dy = y[i] - cm[1], dz = z[i] - cm[2];
  for (int ix = -nrep; ix <= nrep; ix++) {
    const bool hole_x = ix >= -n_replicas && ix <= n_replicas;
    const double rx = dx + ix * L;
    for (int iy = -nrep; iy <= nrep; iy++) {
      const bool hole_xy = hole_x && iy >= -n_replicas && iy <= n_replicas;
      const double ry = dy + iy * L;
      for (int iz = -nrep; iz <= nrep; iz++) {
        const bool hole = hole_xy && iz >= -n_replicas && iz <= n_replicas;
        const double rz = dz + iz * L;
        const double r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > cut2 && !hole) continue;
        double g0, g1, g2, g3;
        if (r2 < inner2) {
          const double s2 = r2 * alpha2;
          double an = ka;
          g0 = an * ((1.0 / 3.0) * s2 - 1.0);
          an *= 2.0 * alpha2;
          g1 = an * ((1.0 / 5.0) * s2 - (1.0 / 3.0));
          an *= 2.0 * alpha2;
          g2 = an * ((1.0 / 7.0) * s2 - (1.0 / 5.0));
          an *= 2.0 * alpha2;
          g3 = an * ((1.0 / 9.0) * s2 - (1.0 / 7.0));
        } else {
          const double r = sqrt(r2);
          const double ir2 = 1.0 / r2;
          const double a = exp(-r2 * alpha2) * ka * ir2;
          g0 = (hole ? -erf(alpha * r) : erfc(alpha * r)) / r;
          g1 = g0 * ir2 + a;
          g2 = 3.0 * g1 * ir2 + 2.0 * alpha2 * a;
          g3 = 5.0 * g2 * ir2 + 4.0 * alpha2 * alpha2 * a;
        }
        const double qx = q[0] * rx + q[1] * ry + q[2] * rz;
        const double qy = q[1] * rx + q[3] * ry + q[4] * rz;
        const double qz = q[2] * rx + q[4] * ry + q[5] * rz;
        const double qrr = 0.5 * (qx * rx + qy * ry + qz * rz);
        const double qta = g1 * mass - g2 * q2 + g3 * qrr;
        p -= g0 * mass - g1 * q2 + g2 * qrr;
        fx += g2 * qx - rx * qta;
        fy += g2 * qy - ry * qta;
        fz += g2 * qz - rz * qta;
      }
    }
  }
  const int nh = *count;
  for (int k = 0; k < nh; k++) {
    const double hd = table[4 * k] * dx + table[4 * k + 1] * dy + table[4 * k + 2] * dz;
    const double c = cos(hd), sn = sin(hd), cf = table[4 * k + 3];
    p += cf * c;
    fx += table[4 * k] * (cf * sn);
    fy += table[4 * k + 1] * (cf * sn);
    fz += table[4 * k + 2] * (cf * sn);
  }
  ax[i] += fx;
  ay[i] += fy;
  az[i] += fz;
  pot[i] += p;
}

// ---------------------------------------------------------------------------
// Ewald sum of the periodic HEXADECAPOLE build: the HEXADECAPOLE branches of
// LeafEwaldFn and EwaldData::EwaldInit (examples/Ewald.h:6-13,31-42,57-67,
// 127-166; examples/Ewald.C:11-51,89-158).  The root's reduced moments to
// fourth order (FMOMR order, completed with their z components by
// tracelessness as momFmomr2Momc does, examples/moments.C:1563-1597) replace
// the quadrupole: g0..g5 in real space, and in Fourier space a cosine
// coefficient -g (m - w^2 M2[h^2]/2 + w^4 M4[h^4]/24) and a sine coefficient
// g w^3 M3[h^3]/6 per vector (w = 2 pi / L, g = e^{-k4 h^2}/(pi h^2 L)).
// Same operations, in the same order, as the oracle's restatement.
// ---------------------------------------------------------------------------

struct Momc {
  double m;
  double xx, yy, xy, xz, yz, zz;
  double xxx, xyy, xxy, yyy, xxz, yyz, xyz, xzz, yzz, zzz;
  double xxxx, xyyy, xxxy, yyyy, xxxz, yyyz, xxyy, xxyz, xyyz, xxzz, xyzz, xzzz, yyzz, yzzz, zzzz;
};

__device__ __forceinline__ void momc_from_reduced(double mass, const double* __restrict__ h,
                                                  Momc& c) {
  c.m = mass;
  c.xx = h[0]; c.yy = h[1]; c.xy = h[2]; c.xz = h[3]; c.yz = h[4];
  c.xxx = h[5]; c.xyy = h[6]; c.xxy = h[7]; c.yyy = h[8]; c.xxz = h[9]; c.yyz = h[10];
  c.xyz = h[11];
  c.xxxx = h[12]; c.xyyy = h[13]; c.xxxy = h[14]; c.yyyy = h[15]; c.xxxz = h[16];
  c.yyyz = h[17]; c.xxyy = h[18]; c.xxyz = h[19]; c.xyyz = h[20];
  c.zz = -(c.xx + c.yy);
  c.xzz = -(c.xxx + c.xyy);
  c.yzz = -(c.xxy + c.yyy);
  c.zzz = -(c.xxz + c.yyz);
  c.xxzz = -(c.xxxx + c.xxyy);
  c.xyzz = -(c.xxxy + c.xyyy);
  c.xzzz = -(c.xxxz + c.xyyz);
  c.yyzz = -(c.xxyy + c.yyyy);
  c.yzzz = -(c.xxyz + c.yyyz);
  c.zzzz = -(c.xxzz + c.yyzz);
}

// M2.v, M3[., v, v] / 2 and M4[., v, v, v] / 6.
__device__ __forceinline__ void momc_contract(const Momc& M, double x, double y, double z,
                                              double q2[3], double q3[3], double q4[3]) {
  q2[0] = M.xx * x + M.xy * y + M.xz * z;
  q2[1] = M.xy * x + M.yy * y + M.yz * z;
  q2[2] = M.xz * x + M.yz * y + M.zz * z;
  const double xx = 0.5 * x * x, yy = 0.5 * y * y, zz = 0.5 * z * z;
  const double xy = x * y, xz = x * z, yz = y * z;
  q3[0] = M.xxx * xx + M.xxy * xy + M.xxz * xz + M.xyy * yy + M.xyz * yz + M.xzz * zz;
  q3[1] = M.xxy * xx + M.xyy * xy + M.xyz * xz + M.yyy * yy + M.yyz * yz + M.yzz * zz;
  q3[2] = M.xxz * xx + M.xyz * xy + M.xzz * xz + M.yyz * yy + M.yzz * yz + M.zzz * zz;
  const double onethird = 1.0 / 3.0;
  const double xxx = onethird * xx * x, xxy = xx * y, xxz = xx * z, xyy = yy * x,
               yyy = onethird * yy * y;
  const double yyz = yy * z, xzz = zz * x, yzz = zz * y, zzz = onethird * zz * z, xyz = xy * z;
  q4[0] = M.xxxx * xxx + M.xxxy * xxy + M.xxxz * xxz + M.xxyy * xyy + M.xxyz * xyz +
          M.xxzz * xzz + M.xyyy * yyy + M.xyyz * yyz + M.xyzz * yzz + M.xzzz * zzz;
  q4[1] = M.xxxy * xxx + M.xxyy * xxy + M.xxyz * xxz + M.xyyy * xyy + M.xyyz * xyz +
          M.xyzz * xzz + M.yyyy * yyy + M.yyyz * yyz + M.yyzz * yzz + M.yzzz * zzz;
  q4[2] = M.xxxz * xxx + M.xxyz * xxy + M.xxzz * xxz + M.xyyz * xyy + M.xyzz * xyz +
          M.xzzz * xzz + M.yyyz * yyy + M.yyzz * yyz + M.yzzz * yzz + M.zzzz * zzz;
}

// One thread: the Fourier table {2 pi h / L, C(h), S(h)} in the lattice order
// of the loops (h_x outermost).
__global__ void ewald_hexa_table_kernel(const double* __restrict__ nd, int64_t nn,
                                        const HexNode* __restrict__ hn, double L,
                                        double* __restrict__ table, int32_t* __restrict__ count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double mass = nd[PT_CD_MASS * nn];
  Momc M;
  momc_from_reduced(mass, hn[0].h, M);
  const double alpha = kEwAlphaFac / L, alpha2 = alpha * alpha;
  const double k4 = kPi * kPi / (alpha2 * L * L);
  const double w = 2.0 * kPi / L;
  const int hr = (int)ceil(kEwHCut);
  int nh = 0;
  for (int hx = -hr; hx <= hr; hx++)
    for (int hy = -hr; hy <= hr; hy++)
      for (int hz = -hr; hz <= hr; hz++) {
        const int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > kEwHCut * kEwHCut) continue;
        const double g = exp(-k4 * h2) / (kPi * h2 * L);
        double q2[3], q3[3], q4[3];
        momc_contract(M, hx, hy, hz, q2, q3, q4);
        const double m2 = 0.5 * (q2[0] * hx + q2[1] * hy + q2[2] * hz);
        const double m3 = (1.0 / 3.0) * (q3[0] * hx + q3[1] * hy + q3[2] * hz);
        const double m4 = 0.25 * (q4[0] * hx + q4[1] * hy + q4[2] * hz);
        table[5 * nh + 0] = w * hx;
        table[5 * nh + 1] = w * hy;
        table[5 * nh + 2] = w * hz;
        table[5 * nh + 3] = -g * (mass - w * w * m2 + w * w * w * w * m4);
        table[5 * nh + 4] = g * w * w * w * m3;
        nh++;
      }
  *count = nh;
}

// One thread per particle (sorted order); adds to the accelerations and
// potentials the replica traversals left.
__global__ void __launch_bounds__(256)
ewald_hexa_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                  const double* __restrict__ z, const double* __restrict__ nd, int64_t nn,
                  const HexNode* __restrict__ hn, double L, int n_replicas,
                  const double* __restrict__ table, const int32_t* __restrict__ count,
                  double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
                  double* __restrict__ pot) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const double mass = nd[PT_CD_MASS * nn];
  const double cm[3] = {nd[PT_CD_CM * nn], nd[(PT_CD_CM + 1) * nn], nd[(PT_CD_CM + 2) * nn]};
  Momc M;
  momc_from_reduced(mass, hn[0].h, M);
  const double Q4xx = 0.5 * (M.xxxx + M.xxyy + M.xxzz), Q4xy = 0.5 * (M.xxxy + M.xyyy + M.xyzz);
  const double Q4xz = 0.5 * (M.xxxz + M.xyyz + M.xzzz), Q4yy = 0.5 * (M.xxyy + M.yyyy + M.yyzz);
  const double Q4yz = 0.5 * (M.xxyz + M.yyyz + M.yzzz), Q4zz = 0.5 * (M.xxzz + M.yyzz + M.zzzz);
  const double Q4 = 0.25 * (Q4xx + Q4yy + Q4zz);
  const double Q3x = 0.5 * (M.xxx + M.xyy + M.xzz), Q3y = 0.5 * (M.xxy + M.yyy + M.yzz);
  const double Q3z = 0.5 * (M.xxz + M.yyz + M.zzz);
  const double Q2 = 0.5 * (M.xx + M.yy + M.zz);
  const double alpha = kEwAlphaFac / L, alpha2 = alpha * alpha;
  const double k1 = kPi / (alpha2 * L * L * L);
  const double ka = 2.0 * alpha / sqrt(kPi);
  int nrep = (int)ceil(kEwCut);
  if (n_replicas > nrep) nrep = n_replicas;
  const double cut2 = kEwCut * kEwCut * L * L, inner2 = 1.2e-3 * L * L;
  double p = mass * k1, fx = 0.0, fy = 0.0, fz = 0.0;
  const double dx = x[i] - cm[0], dy = y[i] - cm[1], dz = z[i] - cm[2];
  for (int ix = -nrep; ix <= nrep; ix++) {
    const bool hole_x = ix >= -n_replicas && ix <= n_replicas;
    const double rx = dx + ix * L;
    for (int iy = -nrep; iy <= nrep; iy++) {
      const bool hole_xy = hole_x && iy >= -n_replicas && iy <= n_replicas;
      const double ry = dy + iy * L;
      for (int iz = -nrep; iz <= nrep; iz++) {
        const bool hole = hole_xy && iz >= -n_replicas && iz <= n_replicas;
        const double rz = dz + iz * L;
        const double r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > cut2 && !hole) continue;
        double g0, g1, g2, g3, g4, g5;
        if (r2 < inner2) {
          const double s2 = r2 * alpha2;
          double an = ka;
          g0 = an * ((1.0 / 3.0) * s2 - 1.0);
          an *= 2.0 * alpha2;
          g1 = an * ((1.0 / 5.0) * s2 - (1.0 / 3.0));
          an *= 2.0 * alpha2;
          g2 = an * ((1.0 / 7.0) * s2 - (1.0 / 5.0));
          an *= 2.0 * alpha2;
          g3 = an * ((1.0 / 9.0) * s2 - (1.0 / 7.0));
          an *= 2.0 * alpha2;
          g4 = an * ((1.0 / 11.0) * s2 - (1.0 / 9.0));
          an *= 2.0 * alpha2;
          g5 = an * ((1.0 / 13.0) * s2 - (1.0 / 11.0));
        } else {
          const double r = sqrt(r2);
          const double ir2 = 1.0 / r2;
          const double a = exp(-r2 * alpha2) * ka * ir2;
          g0 = (hole ? -erf(alpha * r) : erfc(alpha * r)) / r;
          g1 = g0 * ir2 + a;
          double an = 2.0 * alpha2;
          g2 = 3.0 * g1 * ir2 + an * a;
          an *= 2.0 * alpha2;
          g3 = 5.0 * g2 * ir2 + an * a;
          an *= 2.0 * alpha2;
          g4 = 7.0 * g3 * ir2 + an * a;
          an *= 2.0 * alpha2;
          g5 = 9.0 * g4 * ir2 + an * a;
        }
        double q2[3], q3[3], q4[3];
        momc_contract(M, rx, ry, rz, q2, q3, q4);
        const double Q4vx = Q4xx * rx + Q4xy * ry + Q4xz * rz;
        const double Q4vy = Q4xy * rx + Q4yy * ry + Q4yz * rz;
        const double Q4vz = Q4xz * rx + Q4yz * ry + Q4zz * rz;
        const double Q2mir = 0.5 * (q2[0] * rx + q2[1] * ry + q2[2] * rz) -
                             (Q3x * rx + Q3y * ry + Q3z * rz) + Q4;
        const double Q3mir = (1.0 / 3.0) * (q3[0] * rx + q3[1] * ry + q3[2] * rz) -
                             0.5 * (Q4vx * rx + Q4vy * ry + Q4vz * rz);
        const double Q4mir = 0.25 * (q4[0] * rx + q4[1] * ry + q4[2] * rz);
        const double Qta = g1 * mass - g2 * Q2 + g3 * Q2mir + g4 * Q3mir + g5 * Q4mir;
        p -= g0 * mass - g1 * Q2 + g2 * Q2mir + g3 * Q3mir + g4 * Q4mir;
        fx += g2 * (q2[0] - Q3x) + g3 * (q3[0] - Q4vx) + g4 * q4[0] - rx * Qta;
        fy += g2 * (q2[1] - Q3y) + g3 * (q3[1] - Q4vy) + g4 * q4[1] - ry * Qta;
        fz += g2 * (q2[2] - Q3z) + g3 * (q3[2] - Q4vz) + g4 * q4[2] - rz * Qta;
      }
    }
  }
  const int nh = *count;
  for (int k = 0; k < nh; k++) {
    const double hd = table[5 * k] * dx + table[5 * k + 1] * dy + table[5 * k + 2] * dz;
    const double c = cos(hd), sn = sin(hd), cc = table[5 * k + 3], cs = table[5 * k + 4];
    p += cc * c + cs * sn;
    const double f = cc * sn - cs * c;
    fx += table[5 * k] * f;
    fy += table[5 * k + 1] * f;
    fz += table[5 * k + 2] * f;
  }
  ax[i] += fx;
  ay[i] += fy;
  az[i] += fz;
  pot[i] += p;
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
                     double* o3, double* stage, hipStream_t s) {
  if (n <= 0) return 0;
  Res4* st = reinterpret_cast<Res4*>(stage);
  unsort_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, a0, a1, a2, a3, st);
  if (hipGetLastError() != hipSuccess) return -1;
  split4_kernel<<<grid_for(n), 256, 0, s>>>(st, n, o0, o1, o2, o3);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_kick_drift(int64_t n, double dt, int mode, double* x, double* y, double* z,
                         double* vx, double* vy, double* vz, const double* ax, const double* ay,
                         const double* az, hipStream_t s) {
  if (n <= 0) return 0;
  kick_drift_kernel<<<grid_for(n), 256, 0, s>>>(n, dt, mode, x, y, z, vx, vy, vz, ax, ay, az);
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

int pt_launch_min(const double* v, int64_t n, double* parts, int max_parts, double* out,
                  hipStream_t s) {
  int nb = (int)((n + 255) / 256);
  if (nb > max_parts) nb = max_parts;
  if (nb < 1) nb = 1;
  min_partial_kernel<<<nb, 256, 0, s>>>(v, n, parts);
  min_final_kernel<<<1, 256, 0, s>>>(parts, nb, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n <= 0) return 0;
  fill_u64_kernel<<<grid_for(n), 256, 0, s>>>(p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Fourier table entries the Ewald kernels can hold: the integer vectors of a
// 7^3 cube (the cut 2.8 keeps 80 of them); doubles per entry: 4 in the
// quadrupole form, 5 in the HEXADECAPOLE form.
int64_t pt_ewald_table_entries() { return 343; }

int pt_launch_ewald(int64_t n, const double* x, const double* y, const double* z,
                    const double* nd, int64_t nn, double L, int n_replicas, double* table,
                    int32_t* count, double* ax, double* ay, double* az, double* pot,
                    hipStream_t s) {
  if (n <= 0) return 0;
  ewald_table_kernel<<<1, 64, 0, s>>>(nd, nn, L, table, count);
  ewald_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, x, y, z, nd, nn, L, n_replicas,
                                                           table, count, ax, ay, az, pot);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_ewald_hexa(int64_t n, const double* x, const double* y, const double* z,
                         const double* nd, int64_t nn, const HexNode* hn, double L,
                         int n_replicas, double* table, int32_t* count, double* ax, double* ay,
                         double* az, double* pot, hipStream_t s) {
  if (n <= 0) return 0;
  ewald_hexa_table_kernel<<<1, 64, 0, s>>>(nd, nn, hn, L, table, count);
  ewald_hexa_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, x, y, z, nd, nn, hn, L,
                                                                n_replicas, table, count, ax, ay,
                                                                az, pot);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

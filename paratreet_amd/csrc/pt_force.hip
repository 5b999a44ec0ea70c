// Particle-cell and particle-particle interactions over the traversal lists.
//
// Reference: GravityVisitor::node with openSoftening/addGravity/SPLINEQ and
// GravityVisitor::leaf with SPLINE (examples/GravityVisitor.h:37-160,184-224),
// applied through SpatialNode::applyAcceleration/applyPotential
// (src/Node.h:30-38).  Potential is accumulated by the quadrupole branch of
// node() only, as in the reference (addGravity and leaf() add none).
//
// Work mapping (one wavefront per bucket, buckets hold ~5 particles on
// average): the 64 lanes take 64 *sources* at a time (P-C nodes, or source
// particles flattened out of the P-P leaf list with a wave prefix sum), and
// every lane loops over the bucket's targets, which sit in LDS and are read as
// broadcasts.  Each lane keeps one accumulator per target; a wave reduction at
// the end produces each target's sum.  Targets beyond kNT are handled in
// further passes over the same lists.
//
// Two arithmetic forms:
//   * FP64: the reference's expressions in double, sqrt and 1/x included.
//   * FP32C: coordinates are shifted to the bucket's box centre in FP64 and
//     rounded to FP32, the interaction arithmetic runs in FP32 with the
//     hardware reciprocal square root, and every per-target accumulator is a
//     Kahan (compensated) sum; lane sums are combined in FP64.  The openSoftening
//     decision (monopole vs quadrupole) is taken in FP64 in both forms, so both
//     evaluate the same interaction set with the same kernels.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kForceWaves = 4;
constexpr int kForceThreads = kForceWaves * PT_WAVE;
constexpr int kNT = 8;  // targets per pass

// ---------------------------------------------------------------- FP64 ----

__device__ __forceinline__ void spline_b64(double r2, double twoh, double& b) {
  const double r = sqrt(r2);
  if (r < twoh) {
    const double dih = 2.0 / twoh;
    const double u = r * dih;
    if (u < 1.0) {
      b = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
    } else {
      const double dir = 1.0 / r;
      b = -1.0 / 15.0 * dir * dir * dir +
          dih * dih * dih * (8.0 / 3.0 - 3.0 * u + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
    }
  } else {
    const double a = 1.0 / r;
    b = a * a * a;
  }
}

__device__ __forceinline__ void splineq64(double dir, double r2, double twoh, double& a,
                                          double& b, double& c, double& d) {
  if (r2 < twoh * twoh) {
    const double dih = 2.0 / twoh;
    const double u = dih / dir;
    if (u < 1.0) {
      a = dih * (7.0) / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u -
          1.0 / 10.0 * u * u * u * u * u;
      b = dih * dih * dih * (4.0) / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u;
      c = dih * dih * dih * dih * dih * (12.0) / 5.0 - 3.0 / 2.0 * u;
      d = 3.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
    } else {
      a = -1.0 / 15.0 * dir + dih * (8.0 / 5.0) - 4.0 / 3.0 * u * u + u * u * u -
          3.0 / 10.0 * u * u * u * u + 1.0 / 30.0 * u * u * u * u * u;
      b = -1.0 / 15.0 * dir * dir * dir + dih * dih * dih * (8.0) / 3.0 - 3.0 * u +
          6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u;
      c = -1.0 / 5.0 * dir * dir * dir * dir * dir + 3.0 * dih * dih * dih * dih * dir +
          dih * dih * dih * dih * dih * (-12.0) / 5.0 + 1.0 / 2.0 * u;
      d = -dir * dir * dir * dir * dir * dir * dir +
          3.0 * dih * dih * dih * dih * dir * dir * dir -
          1.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
    }
  } else {
    a = dir;
    b = a * a * a;
    c = 3.0 * b * a * a;
    d = 5.0 * c * a * a;
  }
}

// GravityVisitor::openSoftening (examples/GravityVisitor.h:122-128), source
// sphere (cm + offset, 2 soft) against the target bucket's sphere and box.
__device__ __forceinline__ bool open_softening(double sx, double sy, double sz, double ssoft,
                                               const Bucket& b) {
  const double sr = 2.0 * ssoft;
  {
    const double dx = sx - b.cx, dy = sy - b.cy, dz = sz - b.cz;
    const double dsq = dx * dx + dy * dy + dz * dz;
    const double rs = sr + 2.0 * b.soft;
    if (dsq <= rs * rs) return true;
  }
  double dsq = 0.0;
  if (sx < b.lox) { const double t = b.lox - sx; dsq += t * t; }
  else if (sx > b.hix) { const double t = sx - b.hix; dsq += t * t; }
  if (sy < b.loy) { const double t = b.loy - sy; dsq += t * t; }
  else if (sy > b.hiy) { const double t = sy - b.hiy; dsq += t * t; }
  if (sz < b.loz) { const double t = b.loz - sz; dsq += t * t; }
  else if (sz > b.hiz) { const double t = sz - b.hiz; dsq += t * t; }
  return dsq <= sr * sr;
}

struct Tgt64 {
  double x, y, z, e;
};

__global__ void __launch_bounds__(kForceThreads)
force64_kernel(const Bucket* __restrict__ bk, int64_t nb, const GravNode* __restrict__ gn,
               const TopoNode* __restrict__ topo, const double* __restrict__ px,
               const double* __restrict__ py, const double* __restrict__ pz,
               const double* __restrict__ pm, const double* __restrict__ pe,
               const int64_t* __restrict__ pp_off, const int32_t* __restrict__ pp,
               const int64_t* __restrict__ pc_off, const int32_t* __restrict__ pc, double ox,
               double oy, double oz, double* __restrict__ ax, double* __restrict__ ay,
               double* __restrict__ az, double* __restrict__ pot) {
  __shared__ Tgt64 s_t[kForceWaves][kNT];
  __shared__ int32_t s_beg[kForceWaves][PT_WAVE];
  __shared__ int32_t s_exc[kForceWaves][PT_WAVE];
  const int wave = threadIdx.x / PT_WAVE;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t bi = __builtin_amdgcn_readfirstlane((int)((int64_t)blockIdx.x * kForceWaves + wave));
  if (bi >= nb) return;
  const Bucket b = bk[bi];
  const int64_t pc0 = pc_off[bi], pc1 = pc_off[bi + 1];
  const int64_t pp0 = pp_off[bi], pp1 = pp_off[bi + 1];

  for (int g0 = 0; g0 < b.count; g0 += kNT) {
    const int ng = (b.count - g0) < kNT ? (b.count - g0) : kNT;
    if (lane < ng) {
      const int64_t i = (int64_t)b.begin + g0 + lane;
      s_t[wave][lane] = Tgt64{px[i], py[i], pz[i], pe[i]};
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    double sx[kNT], sy[kNT], sz[kNT], sp[kNT];
#pragma unroll
    for (int t = 0; t < kNT; t++) sx[t] = sy[t] = sz[t] = sp[t] = 0.0;

    // ---- particle-cell: GravityVisitor::node
    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int64_t e = e0 + lane;
      if (e < pc1) {
        const GravNode s = gn[pc[e]];
        const double cx = s.cx + ox, cy = s.cy + oy, cz = s.cz + oz;
        if (open_softening(cx, cy, cz, s.soft, b)) {
          // addGravity: unsoftened monopole, no potential
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt64 tg = s_t[wave][t];
              const double dx = cx - tg.x, dy = cy - tg.y, dz = cz - tg.z;
              const double rsq = dx * dx + dy * dy + dz * dz;
              if (rsq != 0) {
                const double f = s.mass / (rsq * sqrt(rsq));
                sx[t] += dx * f;
                sy[t] += dy * f;
                sz[t] += dz * f;
              }
            }
          }
        } else {
          const double tr = 0.5 * (s.qxx + s.qyy + s.qzz);
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt64 tg = s_t[wave][t];
              const double rx = (tg.x - s.cx) - ox;
              const double ry = (tg.y - s.cy) - oy;
              const double rz = (tg.z - s.cz) - oz;
              const double rsq = rx * rx + ry * ry + rz * rz;
              const double dir = 1.0 / sqrt(rsq);
              double a, bb, c, d;
              splineq64(dir, rsq, s.soft + tg.e, a, bb, c, d);
              const double qx = s.qxx * rx + s.qxy * ry + s.qxz * rz;
              const double qy = s.qxy * rx + s.qyy * ry + s.qyz * rz;
              const double qz = s.qxz * rx + s.qyz * ry + s.qzz * rz;
              const double qir = 0.5 * (qx * rx + qy * ry + qz * rz);
              const double qir3 = bb * s.mass + d * qir - c * tr;
              sp[t] += -s.mass * a - c * qir + bb * tr;
              sx[t] += (-qir3 * rx) + (c * qx);
              sy[t] += (-qir3 * ry) + (c * qy);
              sz[t] += (-qir3 * rz) + (c * qz);
            }
          }
        }
      }
    }

    // ---- particle-particle: GravityVisitor::leaf over flattened sources
    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      const int64_t l = l0 + lane;
      int32_t cnt = 0, beg = 0;
      if (l < pp1) {
        const TopoNode tp = topo[pp[l]];
        cnt = tp.count;
        beg = tp.begin;
      }
      const int incl = pt_wave_incl_scan(cnt);
      const int total = __shfl(incl, PT_WAVE - 1, PT_WAVE);
      s_beg[wave][lane] = beg;
      s_exc[wave][lane] = incl - cnt;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      for (int s0 = 0; s0 < total; s0 += PT_WAVE) {
        const int slot = s0 + lane;
        if (slot < total) {
          // last list entry whose exclusive offset is <= slot
          int lo = 0, hi = PT_WAVE - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_exc[wave][mid] <= slot) lo = mid;
            else hi = mid - 1;
          }
          const int64_t j = (int64_t)s_beg[wave][lo] + (slot - s_exc[wave][lo]);
          const double jx = px[j] + ox, jy = py[j] + oy, jz = pz[j] + oz;
          const double jm = pm[j], je = pe[j];
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt64 tg = s_t[wave][t];
              const double dx = jx - tg.x, dy = jy - tg.y, dz = jz - tg.z;
              const double rsq = dx * dx + dy * dy + dz * dz;
              if (rsq != 0) {
                double bv;
                spline_b64(rsq, je + tg.e, bv);
                const double f = bv * jm;
                sx[t] += dx * f;
                sy[t] += dy * f;
                sz[t] += dz * f;
              }
            }
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }

#pragma unroll
    for (int t = 0; t < kNT; t++) {
      if (t < ng) {
        const double rx = pt_wave_sum(sx[t]);
        const double ry = pt_wave_sum(sy[t]);
        const double rz = pt_wave_sum(sz[t]);
        const double rp = pt_wave_sum(sp[t]);
        if (lane == t) {
          const int64_t i = (int64_t)b.begin + g0 + t;
          ax[i] = rx;
          ay[i] = ry;
          az[i] = rz;
          pot[i] = rp;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

// ---------------------------------------------------------------- FP32C ---

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

__device__ __forceinline__ float spline_b32(float r2, float twoh) {
  const float rinv = rsqrt32(r2);
  const float r = r2 * rinv;
  if (r < twoh) {
    const float dih = 2.0f / twoh;
    const float u = r * dih;
    if (u < 1.0f) {
      return dih * dih * dih * (4.0f / 3.0f - 6.0f / 5.0f * u * u + 1.0f / 2.0f * u * u * u);
    }
    return -1.0f / 15.0f * rinv * rinv * rinv +
           dih * dih * dih * (8.0f / 3.0f - 3.0f * u + 6.0f / 5.0f * u * u - 1.0f / 6.0f * u * u * u);
  }
  return rinv * rinv * rinv;
}

__device__ __forceinline__ void splineq32(float dir, float r2, float twoh, float& a, float& b,
                                          float& c, float& d) {
  if (r2 < twoh * twoh) {
    const float dih = 2.0f / twoh;
    const float u = dih / dir;
    if (u < 1.0f) {
      a = dih * (7.0f) / 5.0f - 2.0f / 3.0f * u * u + 3.0f / 10.0f * u * u * u * u -
          1.0f / 10.0f * u * u * u * u * u;
      b = dih * dih * dih * (4.0f) / 3.0f - 6.0f / 5.0f * u * u + 1.0f / 2.0f * u * u * u;
      c = dih * dih * dih * dih * dih * (12.0f) / 5.0f - 3.0f / 2.0f * u;
      d = 3.0f / 2.0f * dih * dih * dih * dih * dih * dih * dir;
    } else {
      a = -1.0f / 15.0f * dir + dih * (8.0f / 5.0f) - 4.0f / 3.0f * u * u + u * u * u -
          3.0f / 10.0f * u * u * u * u + 1.0f / 30.0f * u * u * u * u * u;
      b = -1.0f / 15.0f * dir * dir * dir + dih * dih * dih * (8.0f) / 3.0f - 3.0f * u +
          6.0f / 5.0f * u * u - 1.0f / 6.0f * u * u * u;
      c = -1.0f / 5.0f * dir * dir * dir * dir * dir + 3.0f * dih * dih * dih * dih * dir +
          dih * dih * dih * dih * dih * (-12.0f) / 5.0f + 1.0f / 2.0f * u;
      d = -dir * dir * dir * dir * dir * dir * dir +
          3.0f * dih * dih * dih * dih * dir * dir * dir -
          1.0f / 2.0f * dih * dih * dih * dih * dih * dih * dir;
    }
  } else {
    a = dir;
    b = a * a * a;
    c = 3.0f * b * a * a;
    d = 5.0f * c * a * a;
  }
}

struct Tgt32 {
  float x, y, z, e;
};

__global__ void __launch_bounds__(kForceThreads)
force32c_kernel(const Bucket* __restrict__ bk, int64_t nb, const GravNode32* __restrict__ gn,
                const TopoNode* __restrict__ topo, const double* __restrict__ px,
                const double* __restrict__ py, const double* __restrict__ pz,
                const double* __restrict__ pm, const double* __restrict__ pe,
                const int64_t* __restrict__ pp_off, const int32_t* __restrict__ pp,
                const int64_t* __restrict__ pc_off, const int32_t* __restrict__ pc, double ox,
                double oy, double oz, double* __restrict__ ax, double* __restrict__ ay,
                double* __restrict__ az, double* __restrict__ pot) {
  __shared__ Tgt32 s_t[kForceWaves][kNT];
  __shared__ int32_t s_beg[kForceWaves][PT_WAVE];
  __shared__ int32_t s_exc[kForceWaves][PT_WAVE];
  const int wave = threadIdx.x / PT_WAVE;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t bi = __builtin_amdgcn_readfirstlane((int)((int64_t)blockIdx.x * kForceWaves + wave));
  if (bi >= nb) return;
  const Bucket b = bk[bi];
  // shift origin: the bucket's box centre
  const double ccx = (b.lox + b.hix) * 0.5;
  const double ccy = (b.loy + b.hiy) * 0.5;
  const double ccz = (b.loz + b.hiz) * 0.5;
  const int64_t pc0 = pc_off[bi], pc1 = pc_off[bi + 1];
  const int64_t pp0 = pp_off[bi], pp1 = pp_off[bi + 1];

  for (int g0 = 0; g0 < b.count; g0 += kNT) {
    const int ng = (b.count - g0) < kNT ? (b.count - g0) : kNT;
    if (lane < ng) {
      const int64_t i = (int64_t)b.begin + g0 + lane;
      s_t[wave][lane] = Tgt32{(float)(px[i] - ccx), (float)(py[i] - ccy), (float)(pz[i] - ccz),
                              (float)pe[i]};
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    float sx[kNT], sy[kNT], sz[kNT], sp[kNT];
    float cx_[kNT], cy_[kNT], cz_[kNT], cp_[kNT];
#pragma unroll
    for (int t = 0; t < kNT; t++) {
      sx[t] = sy[t] = sz[t] = sp[t] = 0.f;
      cx_[t] = cy_[t] = cz_[t] = cp_[t] = 0.f;
    }

    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int64_t e = e0 + lane;
      if (e < pc1) {
        const GravNode32 s = gn[pc[e]];
        const double cx = s.cx + ox, cy = s.cy + oy, cz = s.cz + oz;
        const bool mono = open_softening(cx, cy, cz, (double)s.soft, b);
        // source centre in the bucket frame
        const float qx0 = (float)(cx - ccx), qy0 = (float)(cy - ccy), qz0 = (float)(cz - ccz);
        if (mono) {
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt32 tg = s_t[wave][t];
              const float dx = qx0 - tg.x, dy = qy0 - tg.y, dz = qz0 - tg.z;
              const float rsq = dx * dx + dy * dy + dz * dz;
              if (rsq != 0.f) {
                const float ri = rsqrt32(rsq);
                const float f = s.mass * ri * ri * ri;
                kahan_add(sx[t], cx_[t], dx * f);
                kahan_add(sy[t], cy_[t], dy * f);
                kahan_add(sz[t], cz_[t], dz * f);
              }
            }
          }
        } else {
          const float tr = 0.5f * (s.qxx + s.qyy + s.qzz);
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt32 tg = s_t[wave][t];
              const float rx = tg.x - qx0, ry = tg.y - qy0, rz = tg.z - qz0;
              const float rsq = rx * rx + ry * ry + rz * rz;
              const float dir = rsqrt32(rsq);
              float a, bb, c, d;
              splineq32(dir, rsq, s.soft + tg.e, a, bb, c, d);
              const float qx = s.qxx * rx + s.qxy * ry + s.qxz * rz;
              const float qy = s.qxy * rx + s.qyy * ry + s.qyz * rz;
              const float qz = s.qxz * rx + s.qyz * ry + s.qzz * rz;
              const float qir = 0.5f * (qx * rx + qy * ry + qz * rz);
              const float qir3 = bb * s.mass + d * qir - c * tr;
              kahan_add(sp[t], cp_[t], -s.mass * a - c * qir + bb * tr);
              kahan_add(sx[t], cx_[t], (-qir3 * rx) + (c * qx));
              kahan_add(sy[t], cy_[t], (-qir3 * ry) + (c * qy));
              kahan_add(sz[t], cz_[t], (-qir3 * rz) + (c * qz));
            }
          }
        }
      }
    }

    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      const int64_t l = l0 + lane;
      int32_t cnt = 0, beg = 0;
      if (l < pp1) {
        const TopoNode tp = topo[pp[l]];
        cnt = tp.count;
        beg = tp.begin;
      }
      const int incl = pt_wave_incl_scan(cnt);
      const int total = __shfl(incl, PT_WAVE - 1, PT_WAVE);
      s_beg[wave][lane] = beg;
      s_exc[wave][lane] = incl - cnt;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      for (int s0 = 0; s0 < total; s0 += PT_WAVE) {
        const int slot = s0 + lane;
        if (slot < total) {
          int lo = 0, hi = PT_WAVE - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_exc[wave][mid] <= slot) lo = mid;
            else hi = mid - 1;
          }
          const int64_t j = (int64_t)s_beg[wave][lo] + (slot - s_exc[wave][lo]);
          const float jx = (float)((px[j] + ox) - ccx);
          const float jy = (float)((py[j] + oy) - ccy);
          const float jz = (float)((pz[j] + oz) - ccz);
          const float jm = (float)pm[j], je = (float)pe[j];
#pragma unroll
          for (int t = 0; t < kNT; t++) {
            if (t < ng) {
              const Tgt32 tg = s_t[wave][t];
              const float dx = jx - tg.x, dy = jy - tg.y, dz = jz - tg.z;
              const float rsq = dx * dx + dy * dy + dz * dz;
              if (rsq != 0.f) {
                const float f = spline_b32(rsq, je + tg.e) * jm;
                kahan_add(sx[t], cx_[t], dx * f);
                kahan_add(sy[t], cy_[t], dy * f);
                kahan_add(sz[t], cz_[t], dz * f);
              }
            }
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }

#pragma unroll
    for (int t = 0; t < kNT; t++) {
      if (t < ng) {
        const double rx = pt_wave_sum((double)sx[t] - (double)cx_[t]);
        const double ry = pt_wave_sum((double)sy[t] - (double)cy_[t]);
        const double rz = pt_wave_sum((double)sz[t] - (double)cz_[t]);
        const double rp = pt_wave_sum((double)sp[t] - (double)cp_[t]);
        if (lane == t) {
          const int64_t i = (int64_t)b.begin + g0 + t;
          ax[i] = rx;
          ay[i] = ry;
          az[i] = rz;
          pot[i] = rp;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

}  // namespace

int pt_launch_force(int precision, const Bucket* bk, int64_t nb, const GravNode* gn,
                    const GravNode32* gn32, const TopoNode* topo, const double* x,
                    const double* y, const double* z, const double* m, const double* e,
                    const int64_t* pp_off, const int32_t* pp, const int64_t* pc_off,
                    const int32_t* pc, const double off[3], double* ax, double* ay, double* az,
                    double* pot, hipStream_t s) {
  if (nb <= 0) return 0;
  const unsigned grid = (unsigned)((nb + kForceWaves - 1) / kForceWaves);
  if (precision == 0) {
    force64_kernel<<<grid, kForceThreads, 0, s>>>(bk, nb, gn, topo, x, y, z, m, e, pp_off, pp,
                                                  pc_off, pc, off[0], off[1], off[2], ax, ay, az,
                                                  pot);
  } else {
    force32c_kernel<<<grid, kForceThreads, 0, s>>>(bk, nb, gn32, topo, x, y, z, m, e, pp_off,
                                                   pp, pc_off, pc, off[0], off[1], off[2], ax,
                                                   ay, az, pot);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// GravityVisitor interactions on gfx950: list-driven and fused walk+interact.
//
// Reference: GravityVisitor::node with openSoftening / addGravity / SPLINEQ and
// GravityVisitor::leaf with SPLINE (examples/GravityVisitor.h:37-160,184-224),
// applied through SpatialNode::applyAcceleration / applyPotential
// (src/Node.h:30-38).  Potential comes from node()'s quadrupole branch only,
// as in the reference (addGravity and leaf() add none).
//
// Work mapping: one wavefront per target bucket (buckets hold ~5 particles).
// With nt targets the lanes are laid out as lane = s * nt + t: lane holds
// target t in registers and is source slot s of S = 64 / nt slots.  Sources
// (closed nodes for node(), particles of leaf sources for leaf()) are staged
// 64 at a time into LDS by one lane each -- the staging lane also converts
// them (FP32 shift, openSoftening decision) -- and every lane then walks the
// chunk with stride S.  All 64 lanes do useful pair work except the
// 64 - S * nt remainder.  At the end the S partial sums of each target are
// reduced through LDS in a fixed order.
//
// Two arithmetic forms:
//   FP64  -- the reference's expressions in double (sqrt and 1/x included);
//   FP32C -- positions shifted to the bucket's box centre in FP64 and rounded
//            to FP32, interaction arithmetic in FP32 with the hardware
//            reciprocal square root, each chunk's partial sum (<= 64/S terms)
//            added to the per-lane accumulators with Kahan compensation;
//            partials are combined in FP64.  The openSoftening choice
//            (monopole vs quadrupole) is made in FP64 in both forms.
//
// Kernels:
//   interact_kernel  -- reads the per-bucket lists of pt_traverse;
//   gravity_kernel   -- walks the tree itself (pt_walk.h, same decisions) and
//                       evaluates each 64-entry chunk of closed nodes / leaf
//                       sources as soon as the walk has produced it, so no
//                       list ever reaches HBM;
//   grouped_gravity_kernel -- the same per bucket run of at most 64
//                       particles: one walk with the run's union box shares
//                       every source all its buckets agree on, each bucket
//                       walks on its own only below the nodes they decide
//                       differently (the default of pt_compute_forces).
//
// This file is compiled with -ffp-contract=off because it contains the walk's
// opening test; the interaction arithmetic uses explicit fma() where the
// reference's separately rounded order does not matter for the 1e-12 bound.
#include <type_traits>

#include "pt_common.h"
#include "pt_kernels.h"
#include "pt_walk.h"

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = kWaves * PT_WAVE;
constexpr int kGroupMax = PT_WAVE;  // targets per pass (one per lane at most)
// Walk stack entries per wave.  The deepest walks measured on the workloads
// here (a CPU replay of this walk order over 4000 sampled buckets: Plummer
// 1M, theta 0.7: 57 entries; 64-halo mixture 4M, theta 0.5, bucket 32: 114)
// fit the small stack with room to spare; a bucket that overflows it makes
// the whole evaluation rerun with the big one.  The small stack keeps the
// FP32 work-group at 6 per CU in LDS (26.5 KiB each).
constexpr int kSmallStack = 192;
constexpr int kBigStack = 1536;

// ---------------------------------------------------------------------------
// arithmetic
// ---------------------------------------------------------------------------

__device__ __forceinline__ double spline_b64(double r2, double twoh) {
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

template <typename R>
__device__ __forceinline__ void splineq(R dir, R r2, R twoh, R& a, R& b, R& c, R& d) {
  if (r2 < twoh * twoh) {
    const R dih = R(2.0) / twoh;
    const R u = dih / dir;
    if (u < R(1.0)) {
      a = dih * R(7.0) / R(5.0) - R(2.0 / 3.0) * u * u + R(3.0 / 10.0) * u * u * u * u -
          R(1.0 / 10.0) * u * u * u * u * u;
      b = dih * dih * dih * R(4.0) / R(3.0) - R(6.0 / 5.0) * u * u + R(1.0 / 2.0) * u * u * u;
      c = dih * dih * dih * dih * dih * R(12.0) / R(5.0) - R(3.0 / 2.0) * u;
      d = R(3.0 / 2.0) * dih * dih * dih * dih * dih * dih * dir;
    } else {
      a = R(-1.0 / 15.0) * dir + dih * R(8.0 / 5.0) - R(4.0 / 3.0) * u * u + u * u * u -
          R(3.0 / 10.0) * u * u * u * u + R(1.0 / 30.0) * u * u * u * u * u;
      b = R(-1.0 / 15.0) * dir * dir * dir + dih * dih * dih * R(8.0) / R(3.0) - R(3.0) * u +
          R(6.0 / 5.0) * u * u - R(1.0 / 6.0) * u * u * u;
      c = R(-1.0 / 5.0) * dir * dir * dir * dir * dir + R(3.0) * dih * dih * dih * dih * dir +
          dih * dih * dih * dih * dih * R(-12.0) / R(5.0) + R(1.0 / 2.0) * u;
      d = -dir * dir * dir * dir * dir * dir * dir +
          R(3.0) * dih * dih * dih * dih * dir * dir * dir -
          R(1.0 / 2.0) * dih * dih * dih * dih * dih * dih * dir;
    }
  } else {
    a = dir;
    b = a * a * a;
    c = R(3.0) * b * a * a;
    d = R(5.0) * c * a * a;
  }
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }


// Compensated (Kahan) FP32 accumulation.
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// ---------------------------------------------------------------------------
// LDS source records
// ---------------------------------------------------------------------------

// Closed-node records: soft < 0 marks a node that openSoftening sends to
// addGravity (unsoftened monopole); tr = 0.5 * trace(Q).
struct __align__(16) SrcC64 {  // closed node, FP64
  double cx, cy, cz, mass, soft, qxx, qxy, qxz, qyy, qyz, qzz, tr;
};
struct __align__(16) SrcP64 {  // source particle, FP64
  double x, y, z, m, e, pad;
};
// Closed node, FP32, centre in the bucket frame, moments scaled by 3:
// Q' = 3Q and tr = 3 * (0.5 * trace(Q)), the factors node_pair_far's form
// of the quadrupole terms needs.
struct SrcC32 {
  float cx, cy, cz, mass, soft, qxx, qxy, qxz, qyy, qyz, qzz, tr;
};
struct SrcP32 {  // source particle, FP32, bucket frame
  float x, y, z, m, e, pad0, pad1, pad2;
};

// A staged chunk holds up to 64 real sources; the slots the stride-S walk
// reads beyond them hold null sources (zero mass and moments, far away), so
// the evaluation loops need no bounds tests.  With S = 64 / nt the walk reads
// S * ceil(n / S) <= 84 slots for n <= 64 (nt = 3, n = 64).
constexpr int kSlots = 84;
constexpr double kNullPos = 1.0e6;

// Per-wave LDS working set: the staged chunk (also the final reduction area),
// the walk stack and the pending closed-node / leaf-source buffers (a step
// adds at most 64 entries to a buffer holding fewer than 64).
template <bool F32, int STACK>
struct __align__(16) WaveLds {
  static constexpr int kStage = kSlots * (F32 ? sizeof(SrcC32) : sizeof(SrcC64));
  static_assert(64 * 4 * sizeof(double) <= kStage, "reduction area");
  unsigned char stage[kStage];
  int32_t stack[STACK];
  int32_t pcbuf[2 * PT_WAVE];
  int32_t ppbeg[2 * PT_WAVE];
  int32_t ppcnt[2 * PT_WAVE];
  int32_t fl_mark[PT_WAVE];
};

// Per-lane target and accumulator state for one target group.
template <bool F32>
struct Lanes;

template <>
struct Lanes<false> {
  double tx, ty, tz, te;
  double ax, ay, az, ap;
  float emax;  // largest target softening of the group
};
template <>
struct Lanes<true> {
  float tx, ty, tz, te;
  float ax, ay, az, ap;
  float cx, cy, cz, cp;  // Kahan compensations
  float emax;            // largest target softening of the group
};

template <bool F32>
using SrcC = typename std::conditional<F32, SrcC32, SrcC64>::type;
template <bool F32>
using SrcP = typename std::conditional<F32, SrcP32, SrcP64>::type;

// The target bucket in its own (box-centre) frame, FP32, for the quick
// openSoftening rejection of the FP32 path.
struct BucketF {
  float lox, loy, loz, hix, hiy, hiz, cx, cy, cz, soft, ext;
};

__device__ __forceinline__ BucketF bucket_frame(const Bucket& b, double ccx, double ccy,
                                                double ccz) {
  BucketF f;
  f.lox = (float)(b.lox - ccx);
  f.loy = (float)(b.loy - ccy);
  f.loz = (float)(b.loz - ccz);
  f.hix = (float)(b.hix - ccx);
  f.hiy = (float)(b.hiy - ccy);
  f.hiz = (float)(b.hiz - ccz);
  f.cx = (float)(b.cx - ccx);
  f.cy = (float)(b.cy - ccy);
  f.cz = (float)(b.cz - ccz);
  f.soft = (float)b.soft;
  f.ext = (f.hix - f.lox) + (f.hiy - f.loy) + (f.hiz - f.loz);
  return f;
}

// True when the FP32 estimate proves openSoftening false: both distances
// exceed their thresholds by a margin (1e-5 relative on the radius plus
// 1e-6 of the coordinate magnitudes) far above the FP32 rounding of the
// shifted coordinates.  Otherwise the exact FP64 test decides.
__device__ __forceinline__ bool softening_far32(float sx, float sy, float sz, float ssoft,
                                                const BucketF& f) {
  const float eps = 1e-6f * (fabsf(sx) + fabsf(sy) + fabsf(sz) + f.ext);
  const float rb = 2.0f * ssoft * 1.00001f + eps;
  const float ra = (2.0f * ssoft + 2.0f * f.soft) * 1.00001f + eps;
  const float dx = sx - f.cx, dy = sy - f.cy, dz = sz - f.cz;
  const float da = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const float ex = fmaxf(fmaxf(f.lox - sx, sx - f.hix), 0.f);
  const float ey = fmaxf(fmaxf(f.loy - sy, sy - f.hiy), 0.f);
  const float ez = fmaxf(fmaxf(f.loz - sz, sz - f.hiz), 0.f);
  const float db = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
  return da > ra * ra && db > rb * rb;
}

// False when no target of the bucket can come within softening reach of the
// source: the box distance exceeds (e_s + e_max) by a margin (1e-5 relative
// plus 1e-6 of the coordinate magnitudes) far above the rounding of the
// FP32 bucket-frame coordinates, so every pair with this source takes the
// unsoftened branch of SPLINE / SPLINEQ in FP32 and in FP64 alike.  Such
// "far" sources are evaluated by the branch-free loops.
__device__ __forceinline__ bool near32(float sx, float sy, float sz, float se, const BucketF& f,
                                       float emax) {
  const float eps = 1e-6f * (fabsf(sx) + fabsf(sy) + fabsf(sz) + f.ext);
  const float reach = (se + emax) * 1.00001f + eps;
  const float ex = fmaxf(fmaxf(f.lox - sx, sx - f.hix), 0.f);
  const float ey = fmaxf(fmaxf(f.loy - sy, sy - f.hiy), 0.f);
  const float ez = fmaxf(fmaxf(f.loz - sz, sz - f.hiz), 0.f);
  const float db = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
  return !(db > reach * reach);
}

// True when every particle of a leaf source is beyond softening reach of
// every target of the group: the leaf's centre of mass (s, bucket frame) lies
// farther from the bucket box than the leaf's reach bound plus the largest
// source and target softenings, with the margins of near32 (the particles'
// coordinates are bounded by |s| + reach).  All its particles then take the
// branch-free far loop.
__device__ __forceinline__ bool leaf_far32(float sx, float sy, float sz, float lreach,
                                           float lsoft, const BucketF& f, float emax) {
  const float eps = 1e-6f * (fabsf(sx) + fabsf(sy) + fabsf(sz) + f.ext + 3.0f * lreach);
  const float reach = lreach + (lsoft + emax) * 1.00001f + eps;
  const float ex = fmaxf(fmaxf(f.lox - sx, sx - f.hix), 0.f);
  const float ey = fmaxf(fmaxf(f.loy - sy, sy - f.hiy), 0.f);
  const float ez = fmaxf(fmaxf(f.loz - sz, sz - f.hiz), 0.f);
  const float db = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
  return db > reach * reach;
}

// A leaf source's count with the far flag of leaf_far32 in bit 30.
constexpr int32_t kFarLeaf = 1 << 30;

__device__ __forceinline__ int32_t leaf_entry(int32_t count, double cx, double cy, double cz,
                                              float lreach, float lsoft, const BucketF& f,
                                              float emax, double ox, double oy, double oz,
                                              double ccx, double ccy, double ccz) {
  const bool far = leaf_far32((float)((cx + ox) - ccx), (float)((cy + oy) - ccy),
                              (float)((cz + oz) - ccz), lreach, lsoft, f, emax);
  return far ? (count | kFarLeaf) : count;
}

// The same classifications from the walk's exact squared box distance dsq
// of the source's centre (+ offset), with margins (1e-4 relative on the
// radii, 2e-4 on the reaches, 4e-6 of the bucket extent) that cover the FP32
// rounding of the bucket-frame coordinates the pair loops see.  A leaf is far
// when every particle (within leaf_reach of its centre) is beyond softening
// reach of every target; a closed node is far when, in addition, the spheres
// openSoftening tests cannot meet (dsq above (2 soft + 2 bucket soft)^2 rules
// out both of its tests), so node() takes SPLINEQ's unsoftened branch.
__device__ __forceinline__ bool leaf_far_dsq(double dsq, float lreach, float lsoft,
                                             const BucketF& f, float emax) {
  const float reach = fmaf(lreach + lsoft + emax, 1.0002f, fmaf(3e-6f, lreach, 4e-6f * f.ext));
  return (float)dsq > 1.0001f * (reach * reach);
}

__device__ __forceinline__ bool node_far_dsq(double dsq, float nsoft, const BucketF& f,
                                             float emax) {
  const float mono = 2.0f * (nsoft + f.soft);
  const float reach = fmaf(nsoft + emax, 1.0002f, 4e-6f * f.ext);
  const float m = fmaxf(mono, reach);
  return (float)dsq > 1.0001f * (m * m);
}

// A closed node's id with the far flag of node_far_dsq in bit 30.
constexpr int32_t kFarNode = 1 << 30;

__device__ __forceinline__ float frame32(double v, double o, double cc) {
  return (float)((v + o) - cc);
}


struct Group {
  int ng;      // targets in this pass
  int S;       // source slots = 64 / ng
  int magic;   // ceil(2^16 / S): ceil_div_s in integer arithmetic
  int cw;      // sources per full chunk of the FP64 fused walk: the
               // largest multiple of S up to 64, so a full chunk fills every
               // slot (measured faster in FP64, slower in FP32C, which keeps
               // chunks of 64)
  int s, t;    // this lane's slot and target (inactive lanes read slot 0)
  bool act;    // lane participates
  int64_t first;  // sorted index of the group's first target
};

__device__ __forceinline__ Group make_group(const Bucket& b, int g0) {
  Group g;
  const int rem = b.count - g0;
  g.ng = rem < kGroupMax ? rem : kGroupMax;
  g.S = PT_WAVE / g.ng;
  g.magic = (65536 + g.S - 1) / g.S;
  // floor(64 / S) = (64 * magic) >> 16: the excess of magic / 2^16 over 1 / S
  // moves 64 / S by less than 64 / 2^16 < 1 / S
  g.cw = g.S * ((PT_WAVE * g.magic) >> 16);
  const int lane = __lane_id();
  const int s = lane / g.ng;
  g.act = s < g.S;
  g.s = g.act ? s : 0;
  g.t = g.act ? lane - s * g.ng : 0;
  g.first = (int64_t)b.begin + g0;
  return g;
}

// A wave-uniform value moved to a scalar register when U (scalar loop control
// for the FP64 kernel; the FP32C kernel, at its 96-register budget, measured
// faster without it).
template <bool U>
__device__ __forceinline__ int uniform_if(int v) {
  if constexpr (U) return __builtin_amdgcn_readfirstlane(v);
  else return v;
}

// ceil(n / S) for 0 <= n <= 128 without a division: with y = n + S - 1 <= 191
// and magic = ceil(2^16 / S), (y * magic) >> 16 = floor(y / S) because
// y * magic / 2^16 exceeds y / S by less than 191 / 2^16 < 1 / 64 <= 1 / S,
// and y / S is either an integer or at least 1 / S below the next one.  All
// operands are wave-uniform, so this is scalar arithmetic.
__device__ __forceinline__ int ceil_div_s(int n, const Group& g) {
  return ((n + g.S - 1) * g.magic) >> 16;
}

template <bool F32>
__device__ __forceinline__ void load_targets(Lanes<F32>& L, const Group& g,
                                             const double* __restrict__ px,
                                             const double* __restrict__ py,
                                             const double* __restrict__ pz,
                                             const double* __restrict__ pe, double ccx,
                                             double ccy, double ccz) {
  const int64_t i = g.first + g.t;
  if constexpr (F32) {
    L.tx = (float)(px[i] - ccx);
    L.ty = (float)(py[i] - ccy);
    L.tz = (float)(pz[i] - ccz);
    L.te = (float)pe[i];
    L.cx = L.cy = L.cz = L.cp = 0.f;
  } else {
    L.tx = px[i];
    L.ty = py[i];
    L.tz = pz[i];
    L.te = pe[i];
  }
  L.ax = L.ay = L.az = L.ap = 0;
  L.emax = pt_wave_max_f32((float)L.te);
}

// ---- staging (one lane per source) ----------------------------------------

// Closed-node source record; `near` marks a node that takes the monopole
// (openSoftening) branch or may be softened for some target of the group.
template <bool F32>
__device__ __forceinline__ SrcC<F32> node_record(int32_t id, bool far,
                                                 const GravNode* __restrict__ gn,
                                                 const GravNode32* __restrict__ gn32,
                                                 const Bucket& b, const BucketF& bf, float emax,
                                                 double ox, double oy, double oz, double ccx,
                                                 double ccy, double ccz, bool& near) {
  if constexpr (F32) {
    const GravNode32 h = gn32[id];
    const double cx = h.cx + ox, cy = h.cy + oy, cz = h.cz + oz;
    const double soft = h.soft64;
    SrcC32 r;
    r.cx = (float)(cx - ccx);
    r.cy = (float)(cy - ccy);
    r.cz = (float)(cz - ccz);
    bool mono = false;
    near = false;
    if (!far) {
      mono = softening_far32(r.cx, r.cy, r.cz, h.soft, bf) ? false
                                                            : pt_open_softening(cx, cy, cz, soft, b);
      near = mono || near32(r.cx, r.cy, r.cz, h.soft, bf, emax);
    }
    r.mass = h.mass;
    r.soft = mono ? -1.f : h.soft;
    r.qxx = 3.0f * h.qxx;
    r.qxy = 3.0f * h.qxy;
    r.qxz = 3.0f * h.qxz;
    r.qyy = 3.0f * h.qyy;
    r.qyz = 3.0f * h.qyz;
    r.qzz = 3.0f * h.qzz;
    r.tr = 1.5f * (h.qxx + h.qyy + h.qzz);
    return r;
  } else {
    const GravNode s = gn[id];
    const double cx = s.cx + ox, cy = s.cy + oy, cz = s.cz + oz;
    const double soft = s.soft;
    SrcC64 r;
    r.cx = s.cx;
    r.cy = s.cy;
    r.cz = s.cz;
    r.mass = s.mass;
    bool mono = false;
    near = false;
    if (!far) {
      mono = pt_open_softening(cx, cy, cz, s.soft, b);
      near = mono || near32((float)(cx - ccx), (float)(cy - ccy), (float)(cz - ccz),
                            (float)soft, bf, emax);
    }
    r.soft = mono ? -1.0 : s.soft;
    r.qxx = s.qxx;
    r.qxy = s.qxy;
    r.qxz = s.qxz;
    r.qyy = s.qyy;
    r.qyz = s.qyz;
    r.qzz = s.qzz;
    r.tr = s.tr;
    return r;
  }
}

template <bool F32, typename W_t>
__device__ __forceinline__ void stage_null_node(W_t& W, int i) {
  if constexpr (F32) {
    SrcC32 r{};
    r.cx = r.cy = r.cz = (float)kNullPos;
    reinterpret_cast<SrcC32*>(W.stage)[i] = r;
  } else {
    SrcC64 r{};
    r.cx = r.cy = r.cz = kNullPos;
    reinterpret_cast<SrcC64*>(W.stage)[i] = r;
  }
}

// Leaf-source particle record; `near` marks a particle that may be within
// softening reach of some target of the group.
template <bool F32>
__device__ __forceinline__ SrcP<F32> particle_record(int64_t j, const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const double* __restrict__ pm,
                                                     const double* __restrict__ pe,
                                                     const BucketF& bf, float emax, double ox,
                                                     double oy, double oz, double ccx, double ccy,
                                                     double ccz, bool& near) {
  if constexpr (F32) {
    SrcP32 r;
    r.x = (float)((px[j] + ox) - ccx);
    r.y = (float)((py[j] + oy) - ccy);
    r.z = (float)((pz[j] + oz) - ccz);
    r.m = (float)pm[j];
    r.e = (float)pe[j];
    r.pad0 = r.pad1 = r.pad2 = 0.f;
    near = near32(r.x, r.y, r.z, r.e, bf, emax);
    return r;
  } else {
    SrcP64 r;
    r.x = px[j];
    r.y = py[j];
    r.z = pz[j];
    r.m = pm[j];
    r.e = pe[j];
    r.pad = 0.0;
    near = near32(frame32(r.x, ox, ccx), frame32(r.y, oy, ccy), frame32(r.z, oz, ccz), (float)r.e,
                  bf, emax);
    return r;
  }
}

template <bool F32, typename W_t>
__device__ __forceinline__ void stage_null_particle(W_t& W, int i) {
  if constexpr (F32) {
    SrcP32 r{};
    r.x = r.y = r.z = (float)kNullPos;
    reinterpret_cast<SrcP32*>(W.stage)[i] = r;
  } else {
    SrcP64 r{};
    r.x = r.y = r.z = kNullPos;
    reinterpret_cast<SrcP64*>(W.stage)[i] = r;
  }
}

// Null sources for the slots [nsrc, S * ceil(nsrc / S)) the evaluation
// will read.
template <bool F32, bool NODE, typename W_t>
__device__ __forceinline__ void pad_chunk(W_t& W, int nsrc, const Group& g) {
  const int end = g.S * ceil_div_s(nsrc, g);
  const int lane = __lane_id();
  for (int k = nsrc + lane; k < end; k += PT_WAVE) {
    if (NODE) stage_null_node<F32, W_t>(W, k);
    else stage_null_particle<F32, W_t>(W, k);
  }
}

// ---- chunk evaluation (all lanes, no bounds tests) -------------------------
//
// A chunk holds its "near" sources in slots [0, nnear) and the rest after
// them, so iterations k >= kfar = ceil(nnear / S) read far sources (or null
// padding) only and run the unsoftened expressions without branches; the
// first kfar iterations run the general code (monopole branch, SPLINE /
// SPLINEQ softened branches).  The arithmetic of a pair is the same in both
// loops.

// One closed-node pair, general form (addGravity when s.soft < 0).
template <bool F32>
__device__ __forceinline__ void node_pair(const SrcC<F32>& s, const Lanes<F32>& L,
                                          typename std::conditional<F32, float, double>::type& px,
                                          typename std::conditional<F32, float, double>::type& py,
                                          typename std::conditional<F32, float, double>::type& pz,
                                          typename std::conditional<F32, float, double>::type& pp,
                                          double ox, double oy, double oz) {
  if constexpr (F32) {
    if (s.soft < 0.f) {
      // addGravity: unsoftened monopole, no potential
      const float dx = s.cx - L.tx, dy = s.cy - L.ty, dz = s.cz - L.tz;
      const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float ri = fminf(rsqrt32(rsq), 1e10f);
      const float f = s.mass * (ri * ri * ri);
      px = fmaf(dx, f, px);
      py = fmaf(dy, f, py);
      pz = fmaf(dz, f, pz);
    } else {
      const float rx = L.tx - s.cx, ry = L.ty - s.cy, rz = L.tz - s.cz;
      const float rsq = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
      const float dir = rsqrt32(rsq);
      const float twoh = s.soft + L.te;
      float a, bb, c, d;
      if (rsq < twoh * twoh) {
        splineq<float>(dir, rsq, twoh, a, bb, c, d);
      } else {
        const float dir2 = dir * dir;
        a = dir;
        bb = dir * dir2;
        c = 3.0f * bb * dir2;
        d = 5.0f * c * dir2;
      }
      // the staged moments are 3Q and 3 tr: qx = 3 (Qr)_x, qir = r.Q.r / 2
      const float qx = fmaf(s.qxx, rx, fmaf(s.qxy, ry, s.qxz * rz));
      const float qy = fmaf(s.qxy, rx, fmaf(s.qyy, ry, s.qyz * rz));
      const float qz = fmaf(s.qxz, rx, fmaf(s.qyz, ry, s.qzz * rz));
      const float qir = (1.0f / 6.0f) * fmaf(qx, rx, fmaf(qy, ry, qz * rz));
      const float c3 = (1.0f / 3.0f) * c, bb3 = (1.0f / 3.0f) * bb;
      const float qir3 = fmaf(bb, s.mass, fmaf(d, qir, -c3 * s.tr));
      pp = fmaf(-s.mass, a, pp);
      pp = fmaf(-c, qir, pp);
      pp = fmaf(bb3, s.tr, pp);
      px = fmaf(-qir3, rx, px);
      px = fmaf(c3, qx, px);
      py = fmaf(-qir3, ry, py);
      py = fmaf(c3, qy, py);
      pz = fmaf(-qir3, rz, pz);
      pz = fmaf(c3, qz, pz);
    }
  } else {
    if (s.soft < 0.0) {
      const double dx = (s.cx + ox) - L.tx, dy = (s.cy + oy) - L.ty, dz = (s.cz + oz) - L.tz;
      const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
      const double ir = pt_rsqrt64(rsq);
      double f = s.mass * (ir * ir * ir);
      f = rsq != 0 ? f : 0.0;
      px = fma(dx, f, px);
      py = fma(dy, f, py);
      pz = fma(dz, f, pz);
    } else {
      const double rx = (L.tx - s.cx) - ox, ry = (L.ty - s.cy) - oy, rz = (L.tz - s.cz) - oz;
      const double rsq = fma(rx, rx, fma(ry, ry, rz * rz));
      const double dir = pt_rsqrt64(rsq);
      double a, bb, c, d;
      splineq<double>(dir, rsq, s.soft + L.te, a, bb, c, d);
      const double qx = fma(s.qxx, rx, fma(s.qxy, ry, s.qxz * rz));
      const double qy = fma(s.qxy, rx, fma(s.qyy, ry, s.qyz * rz));
      const double qz = fma(s.qxz, rx, fma(s.qyz, ry, s.qzz * rz));
      const double qir = 0.5 * fma(qx, rx, fma(qy, ry, qz * rz));
      const double qir3 = fma(bb, s.mass, fma(d, qir, -c * s.tr));
      pp += fma(-s.mass, a, fma(-c, qir, bb * s.tr));
      px += fma(-qir3, rx, c * qx);
      py += fma(-qir3, ry, c * qy);
      pz += fma(-qir3, rz, c * qz);
    }
  }
}

// One closed-node pair known to take the unsoftened quadrupole branch.
template <bool F32>
__device__ __forceinline__ void node_pair_far(
    const SrcC<F32>& s, const Lanes<F32>& L,
    typename std::conditional<F32, float, double>::type& px,
    typename std::conditional<F32, float, double>::type& py,
    typename std::conditional<F32, float, double>::type& pz,
    typename std::conditional<F32, float, double>::type& pp,
    typename std::conditional<F32, float, double>::type& pq, double ox, double oy, double oz) {
  if constexpr (F32) {
    // The unsoftened branch of SPLINEQ (a = 1/r, b = 1/r^3, c = 3/r^5,
    // d = 15/r^7) regrouped over the staged 3Q and 3 tr: with u = 1/r^2,
    // q = 3Qr, k = u (r.q) = 3 r.Q.r / r^2 and C = b u,
    //   acc = -qir3 r + c Qr      = C (q - (m r^2 + 2.5 k - 3 tr) r),
    //   pot = -m a - c qir + b tr = -m a + (b / 3) (3 tr - 1.5 k),
    // the (b / 3) term summed in pq and scaled once per chunk.
    const float rx = L.tx - s.cx, ry = L.ty - s.cy, rz = L.tz - s.cz;
    const float rsq = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    const float v = rsqrt32(rsq);
    const float u = v * v;
    const float b = v * u;
    const float cu = b * u;
    const float qx = fmaf(s.qxx, rx, fmaf(s.qxy, ry, s.qxz * rz));
    const float qy = fmaf(s.qxy, rx, fmaf(s.qyy, ry, s.qyz * rz));
    const float qz = fmaf(s.qxz, rx, fmaf(s.qyz, ry, s.qzz * rz));
    const float k = u * fmaf(qx, rx, fmaf(qy, ry, qz * rz));
    const float gq = fmaf(s.mass, rsq, fmaf(2.5f, k, -s.tr));
    px = fmaf(cu, fmaf(-gq, rx, qx), px);
    py = fmaf(cu, fmaf(-gq, ry, qy), py);
    pz = fmaf(cu, fmaf(-gq, rz, qz), pz);
    pp = fmaf(-s.mass, v, pp);
    pq = fmaf(b, fmaf(-1.5f, k, s.tr), pq);
  } else {
    const double rx = (L.tx - s.cx) - ox, ry = (L.ty - s.cy) - oy, rz = (L.tz - s.cz) - oz;
    const double rsq = fma(rx, rx, fma(ry, ry, rz * rz));
    const double a = pt_rsqrt64(rsq);
    const double bb = a * a * a;
    const double c = 3.0 * bb * a * a;
    const double d = 5.0 * c * a * a;
    const double qx = fma(s.qxx, rx, fma(s.qxy, ry, s.qxz * rz));
    const double qy = fma(s.qxy, rx, fma(s.qyy, ry, s.qyz * rz));
    const double qz = fma(s.qxz, rx, fma(s.qyz, ry, s.qzz * rz));
    const double qir = 0.5 * fma(qx, rx, fma(qy, ry, qz * rz));
    const double qir3 = fma(bb, s.mass, fma(d, qir, -c * s.tr));
    pp += fma(-s.mass, a, fma(-c, qir, bb * s.tr));
    px += fma(-qir3, rx, c * qx);
    py += fma(-qir3, ry, c * qy);
    pz += fma(-qir3, rz, c * qz);
  }
}

template <bool F32, typename W_t>
__device__ __forceinline__ void eval_nodes(const W_t& W, const Group& g, Lanes<F32>& L, int kfar,
                                           int kend, double ox, double oy, double oz) {
  const SrcC<F32>* src = reinterpret_cast<const SrcC<F32>*>(W.stage);
  if constexpr (F32) {
    float px = 0.f, py = 0.f, pz = 0.f, pp = 0.f, pq = 0.f;
    for (int k = 0; k < kfar; k++) node_pair<true>(src[g.s + k * g.S], L, px, py, pz, pp, ox, oy, oz);
#pragma unroll 2
    for (int k = kfar; k < kend; k++)
      node_pair_far<true>(src[g.s + k * g.S], L, px, py, pz, pp, pq, ox, oy, oz);
    pp = fmaf(1.0f / 3.0f, pq, pp);
    kahan_add(L.ax, L.cx, px);
    kahan_add(L.ay, L.cy, py);
    kahan_add(L.az, L.cz, pz);
    kahan_add(L.ap, L.cp, pp);
  } else {
    for (int k = 0; k < kfar; k++)
      node_pair<false>(src[g.s + k * g.S], L, L.ax, L.ay, L.az, L.ap, ox, oy, oz);
    for (int k = kfar; k < kend; k++)
      node_pair_far<false>(src[g.s + k * g.S], L, L.ax, L.ay, L.az, L.ap, L.ap, ox, oy, oz);
  }
}

template <bool F32, typename W_t>
__device__ __forceinline__ void eval_particles(const W_t& W, const Group& g, Lanes<F32>& L,
                                               int kfar, int kend, double ox, double oy,
                                               double oz) {
  if constexpr (F32) {
    float px = 0.f, py = 0.f, pz = 0.f;
    const SrcP32* src = reinterpret_cast<const SrcP32*>(W.stage);
    for (int k = 0; k < kfar; k++) {
      const SrcP32 s = src[g.s + k * g.S];
      const float dx = s.x - L.tx, dy = s.y - L.ty, dz = s.z - L.tz;
      const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float twoh = s.e + L.te;
      // rsq == 0 (the target itself): the clamp keeps b finite, and dx = dy =
      // dz = 0 then contributes exactly zero
      const float rinv = fminf(rsqrt32(rsq), 1e10f);
      float b = rinv * rinv * rinv;
      if (rsq < twoh * twoh) {
        // GravityVisitor::SPLINE softened branches
        const float r = rsq * rinv;
        const float dih = 2.0f * __builtin_amdgcn_rcpf(twoh);
        const float u = r * dih;
        const float dih3 = dih * dih * dih;
        b = u < 1.0f
                ? dih3 * (4.0f / 3.0f - 6.0f / 5.0f * u * u + 1.0f / 2.0f * u * u * u)
                : -1.0f / 15.0f * rinv * rinv * rinv +
                      dih3 * (8.0f / 3.0f - 3.0f * u + 6.0f / 5.0f * u * u - 1.0f / 6.0f * u * u * u);
      }
      const float f = b * s.m;
      px = fmaf(dx, f, px);
      py = fmaf(dy, f, py);
      pz = fmaf(dz, f, pz);
    }
#pragma unroll 4
    for (int k = kfar; k < kend; k++) {
      // far sources: rsq > 0 and beyond every softening reach
      const SrcP32 s = src[g.s + k * g.S];
      const float dx = s.x - L.tx, dy = s.y - L.ty, dz = s.z - L.tz;
      const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float rinv = rsqrt32(rsq);
      const float f = (rinv * rinv * rinv) * s.m;
      px = fmaf(dx, f, px);
      py = fmaf(dy, f, py);
      pz = fmaf(dz, f, pz);
    }
    kahan_add(L.ax, L.cx, px);
    kahan_add(L.ay, L.cy, py);
    kahan_add(L.az, L.cz, pz);
  } else {
    const SrcP64* src = reinterpret_cast<const SrcP64*>(W.stage);
    for (int k = 0; k < kfar; k++) {
      const SrcP64 s = src[g.s + k * g.S];
      const double dx = (s.x + ox) - L.tx, dy = (s.y + oy) - L.ty, dz = (s.z + oz) - L.tz;
      const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
      double f = spline_b64(rsq, s.e + L.te) * s.m;
      f = rsq != 0 ? f : 0.0;
      L.ax = fma(dx, f, L.ax);
      L.ay = fma(dy, f, L.ay);
      L.az = fma(dz, f, L.az);
    }
    for (int k = kfar; k < kend; k++) {
      const SrcP64 s = src[g.s + k * g.S];
      const double dx = (s.x + ox) - L.tx, dy = (s.y + oy) - L.ty, dz = (s.z + oz) - L.tz;
      const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
      const double a = pt_rsqrt64(rsq);
      const double f = (a * a * a) * s.m;
      L.ax = fma(dx, f, L.ax);
      L.ay = fma(dy, f, L.ay);
      L.az = fma(dz, f, L.az);
    }
  }
}

// Flattens up to 64 leaf sources (begin and flagged count, leaf_entry, in
// the calling lane; count 0 = none) into particles and evaluates them 64 at a
// time.  Leaves not flagged far come first, so each window holds its near
// particles in its first slots and the rest take the branch-free loop.  Each
// window marks the slot where every leaf starts (the leaf's lane in LDS, a
// bit in a ballot); a slot's leaf is the mark at the highest start bit at or
// below it, or the previous window's last leaf when there is none.
template <bool F32, typename W_t>
__device__ __forceinline__ void do_leaves(W_t& W, int32_t beg, int32_t cntf, const Group& g,
                                          Lanes<F32>& L, const double* __restrict__ px,
                                          const double* __restrict__ py,
                                          const double* __restrict__ pz,
                                          const double* __restrict__ pm,
                                          const double* __restrict__ pe, const BucketF& bf,
                                          double ox, double oy, double oz, double ccx, double ccy,
                                          double ccz) {
  const int lane = __lane_id();
  const bool far = (cntf & kFarLeaf) != 0;
  const int cnt = cntf & ~kFarLeaf;
  const int cn = far ? 0 : cnt, cf = far ? cnt : 0;
  const int incl_n = pt_wave_incl_scan(cn);
  const int incl_f = pt_wave_incl_scan(cf);
  const int near_total = __builtin_amdgcn_readlane(incl_n, PT_WAVE - 1);
  const int total = near_total + __builtin_amdgcn_readlane(incl_f, PT_WAVE - 1);
  const int exc = far ? near_total + incl_f - cf : incl_n - cn;
  int carry = 0;
  // windows of cw particles (FP64: a multiple of S, full windows need no
  // padding)
  const int cw = F32 ? PT_WAVE : g.cw;
  for (int s0 = 0; s0 < total; s0 += cw) {
    W.fl_mark[lane] = -1;
    __builtin_amdgcn_wave_barrier();
    if (cnt > 0 && exc >= s0 && exc < s0 + cw) W.fl_mark[exc - s0] = lane;
    __builtin_amdgcn_wave_barrier();
    const unsigned long long starts = __ballot(W.fl_mark[lane] >= 0);
    const unsigned long long upto = starts & (~0ull >> (PT_WAVE - 1 - lane));
    const int leaf = upto ? W.fl_mark[PT_WAVE - 1 - __clzll(upto)] : carry;
    carry = __builtin_amdgcn_readlane(leaf, cw - 1);
    const int lbeg = __shfl(beg, leaf, PT_WAVE);
    const int lexc = __shfl(exc, leaf, PT_WAVE);
    const int slot = s0 + lane;
    if (lane < cw && slot < total) {
      const int64_t j = (int64_t)lbeg + (slot - lexc);
      bool nr;
      reinterpret_cast<SrcP<F32>*>(W.stage)[lane] = particle_record<F32>(
          j, px, py, pz, pm, pe, bf, L.emax, ox, oy, oz, ccx, ccy, ccz, nr);
    }
    const int nsrc = (total - s0) < cw ? (total - s0) : cw;
    const int nn = near_total - s0;
    const int nnear = nn < 0 ? 0 : (nn < nsrc ? nn : nsrc);
    pad_chunk<F32, false, W_t>(W, nsrc, g);
    __builtin_amdgcn_wave_barrier();
    eval_particles<F32, W_t>(W, g, L, uniform_if<!F32>(ceil_div_s(nnear, g)),
                             uniform_if<!F32>(ceil_div_s(nsrc, g)), ox, oy, oz);
    __builtin_amdgcn_wave_barrier();
  }
}

template <bool F32, typename W_t>
__device__ __forceinline__ void do_nodes(W_t& W, int32_t id, int nsrc, const Group& g,
                                         Lanes<F32>& L, const GravNode* __restrict__ gn,
                                         const GravNode32* __restrict__ gn32, const Bucket& b,
                                         const BucketF& bf, double ox, double oy, double oz,
                                         double ccx, double ccy, double ccz) {
  const bool act = (int)__lane_id() < nsrc;
  bool nr = false;
  SrcC<F32> r{};
  if (act)
    r = node_record<F32>(id & ~kFarNode, (id & kFarNode) != 0, gn, gn32, b, bf, L.emax, ox, oy,
                         oz, ccx, ccy, ccz, nr);
  // near (monopole or possibly softened) nodes first, then far ones
  const unsigned long long mn = pt_ballot(act && nr), mf = pt_ballot(act && !nr);
  const int nnear = __popcll(mn);
  if (act) reinterpret_cast<SrcC<F32>*>(W.stage)[nr ? pt_rank_in(mn) : nnear + pt_rank_in(mf)] = r;
  pad_chunk<F32, true, W_t>(W, nsrc, g);
  __builtin_amdgcn_wave_barrier();
  eval_nodes<F32, W_t>(W, g, L, uniform_if<!F32>(ceil_div_s(nnear, g)),
                       uniform_if<!F32>(ceil_div_s(nsrc, g)), ox, oy, oz);
  __builtin_amdgcn_wave_barrier();
}

// Sums the S partials of each target (fixed order) and stores the results.
template <bool F32, typename W_t>
__device__ __forceinline__ void finish_group(W_t& W, const Group& g, const Lanes<F32>& L,
                                             double* __restrict__ ax, double* __restrict__ ay,
                                             double* __restrict__ az, double* __restrict__ pot,
                                             bool acc = false) {
  double* red = reinterpret_cast<double*>(W.stage);
  const int lane = __lane_id();
  double v0, v1, v2, v3;
  if constexpr (F32) {
    v0 = (double)L.ax - (double)L.cx;
    v1 = (double)L.ay - (double)L.cy;
    v2 = (double)L.az - (double)L.cz;
    v3 = (double)L.ap - (double)L.cp;
  } else {
    v0 = L.ax;
    v1 = L.ay;
    v2 = L.az;
    v3 = L.ap;
  }
  red[lane * 4 + 0] = g.act ? v0 : 0.0;
  red[lane * 4 + 1] = g.act ? v1 : 0.0;
  red[lane * 4 + 2] = g.act ? v2 : 0.0;
  red[lane * 4 + 3] = g.act ? v3 : 0.0;
  __builtin_amdgcn_wave_barrier();
  if (lane < g.ng) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int s = 0; s < g.S; s++) {
      const int l = s * g.ng + lane;
      s0 += red[l * 4 + 0];
      s1 += red[l * 4 + 1];
      s2 += red[l * 4 + 2];
      s3 += red[l * 4 + 3];
    }
    const int64_t i = g.first + lane;
    if (acc) {
      // another periodic replica: added to what the earlier ones wrote
      ax[i] += s0;
      ay[i] += s1;
      az[i] += s2;
      pot[i] += s3;
    } else {
      ax[i] = s0;
      ay[i] = s1;
      az[i] = s2;
      pot[i] = s3;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <bool F32>
__global__ void __launch_bounds__(kThreads)
interact_kernel(const Bucket* __restrict__ bk, int64_t nb, const TreeNode* __restrict__ nodes,
                const GravNode* __restrict__ gn, const GravNode32* __restrict__ gn32,
                const double* __restrict__ px,
                const double* __restrict__ py, const double* __restrict__ pz,
                const double* __restrict__ pm, const double* __restrict__ pe,
                const int64_t* __restrict__ pp_off, const int32_t* __restrict__ pp,
                const int64_t* __restrict__ pc_off, const int32_t* __restrict__ pc, double ox,
                double oy, double oz, double* __restrict__ ax, double* __restrict__ ay,
                double* __restrict__ az, double* __restrict__ pot) {
  using W_t = WaveLds<F32, 8>;
  __shared__ W_t lds[kWaves];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t bi = pt_xcd_block(blockIdx.x, gridDim.x) * kWaves + wave;
  if (bi >= nb) return;
  W_t& W = lds[wave];
  const Bucket b = bk[bi];
  const double ccx = (b.lox + b.hix) * 0.5, ccy = (b.loy + b.hiy) * 0.5,
               ccz = (b.loz + b.hiz) * 0.5;
  const BucketF bf = bucket_frame(b, ccx, ccy, ccz);
  const int64_t pc0 = pc_off[bi], pc1 = pc_off[bi + 1];
  const int64_t pp0 = pp_off[bi], pp1 = pp_off[bi + 1];
  for (int g0 = 0; g0 < b.count; g0 += kGroupMax) {
    const Group g = make_group(b, g0);
    Lanes<F32> L;
    load_targets<F32>(L, g, px, py, pz, pe, ccx, ccy, ccz);
    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int nsrc = (pc1 - e0) < PT_WAVE ? (int)(pc1 - e0) : PT_WAVE;
      const int32_t id = lane < nsrc ? pc[e0 + lane] : 0;
      do_nodes<F32, W_t>(W, id, nsrc, g, L, gn, gn32, b, bf, ox, oy, oz, ccx, ccy, ccz);
    }
    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      int32_t beg = 0, cnt = 0;
      if (l0 + lane < pp1) {
        const TreeNode nd = nodes[pp[l0 + lane]];
        beg = nd.begin;
        cnt = leaf_entry(nd.count, nd.cx, nd.cy, nd.cz, nd.leaf_reach, nd.leaf_soft, bf, L.emax,
                         ox, oy, oz, ccx, ccy, ccz);
      }
      do_leaves<F32, W_t>(W, beg, cnt, g, L, px, py, pz, pm, pe, bf, ox, oy, oz, ccx, ccy, ccz);
    }
    finish_group<F32, W_t>(W, g, L, ax, ay, az, pot);
  }
}

// ---------------------------------------------------------------------------
// HEXADECAPOLE build: node() with the reduced moments to fourth order
// (GravityVisitor::node, examples/GravityVisitor.h:184-206 -> momEvalFmomrcm,
// examples/moments.C:1469-1526), over the per-bucket lists, FP64.  leaf() is
// the quadrupole build's.
// ---------------------------------------------------------------------------

// Staged HEXADECAPOLE source: centre of mass, mass, monopole flag (1 when
// openSoftening sends the node to addGravity) and the 21 reduced moments.
struct SrcH {
  double cx, cy, cz, mass, mono;
  double h[PT_HEXA_MOMS];
};
// The same in the compensated-FP32 form: centre in the bucket frame, FP32
// mass and moments (a 64-entry stage takes the bytes of a 32-entry SrcH one).
struct SrcH32 {
  float cx, cy, cz, mass, mono;
  float h[PT_HEXA_MOMS];
};

// Closed nodes are staged kHexaStage at a time (half a chunk), which keeps a
// wave's LDS near 9 KiB, and the fused kernel is held to kHexaWaves waves per
// SIMD (168 VGPRs, a small spill): 1M Plummer, theta 0.7, 13.8 ms against
// 17.7 ms with a 64-entry stage at two waves per SIMD (four waves spill 264
// bytes per lane and take 21.5 ms).
constexpr int kHexaStage = 32;
constexpr int kHexaWaves = 3;
constexpr int kHexaWaves32 = 4;  // the compensated-FP32 form (LDS allows four)

struct __align__(16) HexaLds {
  static constexpr int kStage = kHexaStage * sizeof(SrcH);
  static_assert(kSlots * sizeof(SrcP64) <= kStage, "particle chunk");
  static_assert(64 * 4 * sizeof(double) <= kStage, "reduction area");
  unsigned char stage[kStage];
  int32_t fl_mark[PT_WAVE];
};

// Potential and acceleration of reduced moments h at r = target - cm: with
// c_l = (2l-1)!!/l! (3/2, 5/2, 35/8), v_l = M_l contracted with l-1 copies of
// r and s_l = v_l . r,
//   pot = -(m / r + sum_l c_l s_l / r^(2l+1)),
//   acc = sum_l c_l (l v_l / r^(2l+1) - (2l+1) s_l r / r^(2l+3)) - m r / r^3
// (the trace-free moments give the Legendre series of the potential).
template <typename R>
__device__ __forceinline__ R fma_r(R a, R b, R c) {
  if constexpr (std::is_same<R, float>::value) return fmaf(a, b, c);
  else return fma(a, b, c);
}

template <typename R>
__device__ __forceinline__ void hexa_eval(const R* __restrict__ h, R m, R rx, R ry, R rz, R& pot,
                                          R& gx, R& gy, R& gz) {
  // every multiply-add is an explicit fma (this file is compiled without
  // contraction)
  const R xx = h[0], yy = h[1], xy = h[2], xz = h[3], yz = h[4];
  const R zz = -(xx + yy);
  const R xxx = h[5], xyy = h[6], xxy = h[7], yyy = h[8], xxz = h[9], yyz = h[10],
          xyz = h[11];
  const R xzz = -(xxx + xyy), yzz = -(xxy + yyy), zzz = -(xxz + yyz);
  const R xxxx = h[12], xyyy = h[13], xxxy = h[14], yyyy = h[15], xxxz = h[16],
          yyyz = h[17], xxyy = h[18], xxyz = h[19], xyyz = h[20];
  const R xxzz = -(xxxx + xxyy), yyzz = -(xxyy + yyyy), xyzz = -(xxxy + xyyy);
  const R xzzz = -(xxxz + xyyz), yzzz = -(xxyz + yyyz), zzzz = -(xxzz + yyzz);
  const R rsq = fma_r(rx, rx, fma_r(ry, ry, rz * rz));
  R ir;
  if constexpr (std::is_same<R, float>::value) ir = rsqrt32(rsq);
  else ir = pt_rsqrt64(rsq);
  const R ir2 = ir * ir;
  const R x2 = rx * rx, y2 = ry * ry, z2 = rz * rz;
  const R xyp = rx * ry, xzp = rx * rz, yzp = ry * rz;
  // order 2: v2 = M2 r, s2 = r.M2.r
  const R v2x = fma_r(xx, rx, fma_r(xy, ry, xz * rz));
  const R v2y = fma_r(xy, rx, fma_r(yy, ry, yz * rz));
  const R v2z = fma_r(xz, rx, fma_r(yz, ry, zz * rz));
  const R s2 = fma_r(v2x, rx, fma_r(v2y, ry, v2z * rz));
  // order 3: v3_i = M3_ijk r_j r_k
  const R v3x = fma_r(xxx, x2, fma_r(xyy, y2, fma_r(xzz, z2, R(2.0) * fma_r(xxy, xyp, fma_r(xxz, xzp, xyz * yzp)))));
  const R v3y = fma_r(xxy, x2, fma_r(yyy, y2, fma_r(yzz, z2, R(2.0) * fma_r(xyy, xyp, fma_r(xyz, xzp, yyz * yzp)))));
  const R v3z = fma_r(xxz, x2, fma_r(yyz, y2, fma_r(zzz, z2, R(2.0) * fma_r(xyz, xyp, fma_r(xzz, xzp, yzz * yzp)))));
  const R s3 = fma_r(v3x, rx, fma_r(v3y, ry, v3z * rz));
  // order 4: v4_i = M4_ijkl r_j r_k r_l over the 10 cubic monomials
  const R x3 = x2 * rx, y3 = y2 * ry, z3 = z2 * rz;
  const R x2y = x2 * ry, x2z = x2 * rz, xy2 = rx * y2, xz2 = rx * z2, y2z = y2 * rz,
          yz2 = ry * z2, xyzm = xyp * rz;
  const R v4x =
      fma_r(R(6.0) * xxyz, xyzm,
            fma_r(R(3.0), fma_r(xxxy, x2y, fma_r(xxxz, x2z, fma_r(xxyy, xy2, fma_r(xxzz, xz2, fma_r(xyyz, y2z, xyzz * yz2))))),
                  fma_r(xxxx, x3, fma_r(xyyy, y3, xzzz * z3))));
  const R v4y =
      fma_r(R(6.0) * xyyz, xyzm,
            fma_r(R(3.0), fma_r(xxyy, x2y, fma_r(xxyz, x2z, fma_r(xyyy, xy2, fma_r(xyzz, xz2, fma_r(yyyz, y2z, yyzz * yz2))))),
                  fma_r(xxxy, x3, fma_r(yyyy, y3, yzzz * z3))));
  const R v4z =
      fma_r(R(6.0) * xyzz, xyzm,
            fma_r(R(3.0), fma_r(xxyz, x2y, fma_r(xxzz, x2z, fma_r(xyyz, xy2, fma_r(xzzz, xz2, fma_r(yyzz, y2z, yzzz * yz2))))),
                  fma_r(xxxz, x3, fma_r(yyyz, y3, zzzz * z3))));
  const R s4 = fma_r(v4x, rx, fma_r(v4y, ry, v4z * rz));
  const R ir3 = ir * ir2, ir5 = ir3 * ir2, ir7 = ir5 * ir2, ir9 = ir7 * ir2,
          ir11 = ir9 * ir2;
  // pot = -(m/r + 3/2 s2/r^5 + 5/2 s3/r^7 + 35/8 s4/r^9)
  pot = -fma_r(m, ir, fma_r(R(1.5) * s2, ir5, fma_r(R(2.5) * s3, ir7, R(4.375) * s4 * ir9)));
  const R g = fma_r(m, ir3, fma_r(R(7.5) * s2, ir7, fma_r(R(17.5) * s3, ir9, R(39.375) * s4 * ir11)));
  const R c2 = R(3.0) * ir5, c3 = R(7.5) * ir7, c4 = R(17.5) * ir9;
  gx = fma_r(c2, v2x, fma_r(c3, v3x, fma_r(c4, v4x, -rx * g)));
  gy = fma_r(c2, v2y, fma_r(c3, v3y, fma_r(c4, v4y, -ry * g)));
  gz = fma_r(c2, v2z, fma_r(c3, v3z, fma_r(c4, v4z, -rz * g)));
}

__device__ __forceinline__ void hexa_pair(const SrcH& s, Lanes<false>& L, double ox, double oy,
                                          double oz) {
  if (s.mono != 0.0) {
    // addGravity: unsoftened monopole, no potential
    const double dx = (s.cx + ox) - L.tx, dy = (s.cy + oy) - L.ty, dz = (s.cz + oz) - L.tz;
    const double rsq = dx * dx + dy * dy + dz * dz;
    const double ir = pt_rsqrt64(rsq);
    double f = s.mass * (ir * ir * ir);
    f = rsq != 0 ? f : 0.0;
    L.ax += dx * f;
    L.ay += dy * f;
    L.az += dz * f;
  } else {
    const double rx = (L.tx - s.cx) - ox, ry = (L.ty - s.cy) - oy, rz = (L.tz - s.cz) - oz;
    double p, gx, gy, gz;
    hexa_eval<double>(s.h, s.mass, rx, ry, rz, p, gx, gy, gz);
    L.ap += p;
    L.ax += gx;
    L.ay += gy;
    L.az += gz;
  }
}

// Stages up to 64 closed nodes (node id in the calling lane, kFarNode set
// when no openSoftening test can hold for the group), kHexaStage at a time,
// and evaluates them for the group's targets (slot s of a lane takes sources
// s, s + S, ... of each part; the S partial sums are combined in
// finish_group).
template <typename W_t>
__device__ __forceinline__ void do_nodes_hexa(W_t& W, int32_t id, int nsrc, const Group& g,
                                              Lanes<false>& L, const HexNode* __restrict__ hn,
                                              const Bucket& b, double ox, double oy, double oz) {
  const int lane = __lane_id();
  SrcH* st = reinterpret_cast<SrcH*>(W.stage);
  for (int h = 0; h < nsrc; h += kHexaStage) {
    if (lane >= h && lane < h + kHexaStage && lane < nsrc) {
      const bool far = (id & kFarNode) != 0;
      const HexNode& s = hn[id & ~kFarNode];
      SrcH r;
      r.cx = s.cx;
      r.cy = s.cy;
      r.cz = s.cz;
      r.mass = s.mass;
      r.mono = !far && pt_open_softening(s.cx + ox, s.cy + oy, s.cz + oz, s.soft, b) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < PT_HEXA_MOMS; k++) r.h[k] = s.h[k];
      st[lane - h] = r;
    }
    __builtin_amdgcn_wave_barrier();
    const int n = (nsrc - h) < kHexaStage ? (nsrc - h) : kHexaStage;
    const int kend = ceil_div_s(n, g);
    for (int k = 0; k < kend; k++) {
      const int i = g.s + k * g.S;
      if (i < n) hexa_pair(st[i], L, ox, oy, oz);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// One HEXADECAPOLE pair in FP32 (bucket frame), added to the chunk partials.
__device__ __forceinline__ void hexa_pair32(const SrcH32& s, const Lanes<true>& L, float& px,
                                            float& py, float& pz, float& pp) {
  if (s.mono != 0.f) {
    // addGravity: unsoftened monopole, no potential (rsq == 0 only for a
    // coincident point, where dx = dy = dz = 0 contribute nothing)
    const float dx = s.cx - L.tx, dy = s.cy - L.ty, dz = s.cz - L.tz;
    const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ri = fminf(rsqrt32(rsq), 1e10f);
    const float f = s.mass * (ri * ri * ri);
    px = fmaf(dx, f, px);
    py = fmaf(dy, f, py);
    pz = fmaf(dz, f, pz);
  } else {
    float p, gx, gy, gz;
    hexa_eval<float>(s.h, s.mass, L.tx - s.cx, L.ty - s.cy, L.tz - s.cz, p, gx, gy, gz);
    pp += p;
    px += gx;
    py += gy;
    pz += gz;
  }
}

// do_nodes_hexa in the compensated-FP32 form: up to 64 closed nodes staged at
// once (centres shifted to the bucket frame in FP64, the openSoftening
// decision in FP64), each chunk's FP32 partial sums added to the Kahan
// accumulators.
template <typename W_t>
__device__ __forceinline__ void do_nodes_hexa32(W_t& W, int32_t id, int nsrc, const Group& g,
                                                Lanes<true>& L, const HexNode* __restrict__ hn,
                                                const Bucket& b, double ox, double oy, double oz,
                                                double ccx, double ccy, double ccz) {
  const int lane = __lane_id();
  SrcH32* st = reinterpret_cast<SrcH32*>(W.stage);
  if (lane < nsrc) {
    const bool far = (id & kFarNode) != 0;
    const HexNode& s = hn[id & ~kFarNode];
    const double cx = s.cx + ox, cy = s.cy + oy, cz = s.cz + oz;
    SrcH32 r;
    r.cx = (float)(cx - ccx);
    r.cy = (float)(cy - ccy);
    r.cz = (float)(cz - ccz);
    r.mass = (float)s.mass;
    r.mono = !far && pt_open_softening(cx, cy, cz, s.soft, b) ? 1.f : 0.f;
#pragma unroll
    for (int k = 0; k < PT_HEXA_MOMS; k++) r.h[k] = (float)s.h[k];
    st[lane] = r;
  }
  __builtin_amdgcn_wave_barrier();
  float px = 0.f, py = 0.f, pz = 0.f, pp = 0.f;
  const int kend = ceil_div_s(nsrc, g);
  for (int k = 0; k < kend; k++) {
    const int i = g.s + k * g.S;
    if (i < nsrc) hexa_pair32(st[i], L, px, py, pz, pp);
  }
  kahan_add(L.ax, L.cx, px);
  kahan_add(L.ay, L.cy, py);
  kahan_add(L.az, L.cz, pz);
  kahan_add(L.ap, L.cp, pp);
  __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(kThreads)
interact_hexa_kernel(const Bucket* __restrict__ bk, int64_t nb,
                     const TreeNode* __restrict__ nodes, const HexNode* __restrict__ hn,
                     const double* __restrict__ px, const double* __restrict__ py,
                     const double* __restrict__ pz, const double* __restrict__ pm,
                     const double* __restrict__ pe, const int64_t* __restrict__ pp_off,
                     const int32_t* __restrict__ pp, const int64_t* __restrict__ pc_off,
                     const int32_t* __restrict__ pc, double ox, double oy, double oz,
                     double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
                     double* __restrict__ pot) {
  __shared__ HexaLds lds[kWaves];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t bi = pt_xcd_block(blockIdx.x, gridDim.x) * kWaves + wave;
  if (bi >= nb) return;
  HexaLds& W = lds[wave];
  const Bucket b = bk[bi];
  const double ccx = (b.lox + b.hix) * 0.5, ccy = (b.loy + b.hiy) * 0.5,
               ccz = (b.loz + b.hiz) * 0.5;
  const BucketF bf = bucket_frame(b, ccx, ccy, ccz);
  const int64_t pc0 = pc_off[bi], pc1 = pc_off[bi + 1];
  const int64_t pp0 = pp_off[bi], pp1 = pp_off[bi + 1];
  for (int g0 = 0; g0 < b.count; g0 += kGroupMax) {
    const Group g = make_group(b, g0);
    Lanes<false> L;
    load_targets<false>(L, g, px, py, pz, pe, ccx, ccy, ccz);
    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int nsrc = (pc1 - e0) < PT_WAVE ? (int)(pc1 - e0) : PT_WAVE;
      const int32_t id = lane < nsrc ? pc[e0 + lane] : 0;
      do_nodes_hexa<HexaLds>(W, id, nsrc, g, L, hn, b, ox, oy, oz);
    }
    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      int32_t beg = 0, cnt = 0;
      if (l0 + lane < pp1) {
        const TreeNode nd = nodes[pp[l0 + lane]];
        beg = nd.begin;
        cnt = leaf_entry(nd.count, nd.cx, nd.cy, nd.cz, nd.leaf_reach, nd.leaf_soft, bf, L.emax,
                         ox, oy, oz, ccx, ccy, ccz);
      }
      do_leaves<false, HexaLds>(W, beg, cnt, g, L, px, py, pz, pm, pe, bf, ox, oy, oz, ccx, ccy,
                                ccz);
    }
    finish_group<false, HexaLds>(W, g, L, ax, ay, az, pot);
  }
}

// Fused walk + interactions of the HEXADECAPOLE build (FP64): the walk with
// the hexadecapole open(), closed nodes evaluated with the reduced moments,
// leaves with leaf(); one wave per bucket, chunks evaluated as they fill.
template <int STACK>
struct __align__(16) HexaFusedLds {
  static constexpr int kStage = HexaLds::kStage;
  static_assert(PT_WAVE * sizeof(SrcH32) <= kStage, "FP32 node chunk");
  unsigned char stage[kStage];
  int32_t stack[STACK];
  int32_t pcbuf[2 * PT_WAVE];
  int32_t ppbeg[2 * PT_WAVE];
  int32_t ppcnt[2 * PT_WAVE];
  int32_t fl_mark[PT_WAVE];
};

template <int STACK, bool F32>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(STACK == kSmallStack ? (F32 ? kHexaWaves32 : kHexaWaves) : 1, 8)))
hexa_gravity_kernel(const Bucket* __restrict__ bk, int64_t nb, const TreeNode* __restrict__ nodes,
                    const HexOpen* __restrict__ ho, const HexNode* __restrict__ hn,
                    const double* __restrict__ px, const double* __restrict__ py,
                    const double* __restrict__ pz, const double* __restrict__ pm,
                    const double* __restrict__ pe, double gf, double gm, double ox, double oy,
                    double oz, double* __restrict__ ax, double* __restrict__ ay,
                    double* __restrict__ az, double* __restrict__ pot,
                    unsigned long long* __restrict__ partials, int32_t* __restrict__ err, int acc) {
  using W_t = HexaFusedLds<STACK>;
  __shared__ W_t lds[kWaves];
  __shared__ unsigned long long s_cnt[kWaves][4];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  W_t& W = lds[wave];
  unsigned long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  bool overflow = false;
  const int64_t bi = pt_xcd_block(blockIdx.x, gridDim.x) * kWaves + wave;
  if (bi < nb) {
    const Bucket b = bk[bi];
    const int64_t nt = b.count;
    const double ccx = (b.lox + b.hix) * 0.5, ccy = (b.loy + b.hiy) * 0.5,
                 ccz = (b.loz + b.hiz) * 0.5;
    const BucketF bf = bucket_frame(b, ccx, ccy, ccz);
    for (int g0 = 0; g0 < b.count; g0 += kGroupMax) {
      const Group g = make_group(b, g0);
      const bool count = g0 == 0;  // the walk repeats per target group; count once
      Lanes<F32> L;
      load_targets<F32>(L, g, px, py, pz, pe, ccx, ccy, ccz);
      int sp = 0;
      int npc = 0, npp = 0;
      WalkCounts wc;
      {
        const WalkLane r = pt_walk_root<true>(nodes, b, gf, ox, oy, oz, W.stack, sp, wc, ho, gm);
        if (r.is_closed) {
          if (lane == 0) W.pcbuf[0] = 0;
          npc = 1;
        } else if (r.is_leaf) {
          if (lane == 0) {
            W.ppbeg[0] = r.begin;
            W.ppcnt[0] = r.count;
          }
          npp = 1;
        }
        if (count) {
          c_open += wc.n_open;
          c_close += wc.n_close;
          c_node += (unsigned long long)wc.n_close * nt;
          c_part += (unsigned long long)wc.leaf_parts * nt;
        }
      }
      __builtin_amdgcn_wave_barrier();
      while (sp > 0) {
        unsigned long long leafm, closem;
        const WalkLane r = pt_walk_step<STACK, true>(nodes, b, gf, ox, oy, oz, W.stack, sp,
                                                     overflow, wc, leafm, closem, ho, gm);
        if (count) {
          c_open += wc.n_open;
          c_close += wc.n_close;
          c_node += (unsigned long long)wc.n_close * nt;
          c_part += (unsigned long long)wc.leaf_parts * nt;
        }
        if (r.is_closed) W.pcbuf[npc + pt_rank_in(closem)] = r.node;
        if (r.is_leaf) {
          const int k = npp + pt_rank_in(leafm);
          W.ppbeg[k] = r.begin;
          W.ppcnt[k] = leaf_far_dsq(r.dsq, r.leaf_reach, r.leaf_soft, bf, L.emax)
                           ? (r.count | kFarLeaf)
                           : r.count;
        }
        npc += __popcll(closem);
        npp += __popcll(leafm);
        __builtin_amdgcn_wave_barrier();
        if (npc >= PT_WAVE) {
          if constexpr (F32)
            do_nodes_hexa32<W_t>(W, W.pcbuf[lane], PT_WAVE, g, L, hn, b, ox, oy, oz, ccx, ccy,
                                 ccz);
          else
            do_nodes_hexa<W_t>(W, W.pcbuf[lane], PT_WAVE, g, L, hn, b, ox, oy, oz);
          const int rest = npc - PT_WAVE;
          const int32_t v = lane < rest ? W.pcbuf[PT_WAVE + lane] : 0;
          __builtin_amdgcn_wave_barrier();
          if (lane < rest) W.pcbuf[lane] = v;
          npc = rest;
          __builtin_amdgcn_wave_barrier();
        }
        if (npp >= PT_WAVE) {
          do_leaves<F32, W_t>(W, W.ppbeg[lane], W.ppcnt[lane], g, L, px, py, pz, pm, pe, bf, ox,
                                oy, oz, ccx, ccy, ccz);
          const int rest = npp - PT_WAVE;
          const int32_t vb = lane < rest ? W.ppbeg[PT_WAVE + lane] : 0;
          const int32_t vc = lane < rest ? W.ppcnt[PT_WAVE + lane] : 0;
          __builtin_amdgcn_wave_barrier();
          if (lane < rest) {
            W.ppbeg[lane] = vb;
            W.ppcnt[lane] = vc;
          }
          npp = rest;
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (npc > 0) {
        const int32_t id = lane < npc ? W.pcbuf[lane] : 0;
        if constexpr (F32)
          do_nodes_hexa32<W_t>(W, id, npc, g, L, hn, b, ox, oy, oz, ccx, ccy, ccz);
        else
          do_nodes_hexa<W_t>(W, id, npc, g, L, hn, b, ox, oy, oz);
      }
      if (npp > 0) {
        const int32_t beg = lane < npp ? W.ppbeg[lane] : 0;
        const int32_t cnt = lane < npp ? W.ppcnt[lane] : 0;
        do_leaves<F32, W_t>(W, beg, cnt, g, L, px, py, pz, pm, pe, bf, ox, oy, oz, ccx, ccy,
                              ccz);
      }
      finish_group<F32, W_t>(W, g, L, ax, ay, az, pot, acc != 0);
    }
  }
  if (lane == 0) {
    if (overflow) atomicExch(err, 1);
    s_cnt[wave][0] = c_node;
    s_cnt[wave][1] = c_part;
    s_cnt[wave][2] = c_open;
    s_cnt[wave][3] = c_close;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long v = 0;
    for (int w = 0; w < kWaves; w++) v += s_cnt[w][threadIdx.x];
    partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = v;
  }
}

// The LDS working set of a fused wave: the quadrupole build's, or the
// HEXADECAPOLE build's (larger node records).
template <bool F32, int STACK, bool HEXA>
struct FusedLdsOf {
  using type = WaveLds<F32, STACK>;
};
template <bool F32, int STACK>
struct FusedLdsOf<F32, STACK, true> {
  using type = HexaFusedLds<STACK>;
};

// Stages and evaluates up to one chunk of closed nodes (node id, with the
// kFarNode flag, in the calling lane) in the build's form.
template <bool F32, bool HEXA, typename W_t>
__device__ __forceinline__ void stage_nodes(W_t& W, int32_t id, int nsrc, const Group& g,
                                            Lanes<F32>& L, const GravNode* __restrict__ gn,
                                            const GravNode32* __restrict__ gn32,
                                            const HexNode* __restrict__ hn, const Bucket& b,
                                            const BucketF& bf, double ox, double oy, double oz,
                                            double ccx, double ccy, double ccz) {
  if constexpr (HEXA) {
    if constexpr (F32)
      do_nodes_hexa32<W_t>(W, id, nsrc, g, L, hn, b, ox, oy, oz, ccx, ccy, ccz);
    else
      do_nodes_hexa<W_t>(W, id, nsrc, g, L, hn, b, ox, oy, oz);
  } else {
    do_nodes<F32, W_t>(W, id, nsrc, g, L, gn, gn32, b, bf, ox, oy, oz, ccx, ccy, ccz);
  }
}

// Evaluates every full chunk of the pending closed nodes (cw at a time) and
// leaf sources (64 at a time).  A walk step starts with npc < cw and
// npp < 64 and adds at most 64 to each, so the buffers (128 entries) never
// overflow and each rest (< 64) fits one move.
template <bool F32, bool HEXA, typename W_t>
__device__ __forceinline__ void drain_full(W_t& W, int& npc, int& npp, const Group& g,
                                           Lanes<F32>& L, const GravNode* __restrict__ gn,
                                           const GravNode32* __restrict__ gn32,
                                           const HexNode* __restrict__ hn, const Bucket& b,
                                           const BucketF& bf, const double* __restrict__ px,
                                           const double* __restrict__ py,
                                           const double* __restrict__ pz,
                                           const double* __restrict__ pm,
                                           const double* __restrict__ pe, double ox, double oy,
                                           double oz, double ccx, double ccy, double ccz) {
  const int lane = __lane_id();
  const int cw = (F32 || HEXA) ? PT_WAVE : g.cw;
  while (npc >= cw) {
    stage_nodes<F32, HEXA, W_t>(W, lane < cw ? W.pcbuf[lane] : 0, cw, g, L, gn, gn32, hn, b, bf,
                                ox, oy, oz, ccx, ccy, ccz);
    const int rest = npc - cw;
    const int32_t v = lane < rest ? W.pcbuf[cw + lane] : 0;
    __builtin_amdgcn_wave_barrier();
    if (lane < rest) W.pcbuf[lane] = v;
    npc = rest;
    __builtin_amdgcn_wave_barrier();
  }
  if (npp >= PT_WAVE) {
    do_leaves<F32, W_t>(W, W.ppbeg[lane], W.ppcnt[lane], g, L, px, py, pz, pm, pe, bf, ox, oy, oz,
                        ccx, ccy, ccz);
    const int rest = npp - PT_WAVE;
    const int32_t vb = lane < rest ? W.ppbeg[PT_WAVE + lane] : 0;
    const int32_t vc = lane < rest ? W.ppcnt[PT_WAVE + lane] : 0;
    __builtin_amdgcn_wave_barrier();
    if (lane < rest) {
      W.ppbeg[lane] = vb;
      W.ppcnt[lane] = vc;
    }
    npp = rest;
    __builtin_amdgcn_wave_barrier();
  }
}

// Evaluates what is left in the buffers.
template <bool F32, bool HEXA, typename W_t>
__device__ __forceinline__ void drain_rest(W_t& W, int npc, int npp, const Group& g,
                                           Lanes<F32>& L, const GravNode* __restrict__ gn,
                                           const GravNode32* __restrict__ gn32,
                                           const HexNode* __restrict__ hn, const Bucket& b,
                                           const BucketF& bf, const double* __restrict__ px,
                                           const double* __restrict__ py,
                                           const double* __restrict__ pz,
                                           const double* __restrict__ pm,
                                           const double* __restrict__ pe, double ox, double oy,
                                           double oz, double ccx, double ccy, double ccz) {
  const int lane = __lane_id();
  if (npc > 0) {
    const int32_t id = lane < npc ? W.pcbuf[lane] : 0;
    stage_nodes<F32, HEXA, W_t>(W, id, npc, g, L, gn, gn32, hn, b, bf, ox, oy, oz, ccx, ccy, ccz);
  }
  if (npp > 0) {
    const int32_t beg = lane < npp ? W.ppbeg[lane] : 0;
    const int32_t cnt = lane < npp ? W.ppcnt[lane] : 0;
    do_leaves<F32, W_t>(W, beg, cnt, g, L, px, py, pz, pm, pe, bf, ox, oy, oz, ccx, ccy, ccz);
  }
}

// The walk of one bucket below the opened nodes on the stack ([0, sp); the
// entries from `limit` up are kept), its chunks evaluated as they fill.
// w_open / w_close: this walk's open() outcomes; lane_leaf_parts: the
// source-leaf particles this lane visited.
template <bool F32, int STACK, bool HEXA>
__device__ __forceinline__ void walk_below(
    typename FusedLdsOf<F32, STACK, HEXA>::type& W, int& sp, int& npc, int& npp, int limit, const Bucket& b,
    const BucketF& bf, const Group& g, Lanes<F32>& L, const TreeNode* __restrict__ nodes,
    const GravNode* __restrict__ gn, const GravNode32* __restrict__ gn32,
    const HexOpen* __restrict__ ho, const HexNode* __restrict__ hn, double gm,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ pm, const double* __restrict__ pe, double gf, double ox,
    double oy, double oz, double ccx, double ccy, double ccz, int& w_open, int& w_close,
    int& lane_leaf_parts, bool& overflow) {
  using W_t = typename FusedLdsOf<F32, STACK, HEXA>::type;
  WalkCounts wc;
  while (sp > 0) {
    unsigned long long leafm, closem;
    const WalkLane r = pt_walk_step<STACK, HEXA, false>(nodes, b, gf, ox, oy, oz, W.stack, sp,
                                                        overflow, wc, leafm, closem, ho, gm, limit);
    w_open += wc.n_open;
    w_close += wc.n_close;
    lane_leaf_parts += r.is_leaf ? r.count : 0;
    if (r.is_closed)
      W.pcbuf[npc + pt_rank_in(closem)] =
          node_far_dsq(r.dsq, r.node_soft, bf, L.emax) ? (r.node | kFarNode) : r.node;
    if (r.is_leaf) {
      const int k = npp + pt_rank_in(leafm);
      W.ppbeg[k] = r.begin;
      W.ppcnt[k] = leaf_far_dsq(r.dsq, r.leaf_reach, r.leaf_soft, bf, L.emax)
                       ? (r.count | kFarLeaf)
                       : r.count;
    }
    npc += __popcll(closem);
    npp += __popcll(leafm);
    __builtin_amdgcn_wave_barrier();
    drain_full<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, b, bf, px, py, pz, pm, pe, ox, oy, oz, ccx,
                         ccy, ccz);
  }
}

// One target bucket of the fused kernel: the walk for the bucket box with
// the bucket's chunks evaluated as they fill, counters added to the wave's.
template <bool F32, int STACK, bool HEXA>
__device__ __forceinline__ void gravity_bucket(
    typename FusedLdsOf<F32, STACK, HEXA>::type& W, int64_t bi, const Bucket* __restrict__ bk,
    const TreeNode* __restrict__ nodes, const GravNode* __restrict__ gn,
    const GravNode32* __restrict__ gn32,
    const HexOpen* __restrict__ ho, const HexNode* __restrict__ hn, double gm, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ pm,
    const double* __restrict__ pe, double gf, double ox, double oy, double oz,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ pot, unsigned long long& c_node, unsigned long long& c_part,
    unsigned long long& c_open, unsigned long long& c_close, bool& overflow, bool acc) {
  using W_t = typename FusedLdsOf<F32, STACK, HEXA>::type;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const Bucket b = bk[bi];
  const int64_t nt = b.count;
  const double ccx = (b.lox + b.hix) * 0.5, ccy = (b.loy + b.hiy) * 0.5,
               ccz = (b.loz + b.hiz) * 0.5;
  const BucketF bf = bucket_frame(b, ccx, ccy, ccz);
  for (int g0 = 0; g0 < b.count; g0 += kGroupMax) {
    const Group g = make_group(b, g0);
    const bool count = g0 == 0;  // the walk repeats per target group; count once
    Lanes<F32> L;
    load_targets<F32>(L, g, px, py, pz, pe, ccx, ccy, ccz);
    int sp = 0;
    int npc = 0, npp = 0;
    int w_open = 0, w_close = 0;
    int lane_leaf_parts = 0;
    {
      WalkCounts wc;
      const WalkLane r = pt_walk_root<HEXA>(nodes, b, gf, ox, oy, oz, W.stack, sp, wc, ho, gm);
      if (r.is_closed) {
        if (lane == 0) W.pcbuf[0] = 0;
        npc = 1;
      } else if (r.is_leaf) {
        if (lane == 0) {
          W.ppbeg[0] = r.begin;
          W.ppcnt[0] = r.count;
        }
        npp = 1;
      }
      w_open = wc.n_open;
      w_close = wc.n_close;
      lane_leaf_parts = lane == 0 ? wc.leaf_parts : 0;
    }
    __builtin_amdgcn_wave_barrier();
    walk_below<F32, STACK, HEXA>(W, sp, npc, npp, STACK, b, bf, g, L, nodes, gn, gn32, ho, hn, gm, px, py, pz, pm,
                           pe, gf, ox, oy, oz, ccx, ccy, ccz, w_open, w_close, lane_leaf_parts,
                           overflow);
    drain_rest<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, b, bf, px, py, pz, pm, pe, ox, oy, oz, ccx,
                         ccy, ccz);
    if (count) {
      c_open += (unsigned)w_open;
      c_close += (unsigned)w_close;
      c_node += (unsigned long long)w_close * nt;
      c_part += (unsigned long long)pt_wave_sum_int(lane_leaf_parts) * nt;
    }
    finish_group<F32, W_t>(W, g, L, ax, ay, az, pot, acc);
  }
}

// ---------------------------------------------------------------------------
// Grouped fused evaluation: one wave per run of consecutive buckets (at most
// 64 targets; formed by fgroup_window_kernel).  The group first walks once
// with its union box U (the members' boxes' hull):
//   - a node whose opening sphere misses U is closed for every member, since
//     the rounded box distance to a member box is never below the one to U
//     (pt_box_point_far_dsq states the argument); one that also holds U's
//     farthest corner is opened by every member, as is a node of at most
//     PT_MIN_PARTICLE_NODE particles; otherwise each member's own test
//     decides, and a node opened by all or closed by all is still shared;
//   - a shared closed node that is far (node_far_dsq on U: no openSoftening
//     and no softened pair for any member) is a node() source of every
//     member; a leaf reached by the shared walk is a leaf() source of every
//     member;
//   - every other node is a residual node, kept at the top of the walk
//     stack.
// The shared sources are evaluated once for all the group's targets (frame:
// U's centre).  Then each member bucket applies its own open() to every
// residual node (opened -> its children are walked below it as in the
// per-bucket walk; closed -> node() with the member's own openSoftening) and
// adds the result to its targets' shared sums.  Each member thereby makes
// exactly the calls of its own walk from the root.
// ---------------------------------------------------------------------------

// Target group over particles [first, first + count), count <= 64.
__device__ __forceinline__ Group make_group_range(int64_t first, int count) {
  Group g;
  g.ng = count;
  g.S = PT_WAVE / g.ng;
  g.magic = (65536 + g.S - 1) / g.S;
  g.cw = g.S * ((PT_WAVE * g.magic) >> 16);
  const int lane = __lane_id();
  const int s = lane / g.ng;
  g.act = s < g.S;
  g.s = g.act ? s : 0;
  g.t = g.act ? lane - s * g.ng : 0;
  g.first = first;
  return g;
}

// finish_group for a member whose targets start at offset `off` of a group
// target range whose sums its lanes off + t wrote: the member's sums are moved
// to those lanes, each of which adds to what it wrote itself.
template <bool F32, typename W_t>
__device__ __forceinline__ void finish_member(W_t& W, const Group& g, const Lanes<F32>& L, int off,
                                              double* __restrict__ ax, double* __restrict__ ay,
                                              double* __restrict__ az,
                                              double* __restrict__ pot) {
  double* red = reinterpret_cast<double*>(W.stage);
  const int lane = __lane_id();
  double v0, v1, v2, v3;
  if constexpr (F32) {
    v0 = (double)L.ax - (double)L.cx;
    v1 = (double)L.ay - (double)L.cy;
    v2 = (double)L.az - (double)L.cz;
    v3 = (double)L.ap - (double)L.cp;
  } else {
    v0 = L.ax;
    v1 = L.ay;
    v2 = L.az;
    v3 = L.ap;
  }
  red[lane * 4 + 0] = g.act ? v0 : 0.0;
  red[lane * 4 + 1] = g.act ? v1 : 0.0;
  red[lane * 4 + 2] = g.act ? v2 : 0.0;
  red[lane * 4 + 3] = g.act ? v3 : 0.0;
  __builtin_amdgcn_wave_barrier();
  const int t = lane - off;
  if (t >= 0 && t < g.ng) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int s = 0; s < g.S; s++) {
      const int l = s * g.ng + t;
      s0 += red[l * 4 + 0];
      s1 += red[l * 4 + 1];
      s2 += red[l * 4 + 2];
      s3 += red[l * 4 + 3];
    }
    const int64_t i = g.first + t;
    ax[i] += s0;
    ay[i] += s1;
    az[i] += s2;
    pot[i] += s3;
  }
  __builtin_amdgcn_wave_barrier();
}

// Walk stack entries the residual lists leave to the member walks.  A shared
// walk whose residual list would take more gives up (see gravity_group).
constexpr int kMemberReserve = 96;
// Particles of a sub-run: after the group's shared walk, runs of consecutive
// members of at most this many particles walk once more together below the
// group's residual nodes.
constexpr int kSubTargets = 16;

// One shared-walk visit: each active lane's node is classified for every
// member of the run of buckets mb (mG of them, union box ub) and lands in
// the pending stack (opened by all), the closed-node buffer (closed by all
// and far for every target), the leaf buffer (a leaf), or the residual list
// growing down from res_top (anything else).  used_top: entries above
// res_top held by enclosing residual lists.  Returns false, writing nothing,
// when the entries would not fit with kMemberReserve to spare.
template <bool F32, int STACK, bool HEXA>
__device__ __forceinline__ bool shared_visit(
    typename FusedLdsOf<F32, STACK, HEXA>::type& W, int32_t node, bool active, const TreeNode* __restrict__ nodes,
    const Bucket* __restrict__ mb, int mG, const Bucket& ub, const BucketF& ubf, float emax,
    double gf, const HexOpen* __restrict__ ho, double gm, double ox, double oy, double oz,
    int& sp, int& npc, int& npp, int& nres,
    int res_top, int& sh_open, int& sh_close, int& lane_leaf_parts) {
  const TreeNode nd = nodes[node];
  const bool leaf = active && nd.type == PT_NODE_LEAF;
  const bool internal = active && nd.type == PT_NODE_INTERNAL;
  const double cx = nd.cx + ox, cy = nd.cy + oy, cz = nd.cz + oz;
  const double dsq = pt_box_point_dsq(ub, cx, cy, cz);
  bool op = false, cl = false;
  if (internal) {
    if (nd.count <= PT_MIN_PARTICLE_NODE) {
      op = true;
    } else {
      const double rsq = nd.radius * nd.radius * gf;
      // HEXADECAPOLE: a member also opens the node when the softening spheres
      // meet and the monopole sphere (radius * gm) reaches its box, so only a
      // node beyond both spheres for U is closed for all
      double msq = 0.0;
      if constexpr (HEXA) {
        const double mono_rad = nd.radius * gm;
        msq = mono_rad * mono_rad;
      }
      if (dsq > rsq && (!HEXA || dsq > msq)) {
        cl = true;
      } else if (pt_box_point_far_dsq(ub, cx, cy, cz) <= rsq) {
        op = true;
      } else {
        bool any_open = false, any_closed = false;
        for (int k = 0; k < mG; k++) {
          bool o;
          if constexpr (HEXA) o = pt_gv_open_hexa(nd, ho[node], mb[k], gf, gm, ox, oy, oz);
          else o = pt_box_point_dsq(mb[k], cx, cy, cz) <= rsq;
          any_open |= o;
          any_closed |= !o;
        }
        op = !any_closed;
        cl = !any_open;
      }
    }
  }
  const bool shared_close = cl && node_far_dsq(dsq, nd.node_soft, ubf, emax);
  const bool res = internal && !op && !shared_close;
  const unsigned long long leafm = __ballot(leaf), closem = __ballot(shared_close),
                           openm = __ballot(op), resm = __ballot(res);
  const int n_open = __popcll(openm), n_res = __popcll(resm);
  const int held = STACK - res_top + nres + n_res;  // residual entries after this visit
  if (sp + n_open + held > STACK || held > STACK - kMemberReserve) return false;
  __builtin_amdgcn_wave_barrier();
  if (op) W.stack[sp + pt_rank_in(openm)] = nd.first_child;
  if (res) W.stack[res_top - 1 - (nres + pt_rank_in(resm))] = node;
  if (shared_close) W.pcbuf[npc + pt_rank_in(closem)] = node | kFarNode;
  if (leaf) {
    const int k = npp + pt_rank_in(leafm);
    W.ppbeg[k] = nd.begin;
    W.ppcnt[k] = leaf_far_dsq(dsq, nd.leaf_reach, nd.leaf_soft, ubf, emax) ? (nd.count | kFarLeaf)
                                                                            : nd.count;
  }
  lane_leaf_parts += leaf ? nd.count : 0;
  sp += n_open;
  nres += n_res;
  npc += __popcll(closem);
  npp += __popcll(leafm);
  sh_open += n_open;
  sh_close += __popcll(closem);
  __builtin_amdgcn_wave_barrier();
  return true;
}

// The shared walk of the run of buckets bk[mb0, mb0 + mG) (mG >= 2, their
// particles [first, first + ntot), ntot <= 64), from the root (nseed < 0)
// or from the nseed residual nodes below seed_top (every member visits
// them); residual nodes go below res_top into nres.  On success the shared
// sums are written (base < 0: finish_group, stored or added per acc) or
// added by the lanes base + t that wrote the group's sums (base >= 0), and
// the counters take every member's shared calls.  Returns false with nothing
// written or counted when the lists overflow.
template <bool F32, int STACK, bool HEXA>
__device__ __forceinline__ bool shared_walk(
    typename FusedLdsOf<F32, STACK, HEXA>::type& W, const Bucket* __restrict__ bk, int64_t mb0, int mG, int seed_top,
    int nseed, int res_top, int& nres, int base, const TreeNode* __restrict__ nodes,
    const GravNode* __restrict__ gn, const GravNode32* __restrict__ gn32,
    const HexOpen* __restrict__ ho, const HexNode* __restrict__ hn, double gm,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ pm, const double* __restrict__ pe, double gf, double ox,
    double oy, double oz, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, double* __restrict__ pot, unsigned long long& c_node,
    unsigned long long& c_part, unsigned long long& c_open, unsigned long long& c_close,
    bool acc) {
  using W_t = typename FusedLdsOf<F32, STACK, HEXA>::type;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const Bucket* mb = bk + mb0;
  Bucket ub = mb[0];
  for (int k = 1; k < mG; k++) {
    const Bucket& m = mb[k];
    ub.lox = fmin(ub.lox, m.lox);
    ub.loy = fmin(ub.loy, m.loy);
    ub.loz = fmin(ub.loz, m.loz);
    ub.hix = fmax(ub.hix, m.hix);
    ub.hiy = fmax(ub.hiy, m.hiy);
    ub.hiz = fmax(ub.hiz, m.hiz);
    ub.soft = fmax(ub.soft, m.soft);
  }
  const int64_t first = ub.begin;
  const int ntot = (int)((int64_t)mb[mG - 1].begin + mb[mG - 1].count - first);
  const double gcx = (ub.lox + ub.hix) * 0.5, gcy = (ub.loy + ub.hiy) * 0.5,
               gcz = (ub.loz + ub.hiz) * 0.5;
  const BucketF ubf = bucket_frame(ub, gcx, gcy, gcz);
  const Group g = make_group_range(first, ntot);
  Lanes<F32> L;
  load_targets<F32>(L, g, px, py, pz, pe, gcx, gcy, gcz);
  int sp = 0, npc = 0, npp = 0;
  int sh_open = 0, sh_close = 0, lane_leaf_parts = 0;
  nres = 0;
  // seeds 64 at a time (the root alone from the top), each batch walked to
  // the end
  const int nbatch = nseed < 0 ? 1 : nseed;
  for (int r0 = 0; r0 < nbatch; r0 += PT_WAVE) {
    const bool active = nseed < 0 ? lane == 0 : r0 + lane < nseed;
    const int32_t node = nseed < 0 ? 0 : W.stack[seed_top - 1 - (active ? r0 + lane : r0)];
    if (!shared_visit<F32, STACK, HEXA>(W, node, active, nodes, mb, mG, ub, ubf, L.emax, gf, ho, gm, ox, oy, oz,
                                  sp, npc, npp, nres, res_top, sh_open, sh_close,
                                  lane_leaf_parts))
      return false;
    drain_full<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, ub, ubf, px, py, pz, pm, pe, ox, oy, oz,
                         gcx, gcy, gcz);
    while (sp > 0) {
      const int k = sp < 8 ? sp : 8;
      sp -= k;
      const int slot = lane >> 3;
      const bool act = slot < k;
      const int32_t child = W.stack[sp + (act ? slot : 0)] + (lane & 7);
      if (!shared_visit<F32, STACK, HEXA>(W, child, act, nodes, mb, mG, ub, ubf, L.emax, gf, ho, gm, ox, oy,
                                    oz, sp, npc, npp, nres, res_top, sh_open, sh_close,
                                    lane_leaf_parts))
        return false;
      drain_full<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, ub, ubf, px, py, pz, pm, pe, ox, oy, oz,
                           gcx, gcy, gcz);
    }
  }
  drain_rest<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, ub, ubf, px, py, pz, pm, pe, ox, oy, oz, gcx,
                       gcy, gcz);
  // every member made each shared call
  c_open += (unsigned long long)sh_open * mG;
  c_close += (unsigned long long)sh_close * mG;
  c_node += (unsigned long long)sh_close * ntot;
  c_part += (unsigned long long)pt_wave_sum_int(lane_leaf_parts) * ntot;
  if (base < 0) finish_group<F32, W_t>(W, g, L, ax, ay, az, pot, acc);
  else finish_member<F32, W_t>(W, g, L, base, ax, ay, az, pot);
  return true;
}

// One group of G >= 2 buckets starting at bucket b0 (every member holds at
// most 63 particles, all of them at most 64): the group's shared walk, then
// per sub-run of at most kSubTargets particles a shared walk below the
// group's residual nodes, then each member below what is left residual
// (below the group's residual nodes when its sub-run is a single bucket or
// its shared walk gave up).  Returns false, having written nothing and
// counted nothing, when the group's shared walk gives up; the caller then
// walks each member on its own.
template <bool F32, int STACK, bool HEXA>
__device__ __forceinline__ bool gravity_group(
    typename FusedLdsOf<F32, STACK, HEXA>::type& W, int64_t b0, int G, const Bucket* __restrict__ bk,
    const TreeNode* __restrict__ nodes, const GravNode* __restrict__ gn,
    const GravNode32* __restrict__ gn32,
    const HexOpen* __restrict__ ho, const HexNode* __restrict__ hn, double gm, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ pm,
    const double* __restrict__ pe, double gf, double ox, double oy, double oz,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ pot, unsigned long long& c_node, unsigned long long& c_part,
    unsigned long long& c_open, unsigned long long& c_close, bool& overflow, bool acc) {
  using W_t = typename FusedLdsOf<F32, STACK, HEXA>::type;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t first = bk[b0].begin;
  int nres = 0;
  if (!shared_walk<F32, STACK, HEXA>(W, bk, b0, G, STACK, -1, STACK, nres, -1, nodes, gn, gn32, ho, hn, gm, px, py,
                               pz, pm, pe, gf, ox, oy, oz, ax, ay, az, pot, c_node, c_part, c_open,
                               c_close, acc))
    return false;
  if (nres == 0) return true;
  for (int k0 = 0, k1 = 0; k0 < G; k0 = k1) {
    // the sub-run [k0, k1)
    int tot = bk[b0 + k0].count;
    for (k1 = k0 + 1; k1 < G && tot + bk[b0 + k1].count <= kSubTargets; k1++)
      tot += bk[b0 + k1].count;
    int seed_top = STACK, nseed = nres;
    if (k1 - k0 >= 2 && k1 - k0 < G) {
      int nres2 = 0;
      if (shared_walk<F32, STACK, HEXA>(W, bk, b0 + k0, k1 - k0, STACK, nres, STACK - nres, nres2,
                                  (int)(bk[b0 + k0].begin - first), nodes, gn, gn32, ho, hn, gm, px, py, pz,
                                  pm, pe, gf, ox, oy, oz, ax, ay, az, pot, c_node, c_part, c_open,
                                  c_close, acc)) {
        seed_top = STACK - nres;
        nseed = nres2;
      }
    }
    if (nseed == 0) continue;
    // ---- each member of the sub-run from the residual nodes it was left
    for (int k = k0; k < k1; k++) {
      const Bucket b = bk[b0 + k];
      const double ccx = (b.lox + b.hix) * 0.5, ccy = (b.loy + b.hiy) * 0.5,
                   ccz = (b.loz + b.hiz) * 0.5;
      const BucketF bf = bucket_frame(b, ccx, ccy, ccz);
      const Group g = make_group(b, 0);
      Lanes<F32> L;
      load_targets<F32>(L, g, px, py, pz, pe, ccx, ccy, ccz);
      const int limit = seed_top - nseed;  // >= kMemberReserve
      int sp = 0, npc = 0, npp = 0;
      int w_open = 0, w_close = 0, lane_leaf_parts = 0;
      // 64 residual nodes at a time, each batch walked to the end, so the
      // member's stack holds at most one batch plus what its walk pushes
      for (int r0 = 0; r0 < nseed; r0 += PT_WAVE) {
        const bool active = r0 + lane < nseed;
        const int32_t node = W.stack[seed_top - 1 - (active ? r0 + lane : r0)];
        const TreeNode nd = nodes[node];
        const double dsq = pt_box_point_dsq(b, nd.cx + ox, nd.cy + oy, nd.cz + oz);
        bool op;
        if constexpr (HEXA) op = active && pt_gv_open_hexa(nd, ho[node], b, gf, gm, ox, oy, oz);
        else op = active && pt_gv_open_dsq(nd, gf, dsq);
        const bool cl = active && !op;
        const unsigned long long openm = __ballot(op), closem = __ballot(cl);
        const int n_open = __popcll(openm);
        __builtin_amdgcn_wave_barrier();
        if (op) W.stack[pt_rank_in(openm)] = nd.first_child;
        if (cl)
          W.pcbuf[npc + pt_rank_in(closem)] =
              node_far_dsq(dsq, nd.node_soft, bf, L.emax) ? (node | kFarNode) : node;
        sp = n_open;
        npc += __popcll(closem);
        w_open += n_open;
        w_close += __popcll(closem);
        __builtin_amdgcn_wave_barrier();
        drain_full<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, b, bf, px, py, pz, pm, pe, ox, oy, oz,
                             ccx, ccy, ccz);
        walk_below<F32, STACK, HEXA>(W, sp, npc, npp, limit, b, bf, g, L, nodes, gn, gn32, ho, hn, gm, px, py, pz,
                               pm, pe, gf, ox, oy, oz, ccx, ccy, ccz, w_open, w_close,
                               lane_leaf_parts, overflow);
        if (overflow) return true;
      }
      drain_rest<F32, HEXA, W_t>(W, npc, npp, g, L, gn, gn32, hn, b, bf, px, py, pz, pm, pe, ox, oy, oz,
                           ccx, ccy, ccz);
      const int64_t nt = b.count;
      c_open += (unsigned)w_open;
      c_close += (unsigned)w_close;
      c_node += (unsigned long long)w_close * nt;
      c_part += (unsigned long long)pt_wave_sum_int(lane_leaf_parts) * nt;
      finish_member<F32, W_t>(W, g, L, (int)(b.begin - first), ax, ay, az, pot);
    }
  }
  return true;
}

// Fused walk + interactions, one wave per target bucket.  FP32C: one
// work-group per 4 buckets, mapped XCD-aware (pt_xcd_block).  FP64: a
// resident grid (a multiple of 8 work-groups, dispatched round-robin over the
// 8 XCDs) where XCD x owns the contiguous eighth [b_lo, b_hi) of the buckets
// and its waves take the next one from the XCD's counter (queue[x], zero at
// launch); every wave leaves once the counter has passed the range.  (The
// loop form needs more registers, which the FP32C kernel pays for with
// fewer resident waves; the FP64 kernel gains from the dynamic balance.)
// The FP32C kernel is compiled for six waves per SIMD (80 VGPRs, a 40-byte
// spill): 3.79 ms against 3.95 ms at five on the 1M Plummer.
template <bool F32, int STACK, bool OFF>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(F32 && STACK == kSmallStack ? 6 : 1, 8)))
gravity_kernel(const Bucket* __restrict__ bk, int64_t nb, const TreeNode* __restrict__ nodes,
               const GravNode* __restrict__ gn, const GravNode32* __restrict__ gn32,
               const double* __restrict__ px,
               const double* __restrict__ py, const double* __restrict__ pz,
               const double* __restrict__ pm, const double* __restrict__ pe, double gf,
               double ox_in, double oy_in, double oz_in, double* __restrict__ ax,
               double* __restrict__ ay, double* __restrict__ az, double* __restrict__ pot,
               unsigned long long* __restrict__ partials, int32_t* __restrict__ err,
               unsigned long long* __restrict__ queue, int acc) {
  using W_t = WaveLds<F32, STACK>;
  __shared__ W_t lds[kWaves];
  __shared__ unsigned long long s_cnt[kWaves][4];
  // without a replica offset the offset is -0.0, which every v + offset
  // folds away exactly (v + -0.0 == v for every v, -0.0 included)
  const double ox = OFF ? ox_in : -0.0, oy = OFF ? oy_in : -0.0, oz = OFF ? oz_in : -0.0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  W_t& W = lds[wave];
  unsigned long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  bool overflow = false;
  if constexpr (F32) {
    const int64_t bi = pt_xcd_block(blockIdx.x, gridDim.x) * kWaves + wave;
    if (bi < nb)
      gravity_bucket<F32, STACK, false>(W, bi, bk, nodes, gn, gn32, nullptr, nullptr, 0.0, px, py,
                                        pz, pm, pe, gf, ox, oy, oz,
                                 ax, ay, az, pot, c_node, c_part, c_open, c_close, overflow,
                                 acc != 0);
  } else {
    const int xcd = blockIdx.x % 8;
    const int64_t b_lo = nb * xcd / 8, b_hi = nb * (xcd + 1) / 8;
    unsigned long long first = 0;
    if (lane == 0) first = atomicAdd(&queue[xcd], 1ull);
    for (int64_t bi = b_lo + (int64_t)__shfl(first, 0, PT_WAVE); bi < b_hi;) {
      // the next index is fetched now and used after this bucket
      unsigned long long next = 0;
      if (lane == 0) next = atomicAdd(&queue[xcd], 1ull);
      gravity_bucket<F32, STACK, false>(W, bi, bk, nodes, gn, gn32, nullptr, nullptr, 0.0, px, py,
                                        pz, pm, pe, gf, ox, oy, oz,
                                 ax, ay, az, pot, c_node, c_part, c_open, c_close, overflow,
                                 acc != 0);
      bi = b_lo + (int64_t)__shfl(next, 0, PT_WAVE);
    }
  }
  if (lane == 0) {
    if (overflow) atomicExch(err, 1);
    s_cnt[wave][0] = c_node;
    s_cnt[wave][1] = c_part;
    s_cnt[wave][2] = c_open;
    s_cnt[wave][3] = c_close;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long v = 0;
    for (int w = 0; w < kWaves; w++) v += s_cnt[w][threadIdx.x];
    partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = v;
  }
}

// The grouped form of gravity_kernel: one wave per bucket group (gfirst /
// gnb, *ngroups of them), gravity_group for a group of two or more buckets,
// gravity_bucket for a group of one.  Work-group i runs on XCD i % 8, which
// takes the contiguous eighth [ng x / 8, ng (x + 1) / 8) of the groups in
// order (neighbouring groups share most sources in its L2); the grid is sized
// for an upper bound of the group count and the surplus waves find no group.
// Waves per work-group of the grouped kernel, and its walk stack (pending
// entries and, from the top, the residual list) per precision.  FP32C: 512
// entries (7.9 KiB of LDS per one-wave work-group, five per SIMD would fit);
// the kernel's ~105 VGPRs keep four waves per SIMD without spilling, which
// measured 3.5% faster than a five-wave register budget (95 VGPRs and a
// spill).  FP64: 384.
constexpr int kGWaves = 1;
constexpr int kGroupStack32 = 512;
constexpr int kGroupStack64 = 384;
constexpr int kGroupStackHexa = 384;  // the HEXADECAPOLE build, both forms

template <bool F32, int STACK, bool OFF, bool HEXA>
__global__ void __launch_bounds__(kGWaves * PT_WAVE)
grouped_gravity_kernel(const Bucket* __restrict__ bk, const int32_t* __restrict__ gfirst,
                       const int32_t* __restrict__ gnb, const uint32_t* __restrict__ ngroups,
                       const TreeNode* __restrict__ nodes, const GravNode* __restrict__ gn,
                       const GravNode32* __restrict__ gn32, const HexOpen* __restrict__ ho,
                       const HexNode* __restrict__ hn, const double* __restrict__ px,
                       const double* __restrict__ py, const double* __restrict__ pz,
                       const double* __restrict__ pm, const double* __restrict__ pe, double gf,
                       double gm, double ox_in, double oy_in, double oz_in, double* __restrict__ ax,
                       double* __restrict__ ay, double* __restrict__ az, double* __restrict__ pot,
                       unsigned long long* __restrict__ partials, int32_t* __restrict__ err,
                       int acc) {
  using W_t = typename FusedLdsOf<F32, STACK, HEXA>::type;
  __shared__ W_t lds[kGWaves];
  __shared__ unsigned long long s_cnt[kGWaves][4];
  const double ox = OFF ? ox_in : -0.0, oy = OFF ? oy_in : -0.0, oz = OFF ? oz_in : -0.0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  W_t& W = lds[wave];
  unsigned long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  bool overflow = false;
  const int64_t ng = (int64_t)ngroups[0];
  const int xcd = blockIdx.x % 8;
  const int64_t g_lo = ng * xcd / 8, g_hi = ng * (xcd + 1) / 8;
  const int64_t gi = g_lo + (int64_t)(blockIdx.x / 8) * kGWaves + wave;
  if (gi < g_hi) {
    const int G = gnb[gi];
    const int64_t b0 = gfirst[gi];
    // a group of one bucket, or a group whose residual list overflowed: one
    // walk per bucket
    const bool done =
        G > 1 && gravity_group<F32, STACK, HEXA>(W, b0, G, bk, nodes, gn, gn32, ho, hn, gm, px,
                                                 py, pz, pm, pe, gf, ox, oy, oz, ax, ay, az, pot,
                                                 c_node, c_part, c_open, c_close, overflow,
                                                 acc != 0);
    if (!done) {
      for (int k = 0; k < G; k++)
        gravity_bucket<F32, STACK, HEXA>(W, b0 + k, bk, nodes, gn, gn32, ho, hn, gm, px, py, pz,
                                         pm, pe, gf, ox, oy, oz, ax, ay, az, pot, c_node, c_part,
                                         c_open, c_close, overflow, acc != 0);
    }
  }
  if (lane == 0) {
    if (overflow) atomicExch(err, 1);
    s_cnt[wave][0] = c_node;
    s_cnt[wave][1] = c_part;
    s_cnt[wave][2] = c_open;
    s_cnt[wave][3] = c_close;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long v = 0;
    for (int w = 0; w < kGWaves; w++) v += s_cnt[w][threadIdx.x];
    partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = v;
  }
}

// Bucket groups of the grouped kernel, formed greedily inside windows of
// kFgWindow consecutive buckets: a group closes when the next bucket would
// take it past cap (<= 64) particles, so a bucket of more than cap is a
// group of its own.  One wave per window: the lanes load the window's counts, the greedy
// pass runs on wave-uniform values.  Window w writes its groups to slots
// [w kFgWindow, w kFgWindow + wcount[w]) of tfirst / tnb.
constexpr int kFgWindow = PT_WAVE;

__global__ void __launch_bounds__(256)
fgroup_window_kernel(const Bucket* __restrict__ bk, int64_t nb, int cap,
                     int32_t* __restrict__ tfirst, int32_t* __restrict__ tnb,
                     uint32_t* __restrict__ wcount) {
  const int64_t w = (int64_t)blockIdx.x * 4 + threadIdx.x / PT_WAVE;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t nw = (nb + kFgWindow - 1) / kFgWindow;
  if (w >= nw) return;
  const int64_t b0 = w * kFgWindow;
  const int n = (int)((nb - b0) < kFgWindow ? (nb - b0) : kFgWindow);
  const int cnt = lane < n ? bk[b0 + lane].count : 0;
  int slot = 0, start = 0, tot = 0;
  for (int k = 0; k < n; k++) {
    const int c = __builtin_amdgcn_readlane(cnt, k);
    if (k > start && tot + c > cap) {
      if (lane == 0) {
        tfirst[b0 + slot] = (int32_t)(b0 + start);
        tnb[b0 + slot] = k - start;
      }
      slot++;
      start = k;
      tot = 0;
    }
    tot += c;
  }
  if (lane == 0) {
    tfirst[b0 + slot] = (int32_t)(b0 + start);
    tnb[b0 + slot] = n - start;
    wcount[w] = (uint32_t)(slot + 1);
  }
}

// Compacts the windows' groups in order (woff: exclusive scan of wcount).
__global__ void __launch_bounds__(256)
fgroup_compact_kernel(int64_t nb, const int32_t* __restrict__ tfirst,
                      const int32_t* __restrict__ tnb, const uint32_t* __restrict__ wcount,
                      const uint32_t* __restrict__ woff, int32_t* __restrict__ gfirst,
                      int32_t* __restrict__ gnb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  const int64_t w = i / kFgWindow;
  const uint32_t j = (uint32_t)(i - w * kFgWindow);
  if (j < wcount[w]) {
    gfirst[woff[w] + j] = tfirst[i];
    gnb[woff[w] + j] = tnb[i];
  }
}

inline unsigned grid_for(int64_t nb) { return (unsigned)((nb + kWaves - 1) / kWaves); }

}  // namespace

// Work-groups of the fused kernel: FP32C one per 4 buckets; FP64 the
// resident grid, 8 per CU (at least what fits on a CU; work-groups that do
// not fit start as others finish -- none waits on another), a multiple of 8
// (one share per XCD), capped by the bucket count.
int64_t pt_gravity_blocks(int64_t nb, int precision) {
  const int64_t need = (nb + kWaves - 1) / kWaves;
  if (precision != 0) return need;
  static int64_t cap = 0;
  if (cap == 0) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    cap = (int64_t)(cus > 0 ? cus : 8) * 8;
  }
  int64_t g = need < cap ? need : cap;
  g = ((g + 7) / 8) * 8;
  return g < 8 ? 8 : g;
}

int pt_launch_interact(int precision, const Bucket* bk, int64_t nb, const TreeNode* nodes,
                       const GravNode* gn, const GravNode32* gn32, const double* x, const double* y, const double* z,
                       const double* m, const double* e, const int64_t* pp_off, const int32_t* pp,
                       const int64_t* pc_off, const int32_t* pc, const double off[3], double* ax,
                       double* ay, double* az, double* pot, hipStream_t s) {
  if (nb <= 0) return 0;
  if (precision == 0) {
    interact_kernel<false><<<grid_for(nb), kThreads, 0, s>>>(bk, nb, nodes, gn, gn32, x, y, z, m,
                                                             e, pp_off, pp, pc_off, pc, off[0],
                                                             off[1], off[2], ax, ay, az, pot);
  } else {
    interact_kernel<true><<<grid_for(nb), kThreads, 0, s>>>(bk, nb, nodes, gn, gn32, x, y, z, m,
                                                            e, pp_off, pp, pc_off, pc, off[0], off[1],
                                                            off[2], ax, ay, az, pot);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_interact_hexa(const Bucket* bk, int64_t nb, const TreeNode* nodes,
                            const HexNode* hn, const double* x, const double* y, const double* z,
                            const double* m, const double* e, const int64_t* pp_off,
                            const int32_t* pp, const int64_t* pc_off, const int32_t* pc,
                            const double off[3], double* ax, double* ay, double* az, double* pot,
                            hipStream_t s) {
  if (nb <= 0) return 0;
  interact_hexa_kernel<<<grid_for(nb), kThreads, 0, s>>>(bk, nb, nodes, hn, x, y, z, m, e, pp_off,
                                                         pp, pc_off, pc, off[0], off[1], off[2],
                                                         ax, ay, az, pot);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Fused HEXADECAPOLE evaluation: one work-group per 4 buckets.
int64_t pt_hexa_gravity_blocks(int64_t nb) { return (nb + kWaves - 1) / kWaves; }

int pt_launch_hexa_gravity(int precision, bool big_stack, const Bucket* bk, int64_t nb,
                           const TreeNode* nodes, const HexOpen* ho, const HexNode* hn,
                           const double* x, const double* y, const double* z, const double* m,
                           const double* e, double gf, double gm, const double off[3],
                           double* ax, double* ay, double* az, double* pot,
                           unsigned long long* partials, int32_t* err, bool accumulate,
                           hipStream_t s) {
  if (nb <= 0) return 0;
  const unsigned grid = (unsigned)pt_hexa_gravity_blocks(nb);
#define PT_HEXA_LAUNCH(STACK, F32)                                                               \
  hexa_gravity_kernel<STACK, F32><<<grid, kThreads, 0, s>>>(                                   \
      bk, nb, nodes, ho, hn, x, y, z, m, e, gf, gm, off[0], off[1], off[2], ax, ay, az, pot,    \
      partials, err, accumulate ? 1 : 0)
  if (precision == 0) {
    if (big_stack) PT_HEXA_LAUNCH(kBigStack, false);
    else PT_HEXA_LAUNCH(kSmallStack, false);
  } else {
    if (big_stack) PT_HEXA_LAUNCH(kBigStack, true);
    else PT_HEXA_LAUNCH(kSmallStack, true);
  }
#undef PT_HEXA_LAUNCH
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// big_stack selects the kBigStack-entry walk stack (used after a walk with
// the small stack reported an overflow; 1536 entries keep the FP64
// work-group under 64 KiB of LDS).
int pt_launch_gravity(int precision, bool big_stack, const Bucket* bk, int64_t nb,
                      const TreeNode* nodes, const GravNode* gn, const GravNode32* gn32,
                      const double* x, const double* y, const double* z, const double* m,
                      const double* e, double gf, const double off[3], double* ax, double* ay, double* az,
                      double* pot, unsigned long long* partials, int32_t* err,
                      unsigned long long* queue, bool accumulate, hipStream_t s) {
  if (nb <= 0) return 0;
  if (pt_launch_fill_u64(queue, 8, 0ull, s) != 0) return -1;
  const unsigned grid = (unsigned)pt_gravity_blocks(nb, precision);
  const bool has_off = off[0] != 0.0 || off[1] != 0.0 || off[2] != 0.0;
#define PT_GRAVITY_LAUNCH(F32, STACK, OFF)                                                     \
  gravity_kernel<F32, STACK, OFF><<<grid, kThreads, 0, s>>>(                                   \
      bk, nb, nodes, gn, gn32, x, y, z, m, e, gf, off[0], off[1], off[2], ax, ay, az, pot,      \
      partials, err, queue, accumulate ? 1 : 0)
#define PT_GRAVITY_LAUNCH_OFF(F32, STACK)   \
  if (has_off) PT_GRAVITY_LAUNCH(F32, STACK, true); \
  else PT_GRAVITY_LAUNCH(F32, STACK, false)
  if (precision == 0) {
    if (big_stack) PT_GRAVITY_LAUNCH_OFF(false, kBigStack);
    else PT_GRAVITY_LAUNCH_OFF(false, kSmallStack);
  } else {
    if (big_stack) PT_GRAVITY_LAUNCH_OFF(true, kBigStack);
    else PT_GRAVITY_LAUNCH_OFF(true, kSmallStack);
  }
#undef PT_GRAVITY_LAUNCH_OFF
#undef PT_GRAVITY_LAUNCH
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int64_t pt_fgroup_windows(int64_t nb) { return (nb + kFgWindow - 1) / kFgWindow; }

// Two consecutive groups of a window hold more than cap particles together,
// so a window of T particles has at most 2 T / (cap + 1) + 1 groups.
int64_t pt_fgroup_max_groups(int64_t nb, int64_t n, int cap) {
  return pt_fgroup_windows(nb) + 2 * n / (cap + 1) + 1;
}

int64_t pt_grouped_gravity_blocks(int64_t nb, int64_t n, int cap) {
  const int64_t per_xcd = (pt_fgroup_max_groups(nb, n, cap) + 7) / 8 + 1;
  return 8 * ((per_xcd + kGWaves - 1) / kGWaves);
}

int pt_launch_fgroups(const Bucket* bk, int64_t nb, int cap, int32_t* tfirst, int32_t* tnb,
                      uint32_t* wcount, uint32_t* woff, void* scratch, uint32_t* ngroups,
                      int32_t* gfirst, int32_t* gnb, hipStream_t s) {
  if (nb <= 0) return 0;
  if (cap < 2 || cap > PT_WAVE) return -1;
  const int64_t nw = pt_fgroup_windows(nb);
  fgroup_window_kernel<<<(unsigned)((nw + 3) / 4), 256, 0, s>>>(bk, nb, cap, tfirst, tnb,
                                                                 wcount);
  if (hipGetLastError() != hipSuccess) return -1;
  if (pt_exclusive_scan_u32(wcount, woff, nw, scratch, ngroups, s) != 0) return -1;
  fgroup_compact_kernel<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(nb, tfirst, tnb, wcount,
                                                                      woff, gfirst, gnb);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_grouped_gravity(int precision, bool big_stack, const Bucket* bk,
                              const int32_t* gfirst, const int32_t* gnb, const uint32_t* ngroups,
                              int64_t nb, int64_t n, int cap, const TreeNode* nodes,
                              const GravNode* gn, const GravNode32* gn32, const HexOpen* ho,
                              const HexNode* hn, const double* x, const double* y,
                              const double* z, const double* m, const double* e, double gf,
                              double gm, const double off[3], double* ax, double* ay, double* az,
                              double* pot, unsigned long long* partials, int32_t* err,
                              bool accumulate, hipStream_t s) {
  if (nb <= 0) return 0;
  const unsigned grid = (unsigned)pt_grouped_gravity_blocks(nb, n, cap);
  const bool has_off = off[0] != 0.0 || off[1] != 0.0 || off[2] != 0.0;
  const bool hexa = ho != nullptr;
#define PT_GROUPED_LAUNCH(F32, STACK, OFF, HEXA)                                               \
  grouped_gravity_kernel<F32, STACK, OFF, HEXA><<<grid, kGWaves * PT_WAVE, 0, s>>>(            \
      bk, gfirst, gnb, ngroups, nodes, gn, gn32, ho, hn, x, y, z, m, e, gf, gm, off[0], off[1], \
      off[2], ax, ay, az, pot, partials, err, accumulate ? 1 : 0)
#define PT_GROUPED_LAUNCH_OFF(F32, STACK, HEXA)               \
  if (has_off) PT_GROUPED_LAUNCH(F32, STACK, true, HEXA);     \
  else PT_GROUPED_LAUNCH(F32, STACK, false, HEXA)
  if (hexa) {
    if (precision == 0) {
      if (big_stack) PT_GROUPED_LAUNCH_OFF(false, kBigStack, true);
      else PT_GROUPED_LAUNCH_OFF(false, kGroupStackHexa, true);
    } else {
      if (big_stack) PT_GROUPED_LAUNCH_OFF(true, kBigStack, true);
      else PT_GROUPED_LAUNCH_OFF(true, kGroupStackHexa, true);
    }
  } else if (precision == 0) {
    if (big_stack) PT_GROUPED_LAUNCH_OFF(false, kBigStack, false);
    else PT_GROUPED_LAUNCH_OFF(false, kGroupStack64, false);
  } else {
    if (big_stack) PT_GROUPED_LAUNCH_OFF(true, kBigStack, false);
    else PT_GROUPED_LAUNCH_OFF(true, kGroupStack32, false);
  }
#undef PT_GROUPED_LAUNCH_OFF
#undef PT_GROUPED_LAUNCH
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

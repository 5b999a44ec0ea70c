// Shared device definitions for the gfx950 gravity path (keys, sort, octree,
// moments, traversal, interaction kernels).  Internal header: the public C ABI
// is include/pt_gravity.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PT_WAVE 64
#define PT_KEY_BITS 64
#define PT_MAX_DEPTH 21       // 63 Morton bits / 3 bits per octree level
#define PT_MIN_PARTICLE_NODE 6  // GravityVisitor::nMinParticleNode

// Node types, same numbering as the oracle (EmptyLeaf, Leaf, Internal).
enum { PT_NODE_EMPTY = 0, PT_NODE_LEAF = 1, PT_NODE_INTERNAL = 2 };

// Per-node moment record fields (CentroidData flattened), SoA index order used
// by pt_get_node_data: radius soft mass cmx cmy cmz xx xy xz yy yz zz lox loy
// loz hix hiy hiz.
#define PT_CD_RADIUS 0
#define PT_CD_SOFT 1
#define PT_CD_MASS 2
#define PT_CD_CM 3
#define PT_CD_Q 6
#define PT_CD_LO 12
#define PT_CD_HI 15
#define PT_CD_NFIELDS 18

// Packed node record read by the tree walk: centre of mass, radius and
// topology in one 64-byte line (the 8 children of a node are consecutive
// records, so a wavefront reads them as 512 contiguous bytes).
struct __align__(64) TreeNode {
  double cx, cy, cz, radius;
  int32_t count;
  int32_t first_child;  // -1 for leaves
  int32_t type;
  int32_t begin;
  // leaves: upper bounds of |x_j - cm| (the farthest corner of the particle
  // box) and of the particles' softening, rounded up to FP32; 0 otherwise
  float leaf_reach;
  float leaf_soft;
  // internal nodes: the node's softening (CentroidData soft) rounded up to
  // FP32; 0 for leaves
  float node_soft;
  int32_t pad;
};
// Packed source record for particle-cell interactions (FP64 path).
// tr = 0.5 * (qxx + qyy + qzz) as GravityVisitor::node computes it.
struct __align__(32) GravNode {
  double cx, cy, cz, soft;
  double mass, qxx, qxy, qxz;
  double qyy, qyz, qzz, tr;
};
// Source record of the compensated-FP32 interactions in one 64-byte line: the
// centre of mass and softening in FP64 (the centre is shifted per target
// bucket, the softening decides openSoftening exactly) and FP32 copies of the
// mass, softening and quadrupole.
struct __align__(64) GravNode32 {
  double cx, cy, cz, soft64;
  float mass, soft, qxx, qxy, qxz, qyy, qyz, qzz;
};
// HEXADECAPOLE build (-DHEXADECAPOLE, examples/Makefile:5).  Source record
// of node(): centre of mass, mass, softening and the 21 reduced (trace-free)
// moments of orders 2-4 in FMOMR order (examples/moments.h): xx yy xy xz yz,
// xxx xyy xxy yyy xxz yyz xyz, xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz.
#define PT_HEXA_MOMS 21
#define PT_HEXA_RAW 25  // raw third (10) and fourth (15) moments about the cm
struct __align__(64) HexNode {
  double cx, cy, cz, mass, soft;
  double h[PT_HEXA_MOMS];
  double pad[6];
};
// Source data read by the HEXADECAPOLE open(): the node's box and softening.
struct __align__(64) HexOpen {
  double lox, loy, loz, hix, hiy, hiz, soft, pad;
};

// One particle's inputs in one 64-byte line (original order), so the sort's
// permutation gathers one line per particle instead of five scattered words.
struct __align__(64) ParticleRec {
  double x, y, z, m, e, pad[3];
};

// Bucket (target leaf) record: particle range, box, centre of mass, softening.
struct __align__(32) Bucket {
  double lox, loy, loz, hix;
  double hiy, hiz, cx, cy;
  double cz, soft, pad0, pad1;
  int32_t node, begin, count, pad2;
};

// Interaction counters in the order {n_node_ints, n_part_ints, n_opens,
// n_closes} (src/ThreadStateHolder.h:10-13).
struct PtCounters {
  unsigned long long v[4];
};

__device__ __forceinline__ int pt_lane() { return __lane_id(); }

// Inclusive prefix sum over the 64 lanes: shifts within each row of 16 lanes
// (row_shr 1, 2, 4, 8), then each row's total broadcast into the rows above
// (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3).  DPP
// moves, no LDS round trips; every lane must be active.
__device__ __forceinline__ int pt_wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

// Maximum over the 64 lanes (all active), as a wave-uniform value: the same
// row shifts and broadcasts as pt_wave_incl_scan with max, lanes without a
// source reading -infinity.
__device__ __forceinline__ float pt_wave_max_f32(float v) {
  const int ninf = __float_as_int(-INFINITY);
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x111, 0xf, 0xf, false)));
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x112, 0xf, 0xf, false)));
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x114, 0xf, 0xf, false)));
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x118, 0xf, 0xf, false)));
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x142, 0xa, 0xf, false)));
  v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x143, 0xc, 0xf, false)));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), PT_WAVE - 1));
}

// Sum over the 64 lanes (all active), as a wave-uniform value.
__device__ __forceinline__ int pt_wave_sum_int(int v) {
  return __builtin_amdgcn_readlane(pt_wave_incl_scan(v), PT_WAVE - 1);
}

// A double moved by one DPP control (both 32-bit halves); lanes without a
// source read 0.0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double pt_dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(
      (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}

// Sum over the 64 lanes (all active) with the row shifts and broadcasts of
// pt_wave_incl_scan, as a wave-uniform value (lane 63's inclusive sum).
__device__ __forceinline__ double pt_wave_sum(double v) {
  v += pt_dpp_f64<0x111, 0xf>(v);
  v += pt_dpp_f64<0x112, 0xf>(v);
  v += pt_dpp_f64<0x114, 0xf>(v);
  v += pt_dpp_f64<0x118, 0xf>(v);
  v += pt_dpp_f64<0x142, 0xa>(v);
  v += pt_dpp_f64<0x143, 0xc>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, PT_WAVE - 1);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), PT_WAVE - 1);
  return __longlong_as_double(
      (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}

// 1/sqrt(x), x > 0, in FP64: the hardware estimate refined by two Newton
// steps, within a few ulp of the correctly rounded 1.0 / sqrt(x) (which
// costs a square-root and a divide expansion).
__device__ __forceinline__ double pt_rsqrt64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double t = x * y;
  double e = fma(-t, y, 1.0);
  y = fma(0.5 * y, e, y);
  t = x * y;
  e = fma(-t, y, 1.0);
  return fma(0.5 * y, e, y);
}

__device__ __forceinline__ unsigned long long pt_ballot(bool p) {
  return __ballot(p);
}

// Number of set lanes below the calling lane in mask.
__device__ __forceinline__ int pt_rank_in(unsigned long long mask) {
  return __popcll(mask & ((1ull << __lane_id()) - 1ull));
}

// Host-side error check helper used inside the C ABI implementation.
#define PT_HIP_OK(expr) ((expr) == hipSuccess)

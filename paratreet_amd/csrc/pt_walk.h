// Wavefront tree walk shared by the list-producing traversal and the fused
// walk-and-interact kernel, so both evaluate exactly the same open() decisions.
//
// Reference: TransposedDownTraverser::recurse (src/Traverser.h:154-217) seen
// from one target bucket, with GravityVisitor::open (examples/GravityVisitor.h:
// 162-164,180).  A Leaf source yields leaf() (self included: CallSelfLeaf), an
// EmptyLeaf nothing, an Internal source an open() test: opened -> its 8
// children are visited, closed -> node().  The per-bucket set of calls does not
// depend on the visit order, so the wavefront visits 8 opened nodes' children
// per step (64 lanes): the LDS stack holds the first-child ids of opened
// nodes, whose 8 children are consecutive records.
//
// Included only by translation units compiled with -ffp-contract=off so the
// opening test rounds exactly like the reference's separately rounded
// expression.
#pragma once

#include "pt_common.h"

// Walk stack entries per wavefront in the list-producing traversal.
#define PT_WALK_STACK 1024

// The squared distance from the box to the point as Space::intersect
// (box, point, rsq) sums it: per-axis clamp distances, x then y then z.
__device__ __forceinline__ double pt_box_point_dsq(const Bucket& b, double px, double py,
                                                   double pz) {
  double dsq = 0.0;
  if (px < b.lox) {
    const double t = b.lox - px;
    dsq += t * t;
  } else if (px > b.hix) {
    const double t = px - b.hix;
    dsq += t * t;
  }
  if (py < b.loy) {
    const double t = b.loy - py;
    dsq += t * t;
  } else if (py > b.hiy) {
    const double t = py - b.hiy;
    dsq += t * t;
  }
  if (pz < b.loz) {
    const double t = b.loz - pz;
    dsq += t * t;
  } else if (pz > b.hiz) {
    const double t = pz - b.hiz;
    dsq += t * t;
  }
  return dsq;
}

// The squared distance from the point to the box's farthest corner, per-axis
// terms summed in the order of pt_box_point_dsq.  For any box B inside b the
// rounded pt_box_point_dsq(B, p) is at most this value: each of its per-axis
// terms is a rounded difference of p and a face of B that is never larger
// than the one summed here (rounding is monotone), and so are their squares
// and sums.  Likewise pt_box_point_dsq(b, p) <= pt_box_point_dsq(B, p).
__device__ __forceinline__ double pt_box_point_far_dsq(const Bucket& b, double px, double py,
                                                       double pz) {
  const double tx = fmax(px - b.lox, b.hix - px);
  const double ty = fmax(py - b.loy, b.hiy - py);
  const double tz = fmax(pz - b.loz, b.hiz - pz);
  double dsq = tx * tx;
  dsq += ty * ty;
  dsq += tz * tz;
  return dsq;
}

// Space::intersect(box, point, rsq): dsq <= rsq.
__device__ __forceinline__ bool pt_intersect_box_point(const Bucket& b, double px, double py,
                                                       double pz, double rsq) {
  return pt_box_point_dsq(b, px, py, pz) <= rsq;
}

// GravityVisitor::open (examples/GravityVisitor.h:162-164,180) from the
// squared box distance dsq of the node's centre (+ offset).
__device__ __forceinline__ bool pt_gv_open_dsq(const TreeNode& s, double gf, double dsq) {
  if (s.count <= PT_MIN_PARTICLE_NODE) return true;
  const double dataRsq = s.radius * s.radius * gf;
  return dsq <= dataRsq;
}

// GravityVisitor::open (examples/GravityVisitor.h:162-164,180).
__device__ __forceinline__ bool pt_gv_open(const TreeNode& s, const Bucket& b, double gf,
                                           double ox, double oy, double oz) {
  if (s.count <= PT_MIN_PARTICLE_NODE) return true;
  const double dataRsq = s.radius * s.radius * gf;
  return pt_intersect_box_point(b, s.cx + ox, s.cy + oy, s.cz + oz, dataRsq);
}

// GravityVisitor::openSoftening (examples/GravityVisitor.h:122-128) with the
// bucket as target: the source sphere (cm + offset, 2 soft) against the
// bucket's sphere (cm, 2 soft), then against the bucket's box.
__device__ __forceinline__ bool pt_open_softening(double sx, double sy, double sz, double ssoft,
                                                  const Bucket& b) {
  const double sr = 2.0 * ssoft;
  {
    const double dx = sx - b.cx, dy = sy - b.cy, dz = sz - b.cz;
    const double dsq = dx * dx + dy * dy + dz * dz;
    const double rs = sr + 2.0 * b.soft;
    if (dsq <= rs * rs) return true;
  }
  return pt_intersect_box_point(b, sx, sy, sz, sr * sr);
}

// GravityVisitor::open, HEXADECAPOLE branch (examples/GravityVisitor.h:
// 162-178): closed when the opening sphere misses the bucket box, unless the
// softening spheres meet and the (larger) monopole sphere reaches the box.
// The softening test is called as written there, openSoftening(target.data,
// source.data): the bucket's centre (+ offset) and softening take the source
// role, the node's sphere and box the target role.  gm is
// monopole_gravity_factor.
__device__ __forceinline__ bool pt_gv_open_hexa(const TreeNode& s, const HexOpen& h,
                                                const Bucket& b, double gf, double gm, double ox,
                                                double oy, double oz) {
  if (s.count <= PT_MIN_PARTICLE_NODE) return true;
  const double dataRsq = s.radius * s.radius * gf;
  const double px = s.cx + ox, py = s.cy + oy, pz = s.cz + oz;
  if (pt_intersect_box_point(b, px, py, pz, dataRsq)) return true;
  const double bx = b.cx + ox, by = b.cy + oy, bz = b.cz + oz;
  const double sr = 2.0 * b.soft;
  bool soft_meet;
  {
    const double dx = bx - s.cx, dy = by - s.cy, dz = bz - s.cz;
    const double dsq = dx * dx + dy * dy + dz * dz;
    const double rs = sr + 2.0 * h.soft;
    soft_meet = dsq <= rs * rs;
  }
  if (!soft_meet) {
    Bucket nb;
    nb.lox = h.lox;
    nb.loy = h.loy;
    nb.loz = h.loz;
    nb.hix = h.hix;
    nb.hiy = h.hiy;
    nb.hiz = h.hiz;
    soft_meet = pt_intersect_box_point(nb, bx, by, bz, sr * sr);
  }
  if (!soft_meet) return false;
  const double monoRad = s.radius * gm;
  return pt_intersect_box_point(b, px, py, pz, monoRad * monoRad);
}

// The opening test of the build: quadrupole (ho == nullptr) or HEXADECAPOLE.
template <bool HEXA>
__device__ __forceinline__ bool pt_open_any(const TreeNode& s, int32_t id,
                                            const HexOpen* __restrict__ ho, const Bucket& b,
                                            double gf, double gm, double ox, double oy,
                                            double oz) {
  if constexpr (HEXA) {
    if (s.count <= PT_MIN_PARTICLE_NODE) return true;
    return pt_gv_open_hexa(s, ho[id], b, gf, gm, ox, oy, oz);
  } else {
    return pt_gv_open(s, b, gf, ox, oy, oz);
  }
}

// Result of one walk step for the calling lane.
struct WalkLane {
  int32_t node;      // visited node id (valid if active)
  int32_t count;     // its particle count
  int32_t begin;     // its first particle
  bool is_leaf;      // -> leaf() (P-P)
  bool is_closed;    // -> node() (P-C)
  double cx, cy, cz;        // its centre of mass
  float leaf_reach, leaf_soft;  // leaf bounds (TreeNode)
  float node_soft;              // internal nodes' softening bound (TreeNode)
  double dsq;                   // squared distance of cx + offset from the bucket box
};

// Wave-uniform per-step counts for the interaction counters.
struct WalkCounts {
  int n_open;     // open() returned true
  int n_close;    // open() returned false (node() calls)
  int leaf_parts; // sum of source-leaf particle counts (leaf() calls)
};

// Visits the root (node 0): pushes its first child when it is opened.  sp and
// the counts are wave-uniform; lane 0 owns the stack write.
template <bool HEXA = false>
__device__ __forceinline__ WalkLane pt_walk_root(const TreeNode* __restrict__ nodes,
                                                 const Bucket& b, double gf, double ox,
                                                 double oy, double oz,
                                                 int32_t* __restrict__ stack, int& sp,
                                                 WalkCounts& wc,
                                                 const HexOpen* __restrict__ ho = nullptr,
                                                 double gm = 0.0) {
  WalkLane r{0, 0, 0, false, false, 0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.0};
  wc = WalkCounts{0, 0, 0};
  const TreeNode root = nodes[0];
  r.count = root.count;
  r.begin = root.begin;
  r.cx = root.cx;
  r.cy = root.cy;
  r.cz = root.cz;
  r.leaf_reach = root.leaf_reach;
  r.leaf_soft = root.leaf_soft;
  r.node_soft = root.node_soft;
  r.dsq = pt_box_point_dsq(b, root.cx + ox, root.cy + oy, root.cz + oz);
  if (root.type == PT_NODE_LEAF) {
    r.is_leaf = true;
    wc.leaf_parts = root.count;
  } else if (root.type == PT_NODE_INTERNAL) {
    if (pt_open_any<HEXA>(root, 0, ho, b, gf, gm, ox, oy, oz)) {
      if (__lane_id() == 0) stack[0] = root.first_child;
      sp = 1;
      wc.n_open = 1;
    } else {
      r.is_closed = true;
      wc.n_close = 1;
    }
  }
  return r;
}

// One step: pops up to 8 opened nodes and tests their 64 children.  Returns
// this lane's child and its classification; pushes opened children.  A push
// beyond `limit` (<= STACK) entries sets overflow and ends the walk (the
// caller reports it; nothing is silently dropped).  With SUM_LEAVES false wc.leaf_parts is
// left 0 and the caller sums the leaf lanes' counts itself.
template <int STACK, bool HEXA = false, bool SUM_LEAVES = true>
__device__ __forceinline__ WalkLane pt_walk_step(const TreeNode* __restrict__ nodes,
                                                 const Bucket& b, double gf, double ox,
                                                 double oy, double oz,
                                                 int32_t* __restrict__ stack, int& sp,
                                                 bool& overflow, WalkCounts& wc,
                                                 unsigned long long& leafm,
                                                 unsigned long long& closem,
                                                 const HexOpen* __restrict__ ho = nullptr,
                                                 double gm = 0.0, int limit = STACK) {
  const int lane = __lane_id();
  const int k = sp < 8 ? sp : 8;
  sp -= k;
  const int slot = lane >> 3;
  const bool active = slot < k;
  // lanes beyond the k popped entries read the first one's children again
  // (k >= 1 while the walk runs) and are masked out of every outcome: the
  // loads need no branch and the lane state no defaults
  WalkLane r;
  r.node = stack[sp + (active ? slot : 0)] + (lane & 7);
  const TreeNode nd = nodes[r.node];
  r.count = nd.count;
  r.begin = nd.begin;
  r.cx = nd.cx;
  r.cy = nd.cy;
  r.cz = nd.cz;
  r.leaf_reach = nd.leaf_reach;
  r.leaf_soft = nd.leaf_soft;
  r.node_soft = nd.node_soft;
  const int32_t fc = nd.first_child;
  r.is_leaf = active && nd.type == PT_NODE_LEAF;
  const bool internal = active && nd.type == PT_NODE_INTERNAL;
  // every lane's box distance: the opening test's for internal nodes, the
  // near / far classification's for leaves and closed nodes
  r.dsq = pt_box_point_dsq(b, nd.cx + ox, nd.cy + oy, nd.cz + oz);
  bool op = false;
  if constexpr (HEXA) {
    if (internal) op = pt_open_any<HEXA>(nd, r.node, ho, b, gf, gm, ox, oy, oz);
  } else {
    op = internal && pt_gv_open_dsq(nd, gf, r.dsq);
  }
  r.is_closed = internal && !op;
  leafm = __ballot(r.is_leaf);
  closem = __ballot(r.is_closed);
  const unsigned long long openm = __ballot(op);
  const int n_open = __popcll(openm);
  if (sp + n_open > limit) {
    overflow = true;
    sp = 0;
    leafm = closem = 0;
    r.is_leaf = r.is_closed = false;
    wc = WalkCounts{0, 0, 0};
    return r;
  }
  __builtin_amdgcn_wave_barrier();
  if (op) stack[sp + pt_rank_in(openm)] = fc;
  __builtin_amdgcn_wave_barrier();
  sp += n_open;
  wc.n_open = n_open;
  wc.n_close = __popcll(closem);
  if constexpr (SUM_LEAVES) wc.leaf_parts = pt_wave_sum_int(r.is_leaf ? r.count : 0);
  else wc.leaf_parts = 0;
  return r;
}

// XCD-aware bucket order: work-groups are dispatched round-robin over the 8
// XCDs, so group i is given the (i % 8)-th contiguous eighth of the buckets;
// neighbouring buckets (which share most of their source nodes) then run on
// the same XCD and share its L2.
__device__ __forceinline__ int64_t pt_xcd_block(int64_t i, int64_t g) {
  const int64_t g8 = (g / 8) * 8;
  if (i >= g8) return i;
  return (i % 8) * (g8 / 8) + i / 8;
}

// Node moments: leaf CentroidData and the upward pass, in FP64.
//
// Reference: CentroidData(const Particle*, n, depth) and operator+=
// (examples/CentroidData.h:41-73), MultipoleMoments::operator+= for particles
// and for expansions, quadrupole branch (examples/MultipoleMoments.h:204-353),
// calculateRadiusFarthestCorner / calculateRadiusBox (:458-482), and the merge
// order of Subtree::populateTree (src/Subtree.h:351-384).
//
// populateTree pushes every non-empty leaf (DFS order), then every empty leaf
// (DFS order) into a FIFO and enqueues a parent when its 8th child has been
// merged into it.  The merges are order-dependent at rounding level, so each
// parent here merges its children in the order that FIFO pops them:
// non-empty leaves by child index, then empty leaves by child index, then
// internal children ordered by the pair (h, b) computed below.  A frontier
// node (no internal child) has h = 0 and b = its last-popped leaf child's pop
// rank; any other node takes (h + 1, b) of its last-popped internal child.  Pop
// order among internal nodes is lexicographic in (h, b): every leaf is popped
// before every internal node, and an internal node is enqueued right after its
// last child is popped.  Leaf pop ranks are encoded monotonically as the
// left-aligned key (DFS order of disjoint nodes) with bit 63 marking empties.
//
// The file is compiled with -ffp-contract=off: every multiply and add rounds
// separately, as in the reference source, so the results are bit-exact with a
// CPU evaluation of the same operations.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

struct CData {
  double radius, soft, mass;
  double cm[3];
  double q[6];
  double lo[3], hi[3];
  int64_t count;
};

__device__ __forceinline__ void cd_default(CData& d) {
  d.radius = 0.0;
  d.soft = 0.0;
  d.mass = 0.0;
  for (int k = 0; k < 3; k++) {
    d.cm[k] = 0.0;
    d.lo[k] = INFINITY;
    d.hi[k] = -INFINITY;
  }
  for (int k = 0; k < 6; k++) d.q[k] = 0.0;
  d.count = 0;
}

__device__ __forceinline__ void cd_load(CData& d, const double* __restrict__ nd, int64_t nn,
                                        int64_t i, int32_t count) {
  d.radius = nd[PT_CD_RADIUS * nn + i];
  d.soft = nd[PT_CD_SOFT * nn + i];
  d.mass = nd[PT_CD_MASS * nn + i];
  for (int k = 0; k < 3; k++) {
    d.cm[k] = nd[(PT_CD_CM + k) * nn + i];
    d.lo[k] = nd[(PT_CD_LO + k) * nn + i];
    d.hi[k] = nd[(PT_CD_HI + k) * nn + i];
  }
  for (int k = 0; k < 6; k++) d.q[k] = nd[(PT_CD_Q + k) * nn + i];
  d.count = count;
}

__device__ __forceinline__ void cd_store(const CData& d, double* __restrict__ nd, int64_t nn,
                                         int64_t i) {
  nd[PT_CD_RADIUS * nn + i] = d.radius;
  nd[PT_CD_SOFT * nn + i] = d.soft;
  nd[PT_CD_MASS * nn + i] = d.mass;
  for (int k = 0; k < 3; k++) {
    nd[(PT_CD_CM + k) * nn + i] = d.cm[k];
    nd[(PT_CD_LO + k) * nn + i] = d.lo[k];
    nd[(PT_CD_HI + k) * nn + i] = d.hi[k];
  }
  for (int k = 0; k < 6; k++) nd[(PT_CD_Q + k) * nn + i] = d.q[k];
}

// MultipoleMoments::operator+=(Particle) (examples/MultipoleMoments.h:275-353).
__device__ __forceinline__ void mom_add_particle(CData& d, double px, double py, double pz,
                                                 double pm, double pe) {
  const double m1 = d.mass;
  d.mass += pm;
  if (d.mass == 0.0) {
    d.soft = 0.5 * (d.soft + pe);
    d.cm[0] = 0.5 * (d.cm[0] + px);
    d.cm[1] = 0.5 * (d.cm[1] + py);
    d.cm[2] = 0.5 * (d.cm[2] + pz);
    return;
  }
  d.soft = (m1 * d.soft + pm * pe) / d.mass;
  const double c1[3] = {d.cm[0], d.cm[1], d.cm[2]};
  const double p[3] = {px, py, pz};
  for (int k = 0; k < 3; k++) d.cm[k] = (m1 * d.cm[k] + pm * p[k]) / d.mass;
  double dr[3] = {c1[0] - d.cm[0], c1[1] - d.cm[1], c1[2] - d.cm[2]};
  d.q[0] += m1 * dr[0] * dr[0];
  d.q[3] += m1 * dr[1] * dr[1];
  d.q[5] += m1 * dr[2] * dr[2];
  d.q[1] += m1 * dr[0] * dr[1];
  d.q[2] += m1 * dr[0] * dr[2];
  d.q[4] += m1 * dr[1] * dr[2];
  for (int k = 0; k < 3; k++) dr[k] = p[k] - d.cm[k];
  d.q[0] += pm * dr[0] * dr[0];
  d.q[3] += pm * dr[1] * dr[1];
  d.q[5] += pm * dr[2] * dr[2];
  d.q[1] += pm * dr[0] * dr[1];
  d.q[2] += pm * dr[0] * dr[2];
  d.q[4] += pm * dr[1] * dr[2];
}

// MultipoleMoments::operator+=(MultipoleMoments) (examples/MultipoleMoments.h:204-271).
__device__ __forceinline__ void mom_merge(CData& d, const CData& m) {
  const double m1 = d.mass;
  d.mass += m.mass;
  if (d.mass == 0.0) {
    d.soft = 0.5 * (d.soft + m.soft);
    for (int k = 0; k < 3; k++) d.cm[k] = 0.5 * (d.cm[k] + m.cm[k]);
    return;
  }
  if (m1 == 0.0) {
    d.radius = m.radius;
    d.soft = m.soft;
    d.mass = m.mass;
    for (int k = 0; k < 3; k++) d.cm[k] = m.cm[k];
    for (int k = 0; k < 6; k++) d.q[k] = m.q[k];
    return;
  }
  if (m.mass == 0.0) return;
  d.soft = (m1 * d.soft + m.mass * m.soft) / d.mass;
  const double c1[3] = {d.cm[0], d.cm[1], d.cm[2]};
  for (int k = 0; k < 3; k++) d.cm[k] = (m1 * d.cm[k] + m.mass * m.cm[k]) / d.mass;
  double dr[3] = {c1[0] - d.cm[0], c1[1] - d.cm[1], c1[2] - d.cm[2]};
  d.q[0] += m1 * dr[0] * dr[0];
  d.q[3] += m1 * dr[1] * dr[1];
  d.q[5] += m1 * dr[2] * dr[2];
  d.q[1] += m1 * dr[0] * dr[1];
  d.q[2] += m1 * dr[0] * dr[2];
  d.q[4] += m1 * dr[1] * dr[2];
  for (int k = 0; k < 3; k++) dr[k] = m.cm[k] - d.cm[k];
  d.q[0] += m.q[0] + m.mass * dr[0] * dr[0];
  d.q[3] += m.q[3] + m.mass * dr[1] * dr[1];
  d.q[5] += m.q[5] + m.mass * dr[2] * dr[2];
  d.q[1] += m.q[1] + m.mass * dr[0] * dr[1];
  d.q[2] += m.q[2] + m.mass * dr[0] * dr[2];
  d.q[4] += m.q[4] + m.mass * dr[1] * dr[2];
}

__device__ __forceinline__ double len3(double x, double y, double z) {
  return sqrt(x * x + y * y + z * z);
}

// CentroidData::operator+= (examples/CentroidData.h:61-73) with
// calculateRadiusFarthestCorner (examples/MultipoleMoments.h:458-469).
__device__ __forceinline__ void cd_merge(CData& d, const CData& c) {
  for (int k = 0; k < 3; k++) {
    if (c.lo[k] < d.lo[k]) d.lo[k] = c.lo[k];
    if (c.hi[k] > d.hi[k]) d.hi[k] = c.hi[k];
  }
  mom_merge(d, c);
  d.count += c.count;
  if (d.count > 1) {
    double dl[3];
    for (int k = 0; k < 3; k++) {
      const double d1 = d.cm[k] - d.lo[k];
      const double d2 = d.hi[k] - d.cm[k];
      dl[k] = (d1 > d2 ? d1 : d2);
    }
    d.radius = len3(dl[0], dl[1], dl[2]);
  } else {
    d.radius = 1.0;
  }
}

__global__ void __launch_bounds__(256)
init_kernel(double* __restrict__ nd, int64_t nn) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= nn) return;
  CData d;
  cd_default(d);
  cd_store(d, nd, nn, i);
}

// One thread per bucket: the sequential particle recurrence of the
// CentroidData leaf constructor, in sorted particle order.
__global__ void __launch_bounds__(256)
leaf_kernel(const PtTree t, const int32_t* __restrict__ bucket_node, int64_t nb,
            const double* __restrict__ x, const double* __restrict__ y,
            const double* __restrict__ z, const double* __restrict__ m,
            const double* __restrict__ e, double* __restrict__ nd, int64_t nn) {
  const int64_t bi = blockIdx.x * 256ll + threadIdx.x;
  if (bi >= nb) return;
  const int32_t id = bucket_node[bi];
  const int64_t b = t.begin[id];
  const int64_t n = t.count[id];
  CData d;
  cd_default(d);
  for (int64_t i = b; i < b + n; i++) {
    const double px = x[i], py = y[i], pz = z[i];
    mom_add_particle(d, px, py, pz, m[i], e[i]);
    const double p[3] = {px, py, pz};
    for (int k = 0; k < 3; k++) {
      if (p[k] < d.lo[k]) d.lo[k] = p[k];
      if (p[k] > d.hi[k]) d.hi[k] = p[k];
    }
  }
  d.count = n;
  if (d.count > 1) {
    d.radius = 0.5 * len3(d.hi[0] - d.lo[0], d.hi[1] - d.lo[1], d.hi[2] - d.lo[2]);
  } else {
    d.radius = 1.0;
  }
  cd_store(d, nd, nn, id);
}

__device__ __forceinline__ uint64_t left_align(uint64_t k) {
  const int depth = (63 - __clzll(k)) / 3;
  return k << (PT_KEY_BITS - (3 * depth + 1));
}

// One thread per internal node of one level (deepest level first).  The
// children's pop order is built in registers (fully unrolled loops over the
// 8 children, the order packed as 8 nibbles) so nothing spills to scratch.
__global__ void __launch_bounds__(256)
upward_kernel(const PtTree t, const int32_t* __restrict__ level_nodes, int32_t nl,
              double* __restrict__ nd, int64_t nn, int32_t* __restrict__ ord_h,
              uint64_t* __restrict__ ord_b) {
  const int64_t li = blockIdx.x * 256ll + threadIdx.x;
  if (li >= nl) return;
  const int32_t p = level_nodes[li];
  const int32_t fc = t.first_child[p];
  int ctype[8];
  int32_t h[8];
  uint64_t b[8];
  int n_leaf = 0, n_empty = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    ctype[c] = t.type[fc + c];
    h[c] = -1;
    b[c] = 0;
    if (ctype[c] == PT_NODE_INTERNAL) {
      h[c] = ord_h[fc + c];
      b[c] = ord_b[fc + c];
    }
    n_leaf += ctype[c] == PT_NODE_LEAF;
    n_empty += ctype[c] == PT_NODE_EMPTY;
  }
  // position of every child in the pop order: non-empty leaves, empty leaves
  // (each in child order), then internal children ascending in (h, b)
  uint32_t order = 0;
  int32_t hmax = -1;
  uint64_t bmax = 0;
  int k_leaf = 0, k_empty = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    int pos;
    if (ctype[c] == PT_NODE_LEAF) {
      pos = k_leaf++;
    } else if (ctype[c] == PT_NODE_EMPTY) {
      pos = n_leaf + k_empty++;
    } else {
      int rank = 0;
#pragma unroll
      for (int o = 0; o < 8; o++) {
        const bool before = ctype[o] == PT_NODE_INTERNAL &&
                            (h[o] < h[c] || (h[o] == h[c] && b[o] < b[c]));
        rank += before ? 1 : 0;
      }
      pos = n_leaf + n_empty + rank;
      if (h[c] > hmax || (h[c] == hmax && b[c] > bmax)) {
        hmax = h[c];
        bmax = b[c];
      }
    }
    order |= (uint32_t)c << (4 * pos);
  }
  if (hmax < 0) {
    // frontier node: h = 0, b = rank of the last-popped leaf child
    uint64_t bl = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      uint64_t r = left_align(t.key[fc + c]) >> 1;
      if (ctype[c] == PT_NODE_EMPTY) r |= (1ull << 63);
      bl = r > bl ? r : bl;
    }
    ord_h[p] = 0;
    ord_b[p] = bl;
  } else {
    ord_h[p] = hmax + 1;
    ord_b[p] = bmax;
  }
  CData d;
  cd_load(d, nd, nn, p, 0);
  d.count = 0;
  for (int j = 0; j < 8; j++) {
    const int c = (order >> (4 * j)) & 15;
    CData cd;
    cd_load(cd, nd, nn, fc + c, t.count[fc + c]);
    cd_merge(d, cd);
  }
  cd_store(d, nd, nn, p);
}

// Packs the walk and interaction records from the SoA node table.
__global__ void __launch_bounds__(256)
pack_nodes_kernel(const PtTree t, const double* __restrict__ nd, int64_t nn,
                  TreeNode* __restrict__ tnode, GravNode* __restrict__ gn,
                  GravNode32* __restrict__ gn32) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= nn) return;
  TreeNode a;
  a.cx = nd[(PT_CD_CM + 0) * nn + i];
  a.cy = nd[(PT_CD_CM + 1) * nn + i];
  a.cz = nd[(PT_CD_CM + 2) * nn + i];
  a.radius = nd[PT_CD_RADIUS * nn + i];
  a.count = t.count[i];
  a.first_child = t.first_child[i];
  a.type = t.type[i];
  a.begin = t.begin[i];
  a.pad[0] = a.pad[1] = a.pad[2] = a.pad[3] = 0;
  tnode[i] = a;
  GravNode g;
  g.cx = a.cx;
  g.cy = a.cy;
  g.cz = a.cz;
  g.mass = nd[PT_CD_MASS * nn + i];
  g.soft = nd[PT_CD_SOFT * nn + i];
  g.qxx = nd[(PT_CD_Q + 0) * nn + i];
  g.qxy = nd[(PT_CD_Q + 1) * nn + i];
  g.qxz = nd[(PT_CD_Q + 2) * nn + i];
  g.qyy = nd[(PT_CD_Q + 3) * nn + i];
  g.qyz = nd[(PT_CD_Q + 4) * nn + i];
  g.qzz = nd[(PT_CD_Q + 5) * nn + i];
  g.tr = 0.5 * (g.qxx + g.qyy + g.qzz);
  gn[i] = g;
  GravNode32 h;
  h.mass = (float)g.mass;
  h.soft = (float)g.soft;
  h.qxx = (float)g.qxx;
  h.qxy = (float)g.qxy;
  h.qxz = (float)g.qxz;
  h.qyy = (float)g.qyy;
  h.qyz = (float)g.qyz;
  h.qzz = (float)g.qzz;
  gn32[i] = h;
}

__global__ void __launch_bounds__(256)
pack_buckets_kernel(const PtTree t, const double* __restrict__ nd, int64_t nn,
                    const int32_t* __restrict__ bucket_node, int64_t nb,
                    Bucket* __restrict__ bk) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= nb) return;
  const int32_t id = bucket_node[i];
  Bucket b;
  b.lox = nd[(PT_CD_LO + 0) * nn + id];
  b.loy = nd[(PT_CD_LO + 1) * nn + id];
  b.loz = nd[(PT_CD_LO + 2) * nn + id];
  b.hix = nd[(PT_CD_HI + 0) * nn + id];
  b.hiy = nd[(PT_CD_HI + 1) * nn + id];
  b.hiz = nd[(PT_CD_HI + 2) * nn + id];
  b.cx = nd[(PT_CD_CM + 0) * nn + id];
  b.cy = nd[(PT_CD_CM + 1) * nn + id];
  b.cz = nd[(PT_CD_CM + 2) * nn + id];
  b.soft = nd[PT_CD_SOFT * nn + id];
  b.pad0 = 0.0;
  b.pad1 = 0.0;
  b.node = id;
  b.begin = t.begin[id];
  b.count = t.count[id];
  b.pad2 = 0;
  bk[i] = b;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int pt_launch_moments_init(double* nd, int64_t nn, hipStream_t s) {
  if (nn <= 0) return 0;
  init_kernel<<<grid_for(nn), 256, 0, s>>>(nd, nn);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_leaf_moments(const PtTree& t, const int32_t* bucket_node, int64_t nb,
                           const double* x, const double* y, const double* z, const double* m,
                           const double* e, double* nd, int64_t nn, hipStream_t s) {
  if (nb <= 0) return 0;
  leaf_kernel<<<grid_for(nb), 256, 0, s>>>(t, bucket_node, nb, x, y, z, m, e, nd, nn);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_upward_level(const PtTree& t, const int32_t* level_nodes, int32_t nl, double* nd,
                           int64_t nn, int32_t* ord_h, uint64_t* ord_b, hipStream_t s) {
  if (nl <= 0) return 0;
  upward_kernel<<<grid_for(nl), 256, 0, s>>>(t, level_nodes, nl, nd, nn, ord_h, ord_b);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_pack(const PtTree& t, const double* nd, int64_t nn, const int32_t* bucket_node,
                   int64_t nb, TreeNode* tnode, GravNode* gn, GravNode32* gn32, Bucket* bk,
                   hipStream_t s) {
  if (nn > 0) pack_nodes_kernel<<<grid_for(nn), 256, 0, s>>>(t, nd, nn, tnode, gn, gn32);
  if (nb > 0) pack_buckets_kernel<<<grid_for(nb), 256, 0, s>>>(t, nd, nn, bucket_node, nb, bk);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

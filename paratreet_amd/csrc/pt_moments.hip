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

// CentroidData::operator+= (examples/CentroidData.h:61-73) without its last
// step, the radius (cd_radius): no merge reads the radius it sets (a merge
// into a massless node copies the child's and the next radius overwrites
// it), so after a sequence of merges one cd_radius gives the value of the
// last merge bit for bit.
__device__ __forceinline__ void cd_merge(CData& d, const CData& c) {
  for (int k = 0; k < 3; k++) {
    if (c.lo[k] < d.lo[k]) d.lo[k] = c.lo[k];
    if (c.hi[k] > d.hi[k]) d.hi[k] = c.hi[k];
  }
  mom_merge(d, c);
  d.count += c.count;
}

// The radius step of CentroidData::operator+=: calculateRadiusFarthestCorner
// (examples/MultipoleMoments.h:458-469) when count > 1, else 1.0.
__device__ __forceinline__ void cd_radius(CData& d) {
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

// The smallest float >= v (v >= 0, finite).
__device__ __forceinline__ float round_up_f32(double v) {
  float f = (float)v;
  if ((double)f < v) f = __int_as_float(__float_as_int(f) + 1);
  return f;
}

// Writes node i's walk and interaction records from its final CentroidData
// (the same values a pass over the SoA node table would read back).  Leaves
// also get their reach and softening bounds (leaf_soft: the largest particle
// softening) and, when non-empty, their bucket record (bucket index
// head_pos[begin]).
__device__ __forceinline__ void pack_node(const PtTree& t, int64_t i, const CData& d,
                                          double leaf_soft, const PtPackOut& o) {
  TreeNode a;
  a.cx = d.cm[0];
  a.cy = d.cm[1];
  a.cz = d.cm[2];
  a.radius = d.radius;
  a.count = t.count[i];
  a.first_child = t.first_child[i];
  a.type = t.type[i];
  a.begin = t.begin[i];
  a.leaf_reach = 0.f;
  a.leaf_soft = 0.f;
  a.node_soft = a.type == PT_NODE_INTERNAL ? round_up_f32(d.soft) : 0.f;
  if (a.type == PT_NODE_LEAF) {
    double r2 = 0.0;
    for (int k = 0; k < 3; k++) {
      const double d1 = d.cm[k] - d.lo[k];
      const double d2 = d.hi[k] - d.cm[k];
      const double e = d1 > d2 ? d1 : d2;
      r2 += e * e;
    }
    a.leaf_reach = round_up_f32(sqrt(r2));
    a.leaf_soft = round_up_f32(leaf_soft);
  }
  a.pad = 0;
  o.tnode[i] = a;
  GravNode g;
  g.cx = d.cm[0];
  g.cy = d.cm[1];
  g.cz = d.cm[2];
  g.mass = d.mass;
  g.soft = d.soft;
  g.qxx = d.q[0];
  g.qxy = d.q[1];
  g.qxz = d.q[2];
  g.qyy = d.q[3];
  g.qyz = d.q[4];
  g.qzz = d.q[5];
  g.tr = 0.5 * (g.qxx + g.qyy + g.qzz);
  o.gn[i] = g;
  GravNode32 h;
  h.cx = g.cx;
  h.cy = g.cy;
  h.cz = g.cz;
  h.soft64 = g.soft;
  h.mass = (float)g.mass;
  h.soft = (float)g.soft;
  h.qxx = (float)g.qxx;
  h.qxy = (float)g.qxy;
  h.qxz = (float)g.qxz;
  h.qyy = (float)g.qyy;
  h.qyz = (float)g.qyz;
  h.qzz = (float)g.qzz;
  o.gn32[i] = h;
  if (a.type == PT_NODE_LEAF && a.count > 0) {
    Bucket b;
    b.lox = d.lo[0];
    b.loy = d.lo[1];
    b.loz = d.lo[2];
    b.hix = d.hi[0];
    b.hiy = d.hi[1];
    b.hiz = d.hi[2];
    b.cx = d.cm[0];
    b.cy = d.cm[1];
    b.cz = d.cm[2];
    b.soft = d.soft;
    b.pad0 = 0.0;
    b.pad1 = 0.0;
    b.node = (int32_t)i;
    b.begin = a.begin;
    b.count = a.count;
    b.pad2 = 0;
    o.bk[o.head_pos[a.begin]] = b;
  }
}

// One particle of the CentroidData leaf constructor: moments, box, and the
// largest softening.
__device__ __forceinline__ void leaf_add(CData& d, double& smax, double px, double py, double pz,
                                         double pm, double pe) {
  smax = pe > smax ? pe : smax;
  mom_add_particle(d, px, py, pz, pm, pe);
  const double p[3] = {px, py, pz};
  for (int k = 0; k < 3; k++) {
    if (p[k] < d.lo[k]) d.lo[k] = p[k];
    if (p[k] > d.hi[k]) d.hi[k] = p[k];
  }
}

// One thread per node: a leaf runs the sequential particle recurrence of the
// CentroidData leaf constructor in sorted particle order; an empty leaf stores
// the default CentroidData; internal nodes are left to the upward pass.  The
// node's walk/interaction records (and a leaf's bucket record) follow.
__global__ void __launch_bounds__(256)
leaf_kernel(const PtTree t, int64_t nn, const double* __restrict__ x,
            const double* __restrict__ y, const double* __restrict__ z,
            const double* __restrict__ m, const double* __restrict__ e, double* __restrict__ nd,
            const PtPackOut o) {
  const int64_t id = blockIdx.x * 256ll + threadIdx.x;
  if (id >= nn) return;
  const int ty = t.type[id];
  if (ty == PT_NODE_INTERNAL) return;
  CData d;
  cd_default(d);
  double smax = 0.0;
  if (ty == PT_NODE_LEAF) {
    const int64_t b = t.begin[id];
    const int64_t n = t.count[id];
    // the first kLeafPre particles' inputs are loaded together, then added
    // in order (the same sequence as one at a time, without a memory round
    // trip per particle)
    constexpr int kLeafPre = 16;
    double qx[kLeafPre], qy[kLeafPre], qz[kLeafPre], qm[kLeafPre], qe[kLeafPre];
#pragma unroll
    for (int k = 0; k < kLeafPre; k++) {
      if (k < n) {
        qx[k] = x[b + k];
        qy[k] = y[b + k];
        qz[k] = z[b + k];
        qm[k] = m[b + k];
        qe[k] = e[b + k];
      }
    }
#pragma unroll
    for (int k = 0; k < kLeafPre; k++) {
      if (k < n) leaf_add(d, smax, qx[k], qy[k], qz[k], qm[k], qe[k]);
    }
    for (int64_t i = b + kLeafPre; i < b + n; i++) leaf_add(d, smax, x[i], y[i], z[i], m[i], e[i]);
    d.count = n;
    if (d.count > 1) {
      d.radius = 0.5 * len3(d.hi[0] - d.lo[0], d.hi[1] - d.lo[1], d.hi[2] - d.lo[2]);
    } else {
      d.radius = 1.0;
    }
  }
  cd_store(d, nd, nn, id);
  pack_node(t, id, d, smax, o);
}

__device__ __forceinline__ uint64_t left_align(uint64_t k) {
  const int depth = (63 - __clzll(k)) / 3;
  return k << (PT_KEY_BITS - (3 * depth + 1));
}

// One internal node (its children are final).  The children's pop order is
// built in registers (fully unrolled loops over the 8 children, the order
// packed as 8 nibbles) so nothing spills to scratch.
// A child's CentroidData fields as the merge reads them: soft mass cm[3]
// q[6] lo[3] hi[3] (its radius is not used: the merge recomputes it).
constexpr int kUpFields = 17;
constexpr int kUpThreads = 64;  // threads (= parents) per work-group of the upward pass

// The thread's 8 children's fields are first loaded together into its LDS
// slab (every load in flight at once), then merged from there.
__device__ __forceinline__ void upward_node(const PtTree& t, int32_t p, double* __restrict__ nd,
                                            int64_t nn, int32_t* __restrict__ ord_h,
                                            uint64_t* __restrict__ ord_b,
                                            double (*__restrict__ slab)[kUpFields],
                                            const PtPackOut& o) {
  const int32_t fc = t.first_child[p];
  // every child's fields, type and order keys are loaded in one phase (an
  // empty child's stored defaults and a leaf's order keys are not used)
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int64_t ch = fc + c;
    slab[c][0] = nd[PT_CD_SOFT * nn + ch];
    slab[c][1] = nd[PT_CD_MASS * nn + ch];
#pragma unroll
    for (int k = 0; k < 3; k++) slab[c][2 + k] = nd[(PT_CD_CM + k) * nn + ch];
#pragma unroll
    for (int k = 0; k < 6; k++) slab[c][5 + k] = nd[(PT_CD_Q + k) * nn + ch];
#pragma unroll
    for (int k = 0; k < 3; k++) slab[c][11 + k] = nd[(PT_CD_LO + k) * nn + ch];
#pragma unroll
    for (int k = 0; k < 3; k++) slab[c][14 + k] = nd[(PT_CD_HI + k) * nn + ch];
  }
  int ctype[8];
  int32_t h[8];
  uint64_t b[8];
  uint64_t ckey[8];
  int n_leaf = 0, n_empty = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    ctype[c] = t.type[fc + c];
    ckey[c] = t.key[fc + c];
    const int32_t hc = ord_h[fc + c];
    const uint64_t bc = ord_b[fc + c];
    h[c] = ctype[c] == PT_NODE_INTERNAL ? hc : -1;
    b[c] = ctype[c] == PT_NODE_INTERNAL ? bc : 0;
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
      uint64_t r = left_align(ckey[c]) >> 1;
      if (ctype[c] == PT_NODE_EMPTY) r |= (1ull << 63);
      bl = r > bl ? r : bl;
    }
    ord_h[p] = 0;
    ord_b[p] = bl;
  } else {
    ord_h[p] = hmax + 1;
    ord_b[p] = bmax;
  }
  // the parent starts from the default CentroidData; an empty child is the
  // default too and is not read
  CData d;
  cd_default(d);
  for (int j = 0; j < 8; j++) {
    const int c = (order >> (4 * j)) & 15;
    CData cd;
    cd_default(cd);
    if (ctype[c] != PT_NODE_EMPTY) {
      const double* f = slab[c];
      cd.soft = f[0];
      cd.mass = f[1];
      for (int k = 0; k < 3; k++) cd.cm[k] = f[2 + k];
      for (int k = 0; k < 6; k++) cd.q[k] = f[5 + k];
      for (int k = 0; k < 3; k++) cd.lo[k] = f[11 + k];
      for (int k = 0; k < 3; k++) cd.hi[k] = f[14 + k];
      cd.count = t.count[fc + c];
    }
    cd_merge(d, cd);
  }
  cd_radius(d);
  cd_store(d, nd, nn, p);
  pack_node(t, p, d, 0.0, o);
}

// One thread per internal node of one level (deepest level first).
__global__ void __launch_bounds__(kUpThreads)
upward_kernel(const PtTree t, const int32_t* __restrict__ level_nodes, int32_t nl,
              double* __restrict__ nd, int64_t nn, int32_t* __restrict__ ord_h,
              uint64_t* __restrict__ ord_b, const PtPackOut o) {
  __shared__ double slab[kUpThreads][8][kUpFields];
  const int64_t li = (int64_t)blockIdx.x * kUpThreads + threadIdx.x;
  if (li >= nl) return;
  upward_node(t, level_nodes[li], nd, nn, ord_h, ord_b, slab[threadIdx.x], o);
}

// A run of consecutive small levels in one work-group, deepest first, with a
// work-group barrier between levels (a level reads what the one below wrote;
// all waves of the group share the CU's caches).
__global__ void __launch_bounds__(kUpThreads)
upward_run_kernel(const PtTree t, const int32_t* __restrict__ level_nodes, PtLevelRun run,
                  double* __restrict__ nd, int64_t nn, int32_t* __restrict__ ord_h,
                  uint64_t* __restrict__ ord_b, const PtPackOut o) {
  __shared__ double slab[kUpThreads][8][kUpFields];
  for (int k = 0; k < run.nlev; k++) {
    const int32_t off = run.off[k], cnt = run.cnt[k];
    for (int32_t li = threadIdx.x; li < cnt; li += kUpThreads)
      upward_node(t, level_nodes[off + li], nd, nn, ord_h, ord_b, slab[threadIdx.x], o);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// HEXADECAPOLE moments (the -DHEXADECAPOLE build, examples/Makefile:5).
//
// The reference keeps reduced (trace-free) moments to fourth order per node
// (FMOMR, examples/moments.h), built per particle and shifted to parents
// (examples/MultipoleMoments.h:244-252,317-330).  As written its leaf update
// runs only while the node radius is positive, and the leaf constructor sets
// the radius after adding the particles (examples/CentroidData.h:41-59), so
// every FMOMR stays zero.  These kernels compute what the code intends: the
// raw third and fourth moments R(a,b,c) = sum m dx^a dy^b dz^c about each
// node's centre of mass (leaves sum their particles in sorted order; a parent
// sums its children's moments moved to its centre of mass by the binomial
// expansion, children in index order), then the trace-free parts together
// with the second moments q of the quadrupole pass.  Same operations, in the
// same order, as the oracle's restatement.
// ---------------------------------------------------------------------------

// index of component (a, b, .) among those of order l (a descending, then b)
__device__ __forceinline__ int hx_index(int l, int a, int b) {
  return (l - a) * (l - a + 1) / 2 + (l - a - b);
}

__device__ __forceinline__ double hx_binom(int n, int k) {
  // rows 0-4 of Pascal's triangle
  constexpr double t[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0},
                              {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
  return t[n][k];
}

__device__ __forceinline__ void hx_pow(double v, double p[5]) {
  p[0] = 1.0;
  p[1] = v;
  p[2] = p[1] * v;
  p[3] = p[2] * v;
  p[4] = p[3] * v;
}

// One thread per bucket: raw moments about the leaf's centre of mass.
__global__ void __launch_bounds__(256)
hexa_leaf_kernel(const PtTree t, const int32_t* __restrict__ bucket_node, int64_t nb,
                 const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const double* __restrict__ m,
                 const double* __restrict__ nd, int64_t nn, double* __restrict__ raw) {
  const int64_t bi = blockIdx.x * 256ll + threadIdx.x;
  if (bi >= nb) return;
  const int32_t id = bucket_node[bi];
  const double cx = nd[(PT_CD_CM + 0) * nn + id], cy = nd[(PT_CD_CM + 1) * nn + id],
               cz = nd[(PT_CD_CM + 2) * nn + id];
  double r[PT_HEXA_RAW];
#pragma unroll
  for (int k = 0; k < PT_HEXA_RAW; k++) r[k] = 0.0;
  const int64_t b = t.begin[id], n = t.count[id];
  for (int64_t i = b; i < b + n; i++) {
    double px[5], py[5], pz[5];
    hx_pow(x[i] - cx, px);
    hx_pow(y[i] - cy, py);
    hx_pow(z[i] - cz, pz);
    const double mi = m[i];
#pragma unroll
    for (int l = 3; l <= 4; l++)
#pragma unroll
      for (int a = l; a >= 0; a--)
#pragma unroll
        for (int bb = l - a; bb >= 0; bb--) {
          const int c = l - a - bb;
          r[(l == 3 ? 0 : 10) + hx_index(l, a, bb)] += ((mi * px[a]) * py[bb]) * pz[c];
        }
  }
#pragma unroll
  for (int k = 0; k < PT_HEXA_RAW; k++) raw[k * nn + id] = r[k];
}

// raw moment of exponents (i, j, k) of a child: mass, second moments q
// (xx xy xz yy yz zz), third/fourth from its raw record
__device__ __forceinline__ double hx_child(const double* __restrict__ nd,
                                           const double* __restrict__ raw, int64_t nn,
                                           int64_t ch, double mass, int i, int j, int k) {
  const int l = i + j + k;
  if (l == 0) return mass;
  if (l == 2) return nd[(PT_CD_Q + hx_index(2, i, j)) * nn + ch];
  if (l == 3) return raw[hx_index(3, i, j) * nn + ch];
  return raw[(10 + hx_index(4, i, j)) * nn + ch];
}

// One thread per internal node of one level (deepest level first, after the
// quadrupole pass has fixed every centre of mass).
__global__ void __launch_bounds__(256)
hexa_up_kernel(const PtTree t, const int32_t* __restrict__ level_nodes, int32_t nl,
               const double* __restrict__ nd, int64_t nn, double* __restrict__ raw) {
  const int64_t li = blockIdx.x * 256ll + threadIdx.x;
  if (li >= nl) return;
  const int32_t p = level_nodes[li];
  double r[PT_HEXA_RAW];
#pragma unroll
  for (int k = 0; k < PT_HEXA_RAW; k++) r[k] = 0.0;
  const double pmass = nd[PT_CD_MASS * nn + p];
  if (pmass != 0.0) {
    const double pc[3] = {nd[(PT_CD_CM + 0) * nn + p], nd[(PT_CD_CM + 1) * nn + p],
                          nd[(PT_CD_CM + 2) * nn + p]};
    const int32_t fc = t.first_child[p];
    for (int c = 0; c < 8; c++) {
      const int64_t ch = fc + c;
      const double cm = nd[PT_CD_MASS * nn + ch];
      if (cm == 0.0) continue;
      double px[5], py[5], pz[5];
      hx_pow(nd[(PT_CD_CM + 0) * nn + ch] - pc[0], px);
      hx_pow(nd[(PT_CD_CM + 1) * nn + ch] - pc[1], py);
      hx_pow(nd[(PT_CD_CM + 2) * nn + ch] - pc[2], pz);
#pragma unroll
      for (int l = 3; l <= 4; l++)
#pragma unroll
        for (int a = l; a >= 0; a--)
#pragma unroll
          for (int b = l - a; b >= 0; b--) {
            const int cc = l - a - b;
            double sum = 0.0;
#pragma unroll
            for (int i = 0; i <= a; i++)
#pragma unroll
              for (int j = 0; j <= b; j++)
#pragma unroll
                for (int k = 0; k <= cc; k++) {
                  if (i + j + k == 1) continue;
                  const double coef = hx_binom(a, i) * hx_binom(b, j) * hx_binom(cc, k);
                  sum += (((coef * px[a - i]) * py[b - j]) * pz[cc - k]) *
                         hx_child(nd, raw, nn, ch, cm, i, j, k);
                }
            r[(l == 3 ? 0 : 10) + hx_index(l, a, b)] += sum;
          }
    }
  }
#pragma unroll
  for (int k = 0; k < PT_HEXA_RAW; k++) raw[k * nn + p] = r[k];
}

// Trace-free moments and the HEXADECAPOLE records of every node.
__global__ void __launch_bounds__(256)
hexa_pack_kernel(const PtTree t, const double* __restrict__ nd, const double* __restrict__ raw,
                 int64_t nn, HexNode* __restrict__ hn, HexOpen* __restrict__ ho) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= nn) return;
  const int ty = t.type[i];
  double r[PT_HEXA_RAW];
#pragma unroll
  for (int k = 0; k < PT_HEXA_RAW; k++) r[k] = ty == PT_NODE_EMPTY ? 0.0 : raw[k * nn + i];
  double q[6];
#pragma unroll
  for (int k = 0; k < 6; k++) q[k] = nd[(PT_CD_Q + k) * nn + i];
#define R3(a, b) r[hx_index(3, a, b)]
#define R4(a, b) r[10 + hx_index(4, a, b)]
  const double tr = q[0] + q[3] + q[5];
  const double vx = R3(3, 0) + R3(1, 2) + R3(1, 0);
  const double vy = R3(2, 1) + R3(0, 3) + R3(0, 1);
  const double vz = R3(2, 0) + R3(0, 2) + R3(0, 0);
  const double txx = R4(4, 0) + R4(2, 2) + R4(2, 0);
  const double tyy = R4(2, 2) + R4(0, 4) + R4(0, 2);
  const double tzz = R4(2, 0) + R4(0, 2) + R4(0, 0);
  const double txy = R4(3, 1) + R4(1, 3) + R4(1, 1);
  const double txz = R4(3, 0) + R4(1, 2) + R4(1, 0);
  const double tyz = R4(2, 1) + R4(0, 3) + R4(0, 1);
  const double ss = txx + tyy + tzz;
  HexNode h;
  h.cx = nd[(PT_CD_CM + 0) * nn + i];
  h.cy = nd[(PT_CD_CM + 1) * nn + i];
  h.cz = nd[(PT_CD_CM + 2) * nn + i];
  h.mass = nd[PT_CD_MASS * nn + i];
  h.soft = nd[PT_CD_SOFT * nn + i];
  h.h[0] = q[0] - tr * (1.0 / 3.0);
  h.h[1] = q[3] - tr * (1.0 / 3.0);
  h.h[2] = q[1];
  h.h[3] = q[2];
  h.h[4] = q[4];
  h.h[5] = R3(3, 0) - vx * (3.0 / 5.0);
  h.h[6] = R3(1, 2) - vx * (1.0 / 5.0);
  h.h[7] = R3(2, 1) - vy * (1.0 / 5.0);
  h.h[8] = R3(0, 3) - vy * (3.0 / 5.0);
  h.h[9] = R3(2, 0) - vz * (1.0 / 5.0);
  h.h[10] = R3(0, 2) - vz * (1.0 / 5.0);
  h.h[11] = R3(1, 1);
  h.h[12] = (R4(4, 0) - txx * (6.0 / 7.0)) + ss * (3.0 / 35.0);
  h.h[13] = R4(1, 3) - txy * (3.0 / 7.0);
  h.h[14] = R4(3, 1) - txy * (3.0 / 7.0);
  h.h[15] = (R4(0, 4) - tyy * (6.0 / 7.0)) + ss * (3.0 / 35.0);
  h.h[16] = R4(3, 0) - txz * (3.0 / 7.0);
  h.h[17] = R4(0, 3) - tyz * (3.0 / 7.0);
  h.h[18] = (R4(2, 2) - (txx + tyy) * (1.0 / 7.0)) + ss * (1.0 / 35.0);
  h.h[19] = R4(2, 1) - tyz * (1.0 / 7.0);
  h.h[20] = R4(1, 2) - txz * (1.0 / 7.0);
#undef R3
#undef R4
#pragma unroll
  for (int k = 0; k < 6; k++) h.pad[k] = 0.0;
  hn[i] = h;
  HexOpen o;
  o.lox = nd[(PT_CD_LO + 0) * nn + i];
  o.loy = nd[(PT_CD_LO + 1) * nn + i];
  o.loz = nd[(PT_CD_LO + 2) * nn + i];
  o.hix = nd[(PT_CD_HI + 0) * nn + i];
  o.hiy = nd[(PT_CD_HI + 1) * nn + i];
  o.hiz = nd[(PT_CD_HI + 2) * nn + i];
  o.soft = h.soft;
  o.pad = 0.0;
  ho[i] = o;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int pt_launch_leaf_moments(const PtTree& t, int64_t nn, const double* x, const double* y,
                           const double* z, const double* m, const double* e, double* nd,
                           const PtPackOut& o, hipStream_t s) {
  if (nn <= 0) return 0;
  leaf_kernel<<<grid_for(nn), 256, 0, s>>>(t, nn, x, y, z, m, e, nd, o);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_upward_run(const PtTree& t, const int32_t* level_nodes, const PtLevelRun& run,
                         double* nd, int64_t nn, int32_t* ord_h, uint64_t* ord_b,
                         const PtPackOut& o, hipStream_t s) {
  if (run.nlev <= 0) return 0;
  upward_run_kernel<<<1, kUpThreads, 0, s>>>(t, level_nodes, run, nd, nn, ord_h, ord_b, o);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_upward_level(const PtTree& t, const int32_t* level_nodes, int32_t nl, double* nd,
                           int64_t nn, int32_t* ord_h, uint64_t* ord_b, const PtPackOut& o,
                           hipStream_t s) {
  if (nl <= 0) return 0;
  upward_kernel<<<(unsigned)((nl + kUpThreads - 1) / kUpThreads), kUpThreads, 0, s>>>(
      t, level_nodes, nl, nd, nn, ord_h, ord_b, o);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}


int pt_launch_hexa_leaf(const PtTree& t, const int32_t* bucket_node, int64_t nb, const double* x,
                        const double* y, const double* z, const double* m, const double* nd,
                        int64_t nn, double* raw, hipStream_t s) {
  if (nb <= 0) return 0;
  hexa_leaf_kernel<<<grid_for(nb), 256, 0, s>>>(t, bucket_node, nb, x, y, z, m, nd, nn, raw);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_hexa_level(const PtTree& t, const int32_t* level_nodes, int32_t nl,
                         const double* nd, int64_t nn, double* raw, hipStream_t s) {
  if (nl <= 0) return 0;
  hexa_up_kernel<<<grid_for(nl), 256, 0, s>>>(t, level_nodes, nl, nd, nn, raw);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_hexa_pack(const PtTree& t, const double* nd, const double* raw, int64_t nn,
                        HexNode* hn, HexOpen* ho, hipStream_t s) {
  if (nn <= 0) return 0;
  hexa_pack_kernel<<<grid_for(nn), 256, 0, s>>>(t, nd, raw, nn, hn, ho);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

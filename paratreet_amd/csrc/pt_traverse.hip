// Down-traversal: interaction lists per target bucket.
//
// Reference: Partition::startDown<GravityVisitor> (src/Partition.h:129-134)
// -> TransposedDownTraverser::recurse (src/Traverser.h:154-217) with
// GravityVisitor::open (examples/GravityVisitor.h:162-182, non-HEXADECAPOLE
// branch :180).  For one bucket the reference visits the tree depth-first:
// a Leaf source gets leaf() (self included, CallSelfLeaf), an EmptyLeaf is
// skipped, an Internal source is open()-tested; opened -> children, closed ->
// node().  The set of leaf()/node() calls per bucket does not depend on visit
// order, so one wavefront per bucket walks the tree breadth-first in chunks:
// each step pops up to 8 opened nodes from an LDS stack and tests their 8
// children on the 64 lanes (children of a node are contiguous records, so the
// loads coalesce).  Leaves go to the P-P list, closed nodes to the P-C list,
// opened nodes back on the stack.  Counters follow doOpen/doLeaf/doNode
// (src/Traverser.h:15-37).
//
// Two launches per traversal: a count pass (list lengths + counters) and,
// after an exclusive scan, a fill pass that writes the node ids.  Compiled
// with -ffp-contract=off so the opening test rounds exactly like the
// reference's separately rounded expression.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kTravWaves = 4;
constexpr int kTravThreads = kTravWaves * PT_WAVE;

// Space::intersect(box, point, rsq): per-axis clamp distance, dsq <= rsq.
__device__ __forceinline__ bool intersect_box_point(const Bucket& b, double px, double py,
                                                    double pz, double rsq) {
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
  return dsq <= rsq;
}

// GravityVisitor::open (examples/GravityVisitor.h:162-164,180).
__device__ __forceinline__ bool gv_open(const TravNode& s, int32_t count, const Bucket& b,
                                        double gf, double ox, double oy, double oz) {
  if (count <= PT_MIN_PARTICLE_NODE) return true;
  const double dataRsq = s.radius * s.radius * gf;
  return intersect_box_point(b, s.cx + ox, s.cy + oy, s.cz + oz, dataRsq);
}

template <bool FILL>
__global__ void __launch_bounds__(kTravThreads)
traverse_kernel(const TravNode* __restrict__ tn, const TopoNode* __restrict__ topo,
                const Bucket* __restrict__ bk, int64_t nb, double gf, double ox, double oy,
                double oz, int32_t* __restrict__ pp_cnt, int32_t* __restrict__ pc_cnt,
                const int64_t* __restrict__ pp_off, const int64_t* __restrict__ pc_off,
                int32_t* __restrict__ pp, int32_t* __restrict__ pc,
                unsigned long long* __restrict__ counters, int32_t* __restrict__ err) {
  __shared__ int32_t stack[kTravWaves][PT_TRAV_STACK];
  const int wave = threadIdx.x / PT_WAVE;
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t bi = (int64_t)blockIdx.x * kTravWaves + wave;
  if (bi >= nb) return;
  const Bucket b = bk[bi];
  const int64_t nt = b.count;
  int32_t* st = stack[wave];

  int64_t npp = 0, npc = 0;
  unsigned long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  int32_t* ppw = FILL ? pp + pp_off[bi] : nullptr;
  int32_t* pcw = FILL ? pc + pc_off[bi] : nullptr;
  int sp = 0;

  // the root (node 0)
  {
    const TopoNode r = topo[0];
    if (r.type == PT_NODE_LEAF) {
      if (FILL && lane == 0) ppw[0] = 0;
      npp = 1;
      c_part += (unsigned long long)r.count * (unsigned long long)nt;
    } else if (r.type == PT_NODE_INTERNAL) {
      const bool op = gv_open(tn[0], r.count, b, gf, ox, oy, oz);
      if (op) {
        if (lane == 0) st[0] = 0;
        sp = 1;
        c_open++;
      } else {
        if (FILL && lane == 0) pcw[0] = 0;
        npc = 1;
        c_close++;
        c_node += (unsigned long long)nt;
      }
    }
  }

  bool overflow = false;
  while (sp > 0) {
    const int k = sp < 8 ? sp : 8;
    sp -= k;
    const int slot = lane >> 3;
    const bool active = slot < k;
    int32_t node = -1;
    int type = PT_NODE_EMPTY;
    int32_t cnt = 0;
    bool op = false;
    if (active) {
      const int32_t parent = st[sp + slot];
      node = topo[parent].first_child + (lane & 7);
      const TopoNode tp = topo[node];
      type = tp.type;
      cnt = tp.count;
      if (type == PT_NODE_INTERNAL) op = gv_open(tn[node], cnt, b, gf, ox, oy, oz);
    }
    const unsigned long long leafm = __ballot(active && type == PT_NODE_LEAF);
    const unsigned long long intm = __ballot(active && type == PT_NODE_INTERNAL);
    const unsigned long long openm = __ballot(active && type == PT_NODE_INTERNAL && op);
    const unsigned long long closem = intm & ~openm;
    const int n_open = __popcll(openm);
    if (sp + n_open > PT_TRAV_STACK) {
      overflow = true;
      break;
    }
    // keep every lane's LDS reads of this step ahead of the pushes below
    __builtin_amdgcn_wave_barrier();
    if (openm & (1ull << lane)) st[sp + pt_rank_in(openm)] = node;
    __builtin_amdgcn_wave_barrier();
    sp += n_open;
    if (FILL) {
      if (leafm & (1ull << lane)) ppw[npp + pt_rank_in(leafm)] = node;
      if (closem & (1ull << lane)) pcw[npc + pt_rank_in(closem)] = node;
    }
    npp += __popcll(leafm);
    npc += __popcll(closem);
    c_open += n_open;
    c_close += __popcll(closem);
    c_node += (unsigned long long)__popcll(closem) * (unsigned long long)nt;
    // n_part_ints: sum over leaf lanes of count * nt
    int lc = (leafm & (1ull << lane)) ? cnt : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lc += __shfl_xor(lc, d, PT_WAVE);
    c_part += (unsigned long long)lc * (unsigned long long)nt;
  }

  if (lane == 0) {
    if (overflow) atomicExch(err, 1);
    if (!FILL) {
      pp_cnt[bi] = (int32_t)npp;
      pc_cnt[bi] = (int32_t)npc;
      if (counters) {
        atomicAdd(&counters[0], c_node);
        atomicAdd(&counters[1], c_part);
        atomicAdd(&counters[2], c_open);
        atomicAdd(&counters[3], c_close);
      }
    }
  }
}

}  // namespace

int pt_launch_traverse(bool fill, const TravNode* tn, const TopoNode* topo, const Bucket* bk,
                       int64_t nb, double gf, const double off[3], int32_t* pp_cnt,
                       int32_t* pc_cnt, const int64_t* pp_off, const int64_t* pc_off,
                       int32_t* pp, int32_t* pc, unsigned long long* counters, int32_t* err,
                       hipStream_t s) {
  if (nb <= 0) return 0;
  const unsigned grid = (unsigned)((nb + kTravWaves - 1) / kTravWaves);
  if (fill) {
    traverse_kernel<true><<<grid, kTravThreads, 0, s>>>(tn, topo, bk, nb, gf, off[0], off[1],
                                                        off[2], pp_cnt, pc_cnt, pp_off, pc_off,
                                                        pp, pc, counters, err);
  } else {
    traverse_kernel<false><<<grid, kTravThreads, 0, s>>>(tn, topo, bk, nb, gf, off[0], off[1],
                                                         off[2], pp_cnt, pc_cnt, pp_off, pc_off,
                                                         pp, pc, counters, err);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

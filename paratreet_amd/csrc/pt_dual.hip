// Dual-tree traversal for the gravity visitor, breadth first over node pairs.
//
// Reference: DualTraverser<Data, Visitor> (src/Traverser.h:485-610): doTrav
// walks (source, target) node pairs from Subtree::startDual's pair
// (cm_local->root, local_root) (src/Subtree.h:233-240), and
// runInvertedTraversal walks a target subtree against one source; with
// GravityVisitor's flags TargetMustBeLeaf and ForceEvenDepth and its
// open()/cell() (examples/GravityVisitor.h:11-13,162-182,226-230); the
// counts of doOpen/doLeaf/doNode/doCell (src/Traverser.h:15-45).
//
// The set of visitor calls and the counts do not depend on the order the
// reference's stacks visit the pairs in, so the pairs of one depth of the
// walk are processed together: one thread per pair classifies it exactly as
// doTrav (or, for an inverted pair, as the loop body of runInvertedTraversal)
// would, appends the pairs it pushes to the next frontier and the leaf() /
// node() calls that move particles to two record arrays (target bucket <<
// 32 | source node).  The records are then sorted into per-bucket lists in
// the format of the down traversal's lists, and the list interaction kernel
// evaluates them.
//
// What leaf() and node() do follows the reference's SpatialNode::n_particles
// (src/Node.h:97-107): -1 for an internal node, 0 for an empty leaf.  leaf()
// moves particles only between two leaves with particles and node() only
// onto a leaf; the counters add n_source * n_target and n_target with those
// values, modulo 2^64 like the reference's unsigned long long counters.
//
// Compiled with -ffp-contract=off (opening test rounding).
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

constexpr int kDualThreads = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kDualThreads - 1) / kDualThreads); }

// Frontier item: source node << 32 | target node, bit 31 marking a pair
// inside runInvertedTraversal.
constexpr uint32_t kInverted = 0x80000000u;

__device__ __forceinline__ uint64_t pack(int32_t s, int32_t t, bool inverted) {
  return ((uint64_t)(uint32_t)s << 32) | (uint32_t)t | (inverted ? kInverted : 0u);
}

struct DualNodes {
  const TreeNode* tn;
  const double* nd;      // CentroidData fields, field-major (PT_CD_*)
  const int32_t* depth;
  const int32_t* bucket_of;
  int64_t nn;
};

// Space::intersect(box, point, rsq) with the box of node t: per-axis clamp
// distance squared, x then y then z, dsq <= rsq.
__device__ __forceinline__ bool intersect_node_box(const DualNodes& N, int32_t t, double px,
                                                   double py, double pz, double rsq) {
  const double p[3] = {px, py, pz};
  double dsq = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double lo = N.nd[(PT_CD_LO + k) * N.nn + t];
    const double hi = N.nd[(PT_CD_HI + k) * N.nn + t];
    if (p[k] < lo) {
      const double d = lo - p[k];
      dsq += d * d;
    } else if (p[k] > hi) {
      const double d = p[k] - hi;
      dsq += d * d;
    }
  }
  return dsq <= rsq;
}

// GravityVisitor::open (examples/GravityVisitor.h:162-164,180).
__device__ __forceinline__ bool gv_open(const DualNodes& N, const TreeNode& s, int32_t t,
                                        double gf, double ox, double oy, double oz) {
  if (s.count <= PT_MIN_PARTICLE_NODE) return true;
  const double dataRsq = s.radius * s.radius * gf;
  return intersect_node_box(N, t, s.cx + ox, s.cy + oy, s.cz + oz, dataRsq);
}

// GravityVisitor::cell (examples/GravityVisitor.h:226-230): open the target
// unless the source's box encloses it (lesser_s <= lesser_t and
// greater_t <= greater_s on every axis).
__device__ __forceinline__ bool gv_cell(const DualNodes& N, int32_t s, int32_t t) {
  bool enclose = true;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    enclose = enclose && N.nd[(PT_CD_LO + k) * N.nn + s] <= N.nd[(PT_CD_LO + k) * N.nn + t] &&
              N.nd[(PT_CD_HI + k) * N.nn + t] <= N.nd[(PT_CD_HI + k) * N.nn + s];
  }
  return !enclose;
}

__device__ __forceinline__ int64_t n_particles(const TreeNode& n) {
  return n.type == PT_NODE_INTERNAL ? -1 : (n.type == PT_NODE_LEAF ? n.count : 0);
}

// state: [0] next frontier size, [1] leaf() records, [2] node() records,
// [3..6] counters {n_node_ints, n_part_ints, n_opens, n_closes} of this step.
__global__ void __launch_bounds__(kDualThreads)
dual_step_kernel(const uint64_t* __restrict__ items, int64_t n_items, DualNodes N, double gf,
                 double ox, double oy, double oz, uint64_t* __restrict__ next, int64_t next_cap,
                 uint64_t* __restrict__ pp_rec, int64_t pp_cap, uint64_t* __restrict__ pc_rec,
                 int64_t pc_cap, unsigned long long* __restrict__ state) {
  __shared__ unsigned long long s_cnt[4];
  if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)kDualThreads + threadIdx.x;
  long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  if (i < n_items) {
    const uint64_t it = items[i];
    const int32_t s = (int32_t)(it >> 32);
    const uint32_t lo = (uint32_t)it;
    const bool inverted = (lo & kInverted) != 0;
    const int32_t t = (int32_t)(lo & ~kInverted);
    const TreeNode sn = N.tn[s];
    const TreeNode tn = N.tn[t];
    // pushes of this pair: `np` pairs (source + i or the same source) x
    // (target + j or the same target)
    int np = 0;
    int32_t ps = s, pt = t;
    bool ps_children = false, pt_children = false, p_inv = false;
    // leaf() / node() that move particles
    bool leaf_call = false, node_call = false;
    if (inverted) {
      // runInvertedTraversal's loop body for target node t
      if (tn.type != PT_NODE_INTERNAL) {
        c_part += n_particles(sn) * n_particles(tn);
        leaf_call = sn.type == PT_NODE_LEAF && tn.type == PT_NODE_LEAF;
      } else if (gv_open(N, sn, t, gf, ox, oy, oz)) {
        c_open++;
        np = 8;
        pt = tn.first_child;
        pt_children = true;
        p_inv = true;
      } else {
        c_close++;
        c_node += -1;  // node() onto an internal node: n_particles = -1
      }
    } else if (tn.type != PT_NODE_EMPTY) {
      if (sn.type == PT_NODE_LEAF) {
        if (tn.type == PT_NODE_LEAF) {
          c_part += (long long)sn.count * tn.count;
          leaf_call = true;
        } else {
          np = 1;
          p_inv = true;  // runInvertedTraversal(source leaf, target)
        }
      } else if (sn.type == PT_NODE_INTERNAL) {
        if (tn.type == PT_NODE_LEAF) {
          if (gv_open(N, sn, t, gf, ox, oy, oz)) {
            c_open++;
            np = 8;
            ps = sn.first_child;
            ps_children = true;
          } else {
            c_close++;
            c_node += tn.count;
            node_call = sn.count != 0;
          }
        } else if (!gv_cell(N, s, t)) {
          c_close++;
          np = 1;
          p_inv = true;  // runInvertedTraversal(source node, target)
        } else {
          c_open++;
          ps = sn.first_child;
          ps_children = true;
          if (N.depth[t] <= N.depth[s]) {
            np = 64;
            pt = tn.first_child;
            pt_children = true;
          } else {
            np = 8;
          }
        }
      }
      // an EmptyLeaf source: nothing (the switch's default)
    }
    if (np > 0) {
      const unsigned long long at = atomicAdd(&state[0], (unsigned long long)np);
      for (int k = 0; k < np; k++) {
        const unsigned long long w = at + k;
        if ((int64_t)w < next_cap) {
          const int32_t a = ps_children ? ps + (pt_children && np == 64 ? k >> 3 : k) : ps;
          const int32_t b = pt_children ? pt + (np == 64 ? (k & 7) : k) : pt;
          next[w] = pack(a, b, p_inv);
        }
      }
    }
    if (leaf_call || node_call) {
      const int32_t b = N.bucket_of[t];
      const uint64_t rec = ((uint64_t)(uint32_t)b << 32) | (uint32_t)s;
      const unsigned long long at = atomicAdd(&state[leaf_call ? 1 : 2], 1ull);
      if (leaf_call && (int64_t)at < pp_cap) pp_rec[at] = rec;
      if (node_call && (int64_t)at < pc_cap) pc_rec[at] = rec;
    }
  }
  if (c_node) atomicAdd(&s_cnt[0], (unsigned long long)c_node);
  if (c_part) atomicAdd(&s_cnt[1], (unsigned long long)c_part);
  if (c_open) atomicAdd(&s_cnt[2], (unsigned long long)c_open);
  if (c_close) atomicAdd(&s_cnt[3], (unsigned long long)c_close);
  __syncthreads();
  if (threadIdx.x < 4 && s_cnt[threadIdx.x]) atomicAdd(&state[3 + threadIdx.x], s_cnt[threadIdx.x]);
}

__global__ void __launch_bounds__(kDualThreads)
bucket_of_kernel(const Bucket* __restrict__ bk, int64_t nb, int32_t* __restrict__ bucket_of) {
  const int64_t b = blockIdx.x * (int64_t)kDualThreads + threadIdx.x;
  if (b < nb) bucket_of[bk[b].node] = (int32_t)b;
}

__global__ void __launch_bounds__(kDualThreads)
find_keys_kernel(const uint64_t* __restrict__ key, int64_t nn, uint64_t k0, uint64_t k1,
                 int32_t* __restrict__ ids) {
  const int64_t i = blockIdx.x * (int64_t)kDualThreads + threadIdx.x;
  if (i >= nn) return;
  if (key[i] == k0) ids[0] = (int32_t)i;
  if (key[i] == k1) ids[1] = (int32_t)i;
}

// Per-bucket lists from records sorted by (bucket, source): the source ids in
// order, and off[b] = the first record of bucket b (b = 0 .. nb).
__global__ void __launch_bounds__(kDualThreads)
records_to_list_kernel(const uint64_t* __restrict__ rec, int64_t n, int32_t* __restrict__ ids) {
  const int64_t i = blockIdx.x * (int64_t)kDualThreads + threadIdx.x;
  if (i < n) ids[i] = (int32_t)(uint32_t)rec[i];
}

__global__ void __launch_bounds__(kDualThreads)
records_offsets_kernel(const uint64_t* __restrict__ rec, int64_t n, int64_t nb,
                       int64_t* __restrict__ off) {
  const int64_t b = blockIdx.x * (int64_t)kDualThreads + threadIdx.x;
  if (b > nb) return;
  const uint64_t key = (uint64_t)b << 32;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (rec[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  off[b] = lo;
}

}  // namespace

int pt_launch_dual_bucket_of(const Bucket* bk, int64_t nb, int32_t* bucket_of, hipStream_t s) {
  if (nb <= 0) return 0;
  bucket_of_kernel<<<grid_for(nb), kDualThreads, 0, s>>>(bk, nb, bucket_of);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_dual_find(const uint64_t* key, int64_t nn, uint64_t k0, uint64_t k1, int32_t* ids,
                        hipStream_t s) {
  if (nn <= 0) return 0;
  find_keys_kernel<<<grid_for(nn), kDualThreads, 0, s>>>(key, nn, k0, k1, ids);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_dual_step(const uint64_t* items, int64_t n_items, const TreeNode* tn,
                        const double* nd, const int32_t* depth, const int32_t* bucket_of,
                        int64_t nn, double gf, const double off[3], uint64_t* next,
                        int64_t next_cap, uint64_t* pp_rec, int64_t pp_cap, uint64_t* pc_rec,
                        int64_t pc_cap, unsigned long long* state, hipStream_t s) {
  if (n_items <= 0) return 0;
  const DualNodes N{tn, nd, depth, bucket_of, nn};
  dual_step_kernel<<<grid_for(n_items), kDualThreads, 0, s>>>(
      items, n_items, N, gf, off[0], off[1], off[2], next, next_cap, pp_rec, pp_cap, pc_rec,
      pc_cap, state);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_records_to_lists(const uint64_t* rec, int64_t n, int64_t nb, int32_t* ids,
                               int64_t* off, hipStream_t s) {
  if (n > 0) records_to_list_kernel<<<grid_for(n), kDualThreads, 0, s>>>(rec, n, ids);
  records_offsets_kernel<<<grid_for(nb + 1), kDualThreads, 0, s>>>(rec, n, nb, off);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

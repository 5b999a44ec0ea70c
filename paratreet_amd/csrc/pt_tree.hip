// Level-parallel octree build over sorted Morton keys.
//
// Reference: Subtree::buildTree / recursiveBuild / getType (src/Subtree.h:
// 262-349), OctTree::findChildsLastParticle (src/Modularization.C:60-65) and
// the key helpers of src/Utility.h:9-130.  The reference recurses depth-first;
// here every internal node of one level is split by one thread per
// (node, child): child i of a node covers [first key >= left-aligned(child
// key), first key >= left-aligned(child key + 1)), the last child taking the
// remainder, which is exactly the range recursiveBuild assigns.  The 8
// children of a node get consecutive ids, so a node stores only its first
// child.  Node ids are in level order; node identity across implementations
// is the key.
#include "pt_common.h"
#include "pt_kernels.h"

namespace {

// Utility::mssb64_pos (src/Utility.h:9-21) as a count-leading-zeros.
__device__ __forceinline__ int mssb64_pos(uint64_t x) { return x ? 63 - __clzll(x) : -1; }

// Utility::removeLeadingZeros (src/Utility.h:127-130) for branch factor 8.
__device__ __forceinline__ uint64_t left_align(uint64_t k) {
  const int depth = mssb64_pos(k) / 3;
  return k << (PT_KEY_BITS - (3 * depth + 1));
}

// Utility::binarySearchComp with a >= comparison (src/Utility.h:34-50).
__device__ __forceinline__ int64_t first_ge(const uint64_t* __restrict__ k, int64_t lo,
                                            int64_t hi, uint64_t s) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (k[mid] >= s) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Subtree::getType (src/Subtree.h:262-267).  At the last key level no split
// is possible, so an over-full node there stays a leaf.
__device__ __forceinline__ int node_type(int64_t n, int bucket, int depth) {
  if (n == 0) return PT_NODE_EMPTY;
  if (n <= bucket) return PT_NODE_LEAF;
  if (depth >= PT_MAX_DEPTH) return PT_NODE_LEAF;
  return PT_NODE_INTERNAL;
}

__global__ void root_kernel(int64_t n, int bucket, PtTree t, int32_t* __restrict__ level_cnt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int l = 0; l <= PT_MAX_DEPTH + 1; l++) level_cnt[l] = 0;
  level_cnt[0] = node_type(n, bucket, 0) == PT_NODE_INTERNAL ? 1 : 0;
  t.key[0] = 1;
  t.begin[0] = 0;
  t.count[0] = (int32_t)n;
  t.depth[0] = 0;
  t.type[0] = node_type(n, bucket, 0);
  t.parent[0] = -1;
  t.first_child[0] = -1;
}

// Device-driven level split: level d's internal nodes are
// level_nodes[off_d, off_d + level_cnt[d]) with off_d = sum of the counts of
// the levels above; their children get ids 1 + 8 * off_d + 8 f + c, and the
// internal children are appended to level d + 1 with an atomic counter (their
// order there is irrelevant: children stay contiguous per parent, merge order
// and walk order depend only on the tree).  Writes beyond the capacities set
// *overflow instead.
__global__ void __launch_bounds__(256)
split_level_kernel(const uint64_t* __restrict__ keys, int32_t* __restrict__ level_nodes,
                   int32_t* __restrict__ level_cnt, int d, int bucket, PtTree t,
                   int64_t node_cap, int64_t level_cap, int32_t* __restrict__ overflow) {
  int64_t f0 = 0;
  for (int l = 0; l < d; l++) f0 += level_cnt[l];
  const int64_t nf = level_cnt[d];
  const int64_t gid = blockIdx.x * 256ll + threadIdx.x;
  if (gid >= nf * 8) return;
  const int64_t f = gid >> 3;
  const int c = (int)(gid & 7);
  const int64_t base = 1 + 8 * f0;
  const int64_t id = base + gid;
  // after an overflow at a shallower level the counts run past what was
  // recorded: nothing beyond the capacities is read or written
  if (id >= node_cap || f0 + f >= level_cap) {
    atomicExch(overflow, 1);
    return;
  }
  const int32_t p = level_nodes[f0 + f];
  const int64_t b = t.begin[p];
  const int64_t e = b + t.count[p];
  const uint64_t child_key = (t.key[p] << 3) | (uint64_t)c;
  const int64_t lo = (c == 0) ? b : first_ge(keys, b, e, left_align(child_key));
  int64_t hi = (c == 7) ? e : first_ge(keys, b, e, left_align(child_key + 1));
  // ordered keys give lo <= hi; keys ordered on fewer bits than this split
  // reads (a top-bits sort the host then rejects) must still give in-range,
  // non-negative counts
  if (hi < lo) hi = lo;
  const int ty = node_type(hi - lo, bucket, d + 1);
  t.key[id] = child_key;
  t.begin[id] = (int32_t)lo;
  t.count[id] = (int32_t)(hi - lo);
  t.depth[id] = d + 1;
  t.type[id] = ty;
  t.parent[id] = p;
  t.first_child[id] = -1;
  if (c == 0) t.first_child[p] = (int32_t)(base + 8 * f);
  if (ty == PT_NODE_INTERNAL) {
    const int64_t slot = f0 + nf + atomicAdd(&level_cnt[d + 1], 1);
    if (slot >= level_cap) {
      atomicExch(overflow, 1);
      return;
    }
    level_nodes[slot] = (int32_t)id;
  }
}

// Node count of the finished levels (one thread): status[0] = 1 + 8 * (all
// internal nodes), or 0 when the build overflowed or is incomplete (internal
// nodes at depth `levels`, which no launched split handled): nothing
// downstream runs on such a tree.
__global__ void count_nodes_kernel(const int32_t* __restrict__ level_cnt,
                                   const int32_t* __restrict__ overflow, int64_t node_cap,
                                   int levels, int64_t* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t internal = 0;
  for (int l = 0; l <= PT_MAX_DEPTH + 1; l++) internal += level_cnt[l];
  const int64_t nn = 1 + 8 * internal;
  const bool incomplete = levels <= PT_MAX_DEPTH && level_cnt[levels] > 0;
  status[0] = (*overflow || nn > node_cap || incomplete) ? 0 : nn;
}

// Collects what the host reads back after the build in one place:
// {nn, nb, overflow, level counts}.
__global__ void tree_status_kernel(const int32_t* __restrict__ level_cnt,
                                   const int32_t* __restrict__ overflow,
                                   const uint32_t* __restrict__ nb, int64_t* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  status[1] = status[0] > 0 ? (int64_t)*nb : 0;
  status[2] = *overflow;
  for (int l = 0; l <= PT_MAX_DEPTH + 1; l++) status[3 + l] = level_cnt[l];
}

// Grids cover the node capacity; the live count is status[0].
__global__ void __launch_bounds__(256)
leaf_heads_kernel(const PtTree t, const int64_t* __restrict__ status,
                  uint32_t* __restrict__ head) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < status[0] && t.type[i] == PT_NODE_LEAF) head[t.begin[i]] = 1u;
}

__global__ void __launch_bounds__(256)
bucket_ids_kernel(const PtTree t, const int64_t* __restrict__ status,
                  const uint32_t* __restrict__ head_pos, int32_t* __restrict__ bucket_node) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < status[0] && t.type[i] == PT_NODE_LEAF) bucket_node[head_pos[t.begin[i]]] = (int32_t)i;
}

// After a sort of the keys' top bits only (ties kept in input order), puts
// every leaf's particles in full (key, input index) order: the sort the rest
// of the path expects.  Exact when every split of the build used only sorted
// bits, i.e. each leaf holds whole groups of equal sorted bits.  One thread
// per bucket (non-empty leaf; status[1] of them): up to 16 particles are
// sorted in registers by an odd-even transposition network, up to kFixupMax
// by insertion in place.
__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
  return ka < kb || (ka == kb && va < vb);
}

__global__ void __launch_bounds__(256)
leaf_fixup_kernel(const PtTree t, const int32_t* __restrict__ bucket_node,
                  const int64_t* __restrict__ status, uint64_t* __restrict__ keys,
                  uint32_t* __restrict__ perm) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= status[1]) return;
  const int32_t leaf = bucket_node[i];
  const int64_t b = t.begin[leaf];
  const int32_t n = t.count[leaf];
  if (n < 2 || n > kFixupMax) return;
  if (n <= 16) {
    uint64_t k[16];
    uint32_t v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      k[j] = j < n ? keys[b + j] : ~0ull;
      v[j] = j < n ? perm[b + j] : ~0u;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
      for (int j = r & 1; j + 1 < 16; j += 2) {
        if (pair_less(k[j + 1], v[j + 1], k[j], v[j])) {
          const uint64_t tk = k[j];
          k[j] = k[j + 1];
          k[j + 1] = tk;
          const uint32_t tv = v[j];
          v[j] = v[j + 1];
          v[j + 1] = tv;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (j < n) {
        keys[b + j] = k[j];
        perm[b + j] = v[j];
      }
    }
    return;
  }
  for (int j = 1; j < n; j++) {
    const uint64_t kj = keys[b + j];
    const uint32_t vj = perm[b + j];
    int q = j - 1;
    while (q >= 0 && pair_less(kj, vj, keys[b + q], perm[b + q])) {
      keys[b + q + 1] = keys[b + q];
      perm[b + q + 1] = perm[b + q];
      q--;
    }
    keys[b + q + 1] = kj;
    perm[b + q + 1] = vj;
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int pt_launch_root(int64_t n, int bucket, const PtTree& t, int32_t* level_cnt, hipStream_t s) {
  root_kernel<<<1, 64, 0, s>>>(n, bucket, t, level_cnt);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_leaf_fixup(const PtTree& t, int64_t max_buckets, const int32_t* bucket_node,
                         const int64_t* status, uint64_t* keys, uint32_t* perm, hipStream_t s) {
  if (max_buckets <= 0) return 0;
  leaf_fixup_kernel<<<grid_for(max_buckets), 256, 0, s>>>(t, bucket_node, status, keys, perm);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_split_level(const uint64_t* keys, int32_t* level_nodes, int32_t* level_cnt, int d,
                          int64_t max_frontier, int bucket, const PtTree& t, int64_t node_cap,
                          int64_t level_cap, int32_t* overflow, hipStream_t s) {
  if (max_frontier <= 0) return 0;
  split_level_kernel<<<grid_for(max_frontier * 8), 256, 0, s>>>(keys, level_nodes, level_cnt, d,
                                                                 bucket, t, node_cap, level_cap,
                                                                 overflow);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}


// Buckets of a device-built tree without a host round trip: the node count
// comes from the level counts on the device (status[0]); afterwards status =
// {nn, nb, overflow, level counts [PT_MAX_DEPTH + 2]} for one read-back.
int pt_launch_buckets(const PtTree& t, int64_t node_cap, int64_t n, int levels,
                      const int32_t* level_cnt, const int32_t* overflow, uint32_t* head,
                      uint32_t* head_pos, void* scan_scratch, uint32_t* nb_out,
                      int32_t* bucket_node, int64_t* status, hipStream_t s) {
  count_nodes_kernel<<<1, 64, 0, s>>>(level_cnt, overflow, node_cap, levels, status);
  if (n > 0) {
    if (pt_launch_fill_u32(head, n, 0u, s) != 0) return -1;
    leaf_heads_kernel<<<grid_for(node_cap), 256, 0, s>>>(t, status, head);
    if (pt_exclusive_scan_u32(head, head_pos, n, scan_scratch, nb_out, s) != 0) return -1;
    bucket_ids_kernel<<<grid_for(node_cap), 256, 0, s>>>(t, status, head_pos, bucket_node);
  } else {
    if (pt_launch_fill_u32(nb_out, 1, 0u, s) != 0) return -1;
  }
  tree_status_kernel<<<1, 64, 0, s>>>(level_cnt, overflow, nb_out, status);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

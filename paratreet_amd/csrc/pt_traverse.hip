// Down-traversal producing per-bucket interaction lists.
//
// Reference: Partition::startDown<GravityVisitor> (src/Partition.h:129-134)
// -> TransposedDownTraverser::recurse (src/Traverser.h:154-217); counters of
// doOpen/doLeaf/doNode (src/Traverser.h:15-37).  The walk itself is
// pt_walk.h (shared with the fused kernel).  Two launches: a count pass (list
// lengths, counters) and, after an exclusive scan, a fill pass writing the
// node ids.  The lists feed pt_interact, the parity read-back and the
// adjoint; pt_compute_forces uses the fused kernel instead.
//
// Compiled with -ffp-contract=off (opening test rounding).
#include "pt_common.h"
#include "pt_kernels.h"
#include "pt_walk.h"

namespace {

constexpr int kTravWaves = 4;
constexpr int kTravThreads = kTravWaves * PT_WAVE;

template <bool FILL, bool HEXA>
__global__ void __launch_bounds__(kTravThreads)
traverse_kernel(const TreeNode* __restrict__ nodes, const HexOpen* __restrict__ ho,
                const Bucket* __restrict__ bk, int64_t nb, double gf, double gm, double ox,
                double oy, double oz, int32_t* __restrict__ pp_cnt,
                int32_t* __restrict__ pc_cnt, const int64_t* __restrict__ pp_off,
                const int64_t* __restrict__ pc_off, int32_t* __restrict__ pp,
                int32_t* __restrict__ pc, unsigned long long* __restrict__ partials,
                int32_t* __restrict__ err) {
  __shared__ int32_t stack[kTravWaves][PT_WALK_STACK];
  __shared__ unsigned long long s_cnt[kTravWaves][4];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t blk = pt_xcd_block(blockIdx.x, gridDim.x);
  const int64_t bi_raw = blk * kTravWaves + wave;
  const bool live = bi_raw < nb;
  const int64_t bi = live ? bi_raw : nb - 1;
  const Bucket b = bk[bi];
  const int64_t nt = b.count;
  int32_t* st = stack[wave];

  int sp = 0;
  bool overflow = false;
  unsigned long long c_node = 0, c_part = 0, c_open = 0, c_close = 0;
  int64_t npp = 0, npc = 0;
  int32_t* ppw = FILL ? pp + pp_off[bi] : nullptr;
  int32_t* pcw = FILL ? pc + pc_off[bi] : nullptr;
  if (live) {
    WalkCounts wc;
    const WalkLane r = pt_walk_root<HEXA>(nodes, b, gf, ox, oy, oz, st, sp, wc, ho, gm);
    if (r.is_leaf) {
      if (FILL && lane == 0) ppw[0] = 0;
      npp = 1;
    } else if (r.is_closed) {
      if (FILL && lane == 0) pcw[0] = 0;
      npc = 1;
    }
    c_open += wc.n_open;
    c_close += wc.n_close;
    c_node += (unsigned long long)wc.n_close * nt;
    c_part += (unsigned long long)wc.leaf_parts * nt;
    while (sp > 0) {
      unsigned long long leafm, closem;
      const WalkLane r2 = pt_walk_step<PT_WALK_STACK, HEXA>(nodes, b, gf, ox, oy, oz, st, sp,
                                                            overflow, wc, leafm, closem, ho, gm);
      if (FILL) {
        if (r2.is_leaf) ppw[npp + pt_rank_in(leafm)] = r2.node;
        if (r2.is_closed) pcw[npc + pt_rank_in(closem)] = r2.node;
      }
      npp += __popcll(leafm);
      npc += __popcll(closem);
      c_open += wc.n_open;
      c_close += wc.n_close;
      c_node += (unsigned long long)wc.n_close * nt;
      c_part += (unsigned long long)wc.leaf_parts * nt;
    }
  }
  if (lane == 0) {
    if (live) {
      if (overflow) atomicExch(err, 1);
      if (!FILL) {
        pp_cnt[bi] = (int32_t)npp;
        pc_cnt[bi] = (int32_t)npc;
      }
    }
    s_cnt[wave][0] = c_node;
    s_cnt[wave][1] = c_part;
    s_cnt[wave][2] = c_open;
    s_cnt[wave][3] = c_close;
  }
  if (!FILL) {
    // one partial per block (a global atomic per wavefront would serialise
    // hundreds of thousands of updates on four addresses)
    __syncthreads();
    if (threadIdx.x < 4) {
      unsigned long long v = 0;
      for (int w = 0; w < kTravWaves; w++) v += s_cnt[w][threadIdx.x];
      partials[blk * 4 + threadIdx.x] = v;
    }
  }
}

// Sums the per-block counter partials: each block reduces a strided share and
// adds it to `out` (zeroed beforehand) with four atomics.
__global__ void __launch_bounds__(256)
reduce_counters_kernel(const unsigned long long* __restrict__ part, int64_t nblocks,
                       unsigned long long* __restrict__ out, const int32_t* __restrict__ err) {
  // out[4]: the walk's overflow flag, so one read-back returns both
  if (blockIdx.x == 0 && threadIdx.x == 0) out[4] = err ? (unsigned long long)*err : 0ull;
  __shared__ unsigned long long s[4][256];
  unsigned long long v[4] = {0, 0, 0, 0};
  for (int64_t b = blockIdx.x * 256ll + threadIdx.x; b < nblocks; b += (int64_t)gridDim.x * 256)
    for (int k = 0; k < 4; k++) v[k] += part[b * 4 + k];
  for (int k = 0; k < 4; k++) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) atomicAdd(&out[threadIdx.x], s[threadIdx.x][0]);
}

// Groups of consecutive buckets (see pt_group_slots): one thread per window
// of PT_GROUP_BUCKETS buckets, greedy in bucket order.
__global__ void __launch_bounds__(256)
group_buckets_kernel(const Bucket* __restrict__ bk, int64_t nb, int32_t* __restrict__ gfirst,
                     int32_t* __restrict__ gnb) {
  const int64_t w = blockIdx.x * 256ll + threadIdx.x;
  if (w * PT_GROUP_BUCKETS >= nb) return;
  const int64_t b0 = w * PT_GROUP_BUCKETS;
  const int64_t b1 = b0 + PT_GROUP_BUCKETS < nb ? b0 + PT_GROUP_BUCKETS : nb;
  int32_t* f = gfirst + b0;
  int32_t* c = gnb + b0;
  int slot = 0;
  int64_t start = b0;
  int tot = 0;
  for (int64_t b = b0; b < b1; b++) {
    const int cnt = bk[b].count;
    if (b > start && tot + cnt > PT_GROUP_TARGETS) {
      f[slot] = (int32_t)start;
      c[slot] = (int32_t)(b - start);
      slot++;
      start = b;
      tot = 0;
    }
    tot += cnt;
  }
  f[slot] = (int32_t)start;
  c[slot] = (int32_t)(b1 - start);
  for (slot++; slot < PT_GROUP_BUCKETS; slot++) {
    f[slot] = (int32_t)b0;
    c[slot] = 0;
  }
}

// One wavefront per bucket group.  The walk of pt_walk_step with a member
// mask per stack entry: a popped entry's 8 children inherit the mask of the
// buckets that opened their parent; each lane tests its child against every
// bucket of that mask (GravityVisitor::open, the same FP64 test as the
// per-bucket walk), so each bucket sees exactly its own decisions.
constexpr int kGroupStack = 1024;

template <bool FILL>
__global__ void __launch_bounds__(kTravThreads)
group_traverse_kernel(const TreeNode* __restrict__ nodes, const Bucket* __restrict__ bk,
                      const int32_t* __restrict__ gfirst, const int32_t* __restrict__ gnb,
                      int64_t ns, double gf, double ox, double oy, double oz,
                      int32_t* __restrict__ pp_cnt, int32_t* __restrict__ pc_cnt,
                      const int64_t* __restrict__ pp_off, const int64_t* __restrict__ pc_off,
                      int32_t* __restrict__ pp, uint8_t* __restrict__ ppm,
                      int32_t* __restrict__ pc, uint8_t* __restrict__ pcm,
                      int32_t* __restrict__ err) {
  __shared__ int32_t stack[kTravWaves][kGroupStack];
  __shared__ uint8_t smask[kTravWaves][kGroupStack];
  __shared__ Bucket s_bk[kTravWaves][PT_GROUP_BUCKETS];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t gi = pt_xcd_block(blockIdx.x, gridDim.x) * kTravWaves + wave;
  if (gi >= ns) return;
  const int nbk = gnb[gi];
  if (nbk <= 0) {
    if (!FILL && lane == 0) {
      pp_cnt[gi] = 0;
      pc_cnt[gi] = 0;
    }
    return;
  }
  const int64_t b0 = gfirst[gi];
  if (lane < nbk) s_bk[wave][lane] = bk[b0 + lane];
  __builtin_amdgcn_wave_barrier();
  int32_t* st = stack[wave];
  uint8_t* sm = smask[wave];
  int32_t* ppw = FILL ? pp + pp_off[gi] : nullptr;
  uint8_t* ppmw = FILL ? ppm + pp_off[gi] : nullptr;
  int32_t* pcw = FILL ? pc + pc_off[gi] : nullptr;
  uint8_t* pcmw = FILL ? pcm + pc_off[gi] : nullptr;
  const unsigned all = (1u << nbk) - 1u;
  int sp = 0;
  bool overflow = false;
  int64_t npp = 0, npc = 0;
  {
    const TreeNode root = nodes[0];
    if (root.type == PT_NODE_LEAF) {
      if (FILL && lane == 0) {
        ppw[0] = 0;
        ppmw[0] = (uint8_t)all;
      }
      npp = 1;
    } else if (root.type == PT_NODE_INTERNAL) {
      unsigned om = 0;
      for (int k = 0; k < nbk; k++)
        if (pt_gv_open(root, s_bk[wave][k], gf, ox, oy, oz)) om |= 1u << k;
      const unsigned cm = all & ~om;
      if (cm) {
        if (FILL && lane == 0) {
          pcw[0] = 0;
          pcmw[0] = (uint8_t)cm;
        }
        npc = 1;
      }
      if (om) {
        if (lane == 0) {
          st[0] = root.first_child;
          sm[0] = (uint8_t)om;
        }
        sp = 1;
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  while (sp > 0) {
    const int k = sp < 8 ? sp : 8;
    sp -= k;
    const int slot = lane >> 3;
    int32_t node = -1, fc = -1;
    unsigned vm = 0, om = 0, cm = 0;
    bool leaf = false;
    if (slot < k) {
      node = st[sp + slot] + (lane & 7);
      vm = sm[sp + slot];
      const TreeNode nd = nodes[node];
      fc = nd.first_child;
      if (nd.type == PT_NODE_LEAF) {
        leaf = true;
      } else if (nd.type == PT_NODE_INTERNAL) {
        for (unsigned m = vm; m != 0u; m &= m - 1u) {
          const int b = __builtin_ctz(m);
          if (pt_gv_open(nd, s_bk[wave][b], gf, ox, oy, oz)) om |= 1u << b;
        }
        cm = vm & ~om;
      }
    }
    const unsigned long long leafm = __ballot(leaf);
    const unsigned long long closem = __ballot(cm != 0u);
    const unsigned long long openm = __ballot(om != 0u);
    const int n_open = __popcll(openm);
    if (sp + n_open > kGroupStack) {
      overflow = true;
      break;
    }
    __builtin_amdgcn_wave_barrier();
    if (om != 0u) {
      const int p = sp + pt_rank_in(openm);
      st[p] = fc;
      sm[p] = (uint8_t)om;
    }
    if (FILL) {
      if (leaf) {
        const int64_t p = npp + pt_rank_in(leafm);
        ppw[p] = node;
        ppmw[p] = (uint8_t)vm;
      }
      if (cm != 0u) {
        const int64_t p = npc + pt_rank_in(closem);
        pcw[p] = node;
        pcmw[p] = (uint8_t)cm;
      }
    }
    npp += __popcll(leafm);
    npc += __popcll(closem);
    sp += n_open;
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) {
    if (overflow) atomicExch(err, 1);
    if (!FILL) {
      pp_cnt[gi] = (int32_t)npp;
      pc_cnt[gi] = (int32_t)npc;
    }
  }
}

}  // namespace

int64_t pt_traverse_blocks(int64_t nb) { return (nb + kTravWaves - 1) / kTravWaves; }

int pt_launch_reduce_counters(const unsigned long long* part, int64_t nblocks,
                              unsigned long long* out, const int32_t* err, hipStream_t s,
                              bool accumulate) {
  if (!accumulate && pt_launch_fill_u64(out, 4, 0ull, s) != 0) return -1;
  int64_t grid = (nblocks + 255) / 256;
  if (grid > 256) grid = 256;
  if (grid < 1) grid = 1;
  reduce_counters_kernel<<<(unsigned)grid, 256, 0, s>>>(part, nblocks, out, err);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ho != nullptr selects the HEXADECAPOLE open() (gm = monopole_gravity_factor).
int pt_launch_traverse(bool fill, const TreeNode* nodes, const HexOpen* ho, const Bucket* bk,
                       int64_t nb, double gf, double gm, const double off[3], int32_t* pp_cnt,
                       int32_t* pc_cnt, const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                       int32_t* pc, unsigned long long* partials, int32_t* err, hipStream_t s) {
  if (nb <= 0) return 0;
  const unsigned grid = (unsigned)pt_traverse_blocks(nb);
#define PT_TRAV(F, H)                                                                        \
  traverse_kernel<F, H><<<grid, kTravThreads, 0, s>>>(nodes, ho, bk, nb, gf, gm, off[0],     \
                                                      off[1], off[2], pp_cnt, pc_cnt, pp_off, \
                                                      pc_off, pp, pc, partials, err)
  if (ho) {
    if (fill) PT_TRAV(true, true);
    else PT_TRAV(false, true);
  } else {
    if (fill) PT_TRAV(true, false);
    else PT_TRAV(false, false);
  }
#undef PT_TRAV
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int64_t pt_group_slots(int64_t nb) {
  return (nb + PT_GROUP_BUCKETS - 1) / PT_GROUP_BUCKETS * PT_GROUP_BUCKETS;
}

int pt_launch_group_buckets(const Bucket* bk, int64_t nb, int32_t* gfirst, int32_t* gnb,
                            hipStream_t s) {
  if (nb <= 0) return 0;
  const int64_t nw = (nb + PT_GROUP_BUCKETS - 1) / PT_GROUP_BUCKETS;
  group_buckets_kernel<<<(unsigned)((nw + 255) / 256), 256, 0, s>>>(bk, nb, gfirst, gnb);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_group_traverse(bool fill, const TreeNode* nodes, const Bucket* bk,
                             const int32_t* gfirst, const int32_t* gnb, int64_t ns, double gf,
                             const double off[3], int32_t* pp_cnt, int32_t* pc_cnt,
                             const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                             uint8_t* ppm, int32_t* pc, uint8_t* pcm, int32_t* err,
                             hipStream_t s) {
  if (ns <= 0) return 0;
  const unsigned grid = (unsigned)((ns + kTravWaves - 1) / kTravWaves);
  if (fill)
    group_traverse_kernel<true><<<grid, kTravThreads, 0, s>>>(
        nodes, bk, gfirst, gnb, ns, gf, off[0], off[1], off[2], pp_cnt, pc_cnt, pp_off, pc_off,
        pp, ppm, pc, pcm, err);
  else
    group_traverse_kernel<false><<<grid, kTravThreads, 0, s>>>(
        nodes, bk, gfirst, gnb, ns, gf, off[0], off[1], off[2], pp_cnt, pc_cnt, pp_off, pc_off,
        pp, ppm, pc, pcm, err);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

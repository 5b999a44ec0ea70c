// Host-side launch wrappers of the gfx950 kernels (internal to
// libpt_gravity.so).  Every wrapper enqueues on the given stream and returns 0
// or -1 (launch error).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_common.h"

// Device node table (SoA, level order).
struct PtTree {
  uint64_t* key;
  int32_t* begin;
  int32_t* count;
  int32_t* depth;
  int32_t* type;
  int32_t* parent;
  int32_t* first_child;
};

int64_t pt_scan_scratch_bytes(int64_t n);
int pt_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* scratch,
                          uint32_t* total, hipStream_t s);
int pt_exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, void* scratch,
                                 int64_t* total, hipStream_t s);

int pt_launch_universe(const double* x, const double* y, const double* z, int64_t n,
                       double* scratch_parts, int max_parts, double* bbox_out, double* box,
                       hipStream_t s);
// Morton keys and the 64-byte particle records the sort's gather reads.
int pt_launch_keys(const double* x, const double* y, const double* z, const double* m,
                   const double* e, int64_t n, const double* box, uint64_t* keys,
                   ParticleRec* rec, hipStream_t s);

int64_t pt_sort_scratch_bytes(int64_t n);
// Stable LSD sort of src (read only) into keys, with vals = input index.
int pt_radix_sort(const uint64_t* src, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                  int end_bit, void* scratch, hipStream_t s);
int pt_launch_iota(uint32_t* v, int64_t n, hipStream_t s);
// Message of the calling thread's last failed pt_sort_noalloc, or NULL.
const char* pt_noalloc_last_error();

int pt_launch_root(int64_t n, int bucket, const PtTree& t, int32_t* level_cnt, hipStream_t s);
// Largest leaf the fixup after a top-bits sort handles (larger buckets use
// the full sort).
constexpr int kFixupMax = 64;
// status: the build read-back ({nn, nb, ...}); max_buckets bounds nb.
int pt_launch_leaf_fixup(const PtTree& t, int64_t max_buckets, const int32_t* bucket_node,
                         const int64_t* status, uint64_t* keys, uint32_t* perm, hipStream_t s);
int pt_launch_split_level(const uint64_t* keys, int32_t* level_nodes, int32_t* level_cnt, int d,
                          int64_t max_frontier, int bucket, const PtTree& t, int64_t node_cap,
                          int64_t level_cap, int32_t* overflow, hipStream_t s);
// levels: the split launches done (depths 0 .. levels-1).
int pt_launch_buckets(const PtTree& t, int64_t node_cap, int64_t n, int levels,
                      const int32_t* level_cnt, const int32_t* overflow, uint32_t* head,
                      uint32_t* head_pos, void* scan_scratch, uint32_t* nb_out,
                      int32_t* bucket_node, int64_t* status, hipStream_t s);

// HEXADECAPOLE moments: raw third/fourth moments (SoA, PT_HEXA_RAW x nn) of
// the buckets, then of each internal level (deepest first), then the records.
int pt_launch_hexa_leaf(const PtTree& t, const int32_t* bucket_node, int64_t nb, const double* x,
                        const double* y, const double* z, const double* m, const double* nd,
                        int64_t nn, double* raw, hipStream_t s);
int pt_launch_hexa_level(const PtTree& t, const int32_t* level_nodes, int32_t nl,
                         const double* nd, int64_t nn, double* raw, hipStream_t s);
int pt_launch_hexa_pack(const PtTree& t, const double* nd, const double* raw, int64_t nn,
                        HexNode* hn, HexOpen* ho, hipStream_t s);
// Leaf CentroidData of every leaf and the default record of every empty leaf.
// Where the moment kernels write each node's walk and interaction records as
// soon as its CentroidData is final, and each bucket's record (index
// head_pos[first particle]).
struct PtPackOut {
  TreeNode* tnode;
  GravNode* gn;
  GravNode32* gn32;
  Bucket* bk;
  const uint32_t* head_pos;
};
int pt_launch_leaf_moments(const PtTree& t, int64_t nn, const double* x, const double* y,
                           const double* z, const double* m, const double* e, double* nd,
                           const PtPackOut& o, hipStream_t s);
// Consecutive small levels of the upward pass (deepest first) in one
// work-group: off/cnt of each level within level_nodes.
#define PT_MAX_RUN 24
struct PtLevelRun {
  int32_t nlev;
  int32_t off[PT_MAX_RUN];
  int32_t cnt[PT_MAX_RUN];
};
int pt_launch_upward_run(const PtTree& t, const int32_t* level_nodes, const PtLevelRun& run,
                         double* nd, int64_t nn, int32_t* ord_h, uint64_t* ord_b,
                         const PtPackOut& o, hipStream_t s);
int pt_launch_upward_level(const PtTree& t, const int32_t* level_nodes, int32_t nl, double* nd,
                           int64_t nn, int32_t* ord_h, uint64_t* ord_b, const PtPackOut& o,
                           hipStream_t s);

// List-producing traversal (count pass when fill is false).  partials holds
// pt_traverse_blocks(nb) x 4 counters, reduced by pt_launch_reduce_counters.
int64_t pt_traverse_blocks(int64_t nb);
// ho != nullptr: HEXADECAPOLE open() (gm = monopole_gravity_factor).
int pt_launch_traverse(bool fill, const TreeNode* nodes, const HexOpen* ho, const Bucket* bk,
                       int64_t nb, double gf, double gm, const double off[3], int32_t* pp_cnt,
                       int32_t* pc_cnt, const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                       int32_t* pc, unsigned long long* partials, int32_t* err, hipStream_t s);
// out[0..3] = summed counters (added to out with accumulate), out[4] = *err
// (err may be null).
int pt_launch_reduce_counters(const unsigned long long* part, int64_t nblocks,
                              unsigned long long* out, const int32_t* err, hipStream_t s,
                              bool accumulate = false);

// Interactions over the lists (precision 0 = FP64, 1 = FP32C).
int pt_launch_interact(int precision, const Bucket* bk, int64_t nb, const TreeNode* nodes,
                       const GravNode* gn, const GravNode32* gn32, const double* x, const double* y, const double* z,
                       const double* m, const double* e, const int64_t* pp_off, const int32_t* pp,
                       const int64_t* pc_off, const int32_t* pc, const double off[3], double* ax,
                       double* ay, double* az, double* pot, hipStream_t s);

// Fused walk + interactions (resident grid of pt_gravity_blocks work-groups;
// partials holds 4 counters per work-group; queue holds 8 work counters).
// node() / leaf() of the HEXADECAPOLE build over the lists (FP64).
int pt_launch_interact_hexa(const Bucket* bk, int64_t nb, const TreeNode* nodes,
                            const HexNode* hn, const double* x, const double* y, const double* z,
                            const double* m, const double* e, const int64_t* pp_off,
                            const int32_t* pp, const int64_t* pc_off, const int32_t* pc,
                            const double off[3], double* ax, double* ay, double* az, double* pot,
                            hipStream_t s);
// Fused walk + interactions of the HEXADECAPOLE build (FP64), counters as
// partials of pt_hexa_gravity_blocks(nb) work-groups.
int64_t pt_hexa_gravity_blocks(int64_t nb);
// precision 0: FP64; otherwise the compensated-FP32 form of node() and leaf().
int pt_launch_hexa_gravity(int precision, bool big_stack, const Bucket* bk, int64_t nb,
                           const TreeNode* nodes, const HexOpen* ho, const HexNode* hn,
                           const double* x, const double* y, const double* z, const double* m,
                           const double* e, double gf, double gm, const double off[3],
                           double* ax, double* ay, double* az, double* pot,
                           unsigned long long* partials, int32_t* err, bool accumulate,
                           hipStream_t s);
int64_t pt_gravity_blocks(int64_t nb, int precision);
int pt_launch_gravity(int precision, bool big_stack, const Bucket* bk, int64_t nb,
                      const TreeNode* nodes, const GravNode* gn, const GravNode32* gn32,
                      const double* x, const double* y, const double* z, const double* m,
                      const double* e, double gf, const double off[3], double* ax, double* ay, double* az,
                      double* pot, unsigned long long* partials, int32_t* err,
                      unsigned long long* queue, bool accumulate, hipStream_t s);

// Bucket groups of the grouped fused kernel: runs of consecutive buckets of
// at most cap (2..64) particles in all (a larger bucket alone), formed inside
// windows of 64 buckets and compacted in order into gfirst / gnb, their
// number in *ngroups (device).  tfirst / tnb / gfirst / gnb hold nb entries,
// wcount / woff pt_fgroup_windows(nb); scratch: pt_scan_scratch_bytes of the
// window count.
int64_t pt_fgroup_windows(int64_t nb);
int64_t pt_fgroup_max_groups(int64_t nb, int64_t n, int cap);
int pt_launch_fgroups(const Bucket* bk, int64_t nb, int cap, int32_t* tfirst, int32_t* tnb,
                      uint32_t* wcount, uint32_t* woff, void* scratch, uint32_t* ngroups,
                      int32_t* gfirst, int32_t* gnb, hipStream_t s);
// The fused walk + interactions with the shared walk per bucket group; the
// same calls, counters and partials layout as pt_launch_gravity (partials:
// pt_grouped_gravity_blocks(nb, n, cap) work-groups).  ho / hn non-null: the
// HEXADECAPOLE build (open() with gm, node() from hn), else the quadrupole
// build (gn / gn32).
int64_t pt_grouped_gravity_blocks(int64_t nb, int64_t n, int cap);
int pt_launch_grouped_gravity(int precision, bool big_stack, const Bucket* bk,
                              const int32_t* gfirst, const int32_t* gnb, const uint32_t* ngroups,
                              int64_t nb, int64_t n, int cap, const TreeNode* nodes,
                              const GravNode* gn, const GravNode32* gn32, const HexOpen* ho,
                              const HexNode* hn, const double* x, const double* y,
                              const double* z, const double* m, const double* e, double gf,
                              double gm, const double off[3], double* ax, double* ay, double* az,
                              double* pot, unsigned long long* partials, int32_t* err,
                              bool accumulate, hipStream_t s);

// Ewald lattice correction of the periodic build (quadrupole form) from the
// root's moments (node 0 of nd), added to the sorted results.  table holds
// 4 * pt_ewald_table_entries() doubles, count one int.
int64_t pt_ewald_table_entries();
int pt_launch_ewald(int64_t n, const double* x, const double* y, const double* z,
                    const double* nd, int64_t nn, double L, int n_replicas, double* table,
                    int32_t* count, double* ax, double* ay, double* az, double* pot,
                    hipStream_t s);

// LeafEwaldFn of the HEXADECAPOLE build: the root's moments to fourth order
// (nd for mass and centre, hn[0] for the reduced moments); table holds
// 5 * pt_ewald_table_entries() doubles.
int pt_launch_ewald_hexa(int64_t n, const double* x, const double* y, const double* z,
                         const double* nd, int64_t nn, const HexNode* hn, double L,
                         int n_replicas, double* table, int32_t* count, double* ax, double* ay,
                         double* az, double* pot, hipStream_t s);

int pt_launch_gather(const uint32_t* perm, int64_t n, const ParticleRec* rec, double* xs,
                     double* ys, double* zs, double* ms, double* es, hipStream_t s);
// Sorted slot i -> original slot perm[i] for four arrays; stage: 4 n doubles
// (32-byte aligned).
int pt_launch_unsort(const uint32_t* perm, int64_t n, const double* a0, const double* a1,
                     const double* a2, const double* a3, double* o0, double* o1, double* o2,
                     double* o3, double* stage, hipStream_t s);
// Dual-tree traversal (pt_dual.hip): node -> bucket map, the node ids of two
// keys, one breadth-first step over node pairs, and per-bucket lists from
// records sorted by (bucket << 32 | source).
int pt_launch_dual_bucket_of(const Bucket* bk, int64_t nb, int32_t* bucket_of, hipStream_t s);
int pt_launch_dual_find(const uint64_t* key, int64_t nn, uint64_t k0, uint64_t k1, int32_t* ids,
                        hipStream_t s);
int pt_launch_dual_step(const uint64_t* items, int64_t n_items, const TreeNode* tn,
                        const double* nd, const int32_t* depth, const int32_t* bucket_of,
                        int64_t nn, double gf, const double off[3], uint64_t* next,
                        int64_t next_cap, uint64_t* pp_rec, int64_t pp_cap, uint64_t* pc_rec,
                        int64_t pc_cap, unsigned long long* state, hipStream_t s);
int pt_launch_records_to_lists(const uint64_t* rec, int64_t n, int64_t nb, int32_t* ids,
                               int64_t* off, hipStream_t s);

// Step modes of pt_launch_kick_drift (Particle::kick, Particle::perturb).
enum { PT_STEP_KICK_DRIFT = 0, PT_STEP_KICK = 1, PT_STEP_PERTURB = 2 };
int pt_launch_kick_drift(int64_t n, double dt, int mode, double* x, double* y, double* z,
                         double* vx, double* vy, double* vz, const double* ax, const double* ay,
                         const double* az, hipStream_t s);
int pt_launch_direct(const int64_t* targets, int64_t nt, const double* x, const double* y,
                     const double* z, const double* m, const double* e, int64_t n, double* out,
                     hipStream_t s);
int pt_launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
int pt_launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
// *out = min(v[0..n)) (+inf for n = 0); parts: max_parts doubles of scratch,
// out outside them.
int pt_launch_min(const double* v, int64_t n, double* parts, int max_parts, double* out,
                  hipStream_t s);

// Bucket groups of the adjoint: runs of consecutive buckets (their particles
// are consecutive) of at most PT_GROUP_BUCKETS buckets and, unless a single
// bucket is larger, PT_GROUP_TARGETS particles, formed greedily inside aligned
// windows of PT_GROUP_BUCKETS buckets.  Window w owns the slots
// [w * PT_GROUP_BUCKETS, (w + 1) * PT_GROUP_BUCKETS); unused slots hold
// gnb = 0.  pt_group_slots(nb) slots in all.
#define PT_GROUP_BUCKETS 8
#define PT_GROUP_TARGETS 16
int64_t pt_group_slots(int64_t nb);
int pt_launch_group_buckets(const Bucket* bk, int64_t nb, int32_t* gfirst, int32_t* gnb,
                            hipStream_t s);
// The walk of pt_launch_traverse for every bucket of a group at once: a node
// is visited when some member bucket visits it, and each list entry carries
// the member mask (bit k = the group's k-th bucket) for which the node is a
// node() source (pc/pcm) or a leaf() source (pp/ppm).  For each member bucket
// the entries with its bit set are exactly its pt_launch_traverse list.
// Count pass (fill false: pp_cnt/pc_cnt per slot) and fill pass; err is set
// when a walk stack overflows.
int pt_launch_group_traverse(bool fill, const TreeNode* nodes, const Bucket* bk,
                             const int32_t* gfirst, const int32_t* gnb, int64_t ns, double gf,
                             const double off[3], int32_t* pp_cnt, int32_t* pc_cnt,
                             const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                             uint8_t* ppm, int32_t* pc, uint8_t* pcm, int32_t* err,
                             hipStream_t s);

// Adjoint of the interactions over the group lists (precision 0: FP64;
// otherwise FP32 pair arithmetic on group-centre-shifted coordinates with
// FP64 sums).  lam_node holds pt_adjoint_node_scratch(nn) doubles, all zero
// on entry (node cotangents field-major, then private copies of the hot
// ones); g* accumulate.
int pt_node_adjoint_fields();
int64_t pt_adjoint_node_scratch(int64_t nn);
int pt_launch_adjoint_groups(int precision, const Bucket* bk, const int32_t* gfirst,
                             const int32_t* gnb, int64_t ns, const TreeNode* nodes,
                             const GravNode* gn, const GravNode32* gn32, const double* x,
                             const double* y, const double* z, const double* m, const double* e,
                             const int64_t* pp_off, const int32_t* pp, const uint8_t* ppm,
                             const int64_t* pc_off, const int32_t* pc, const uint8_t* pcm,
                             const double off[3], const double* wx, const double* wy,
                             const double* wz, double* gx, double* gy, double* gz, double* gm,
                             double* lam_node, int64_t nn, hipStream_t s);
int pt_launch_adjoint_down(int64_t n, const uint32_t* head, const uint32_t* head_pos,
                           const int32_t* bucket_node, const int32_t* parent, const GravNode* gn,
                           const double* lam_node, int64_t nn, const double* x, const double* y,
                           const double* z, const double* m, const double* e, double* gx,
                           double* gy, double* gz, double* gm, hipStream_t s);
int pt_launch_direct_all(int64_t n, const double* x, const double* y, const double* z,
                         const double* m, const double* e, double* ax, double* ay, double* az,
                         double* pot, hipStream_t s);
int pt_launch_direct_adjoint(int64_t n, const double* x, const double* y, const double* z,
                             const double* m, const double* e, const double* wx,
                             const double* wy, const double* wz, double* gx, double* gy,
                             double* gz, double* gm, hipStream_t s);
int pt_launch_adjoint_w(int64_t n, double dt, const double* lx, const double* ly,
                        const double* lz, double* lvx, double* lvy, double* lvz, double* wx,
                        double* wy, double* wz, hipStream_t s);
int pt_launch_gather3(const uint32_t* perm, int64_t n, const double* a, const double* b,
                      const double* c, double* as, double* bs, double* cs, hipStream_t s);
int pt_launch_scatter_add4(const uint32_t* perm, int64_t n, const double* a, const double* b,
                           const double* c, const double* d, double* oa, double* ob, double* oc,
                           double* od, hipStream_t s);
int pt_launch_split3(int64_t n, const double* v, double* a, double* b, double* c, hipStream_t s);
int pt_launch_join3(int64_t n, const double* a, const double* b, const double* c, double* v,
                    hipStream_t s);

/*
 * pt_gravity.h -- C ABI of the MI355X (gfx950) gravity path:
 * universe -> Morton keys -> radix sort -> octree -> CentroidData upward pass
 * -> TransposedDownTraverser<GravityVisitor> lists -> P-C / P-P interactions
 * -> kick/drift.
 *
 * The reference (paratreet/paratreet) has no FFI: its path is driven by
 * Charm++ entry methods and C++ templates.  Each entry point below names the
 * reference function or entry method it stands in for, so a host that used
 * the reference's Driver loop (src/Driver.h:183-288) can call these in the
 * same order.  All pointers are plain host or device pointers; no torch or
 * HIP types appear in the signatures (the stream is passed as void*).
 *
 * Error behaviour: every call returns a pt_status (0 = PT_OK, < 0 = error)
 * and records a message retrievable with pt_last_error(ctx).  Where the
 * reference aborts (CkAbort / CkAssert / CmiAbort) the library returns an
 * error instead; it never aborts the process.
 *
 * Threading: one context per host thread; a context owns one HIP stream and
 * all its device buffers.  Calls are synchronous with respect to the host
 * unless documented otherwise.
 */
#ifndef PT_GRAVITY_H_
#define PT_GRAVITY_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_ctx pt_ctx;

typedef enum {
  PT_OK = 0,
  PT_ERR_ARG = -1,         /* invalid argument */
  PT_ERR_HIP = -2,         /* HIP runtime / launch failure */
  PT_ERR_UNSUPPORTED = -3, /* configuration outside this path (e.g. kd tree) */
  PT_ERR_CAPACITY = -4,    /* a bounded device structure overflowed */
  PT_ERR_STATE = -5        /* stage called before its inputs exist */
} pt_status;

/* paratreet::TreeType and paratreet::DecompType (src/Configuration.h:14-29). */
typedef enum { PT_TREE_OCT = 0, PT_TREE_BINARY_OCT = 1, PT_TREE_KD = 2, PT_TREE_LONGEST = 3 } pt_tree_type;
typedef enum { PT_DECOMP_OCT = 0, PT_DECOMP_BINARY_OCT = 1, PT_DECOMP_SFC = 2, PT_DECOMP_KD = 3,
               PT_DECOMP_LONGEST = 4 } pt_decomp_type;

/* Arithmetic form of the interaction kernels. */
typedef enum { PT_FP64 = 0, PT_FP32C = 1 } pt_precision;

/* Force model of a trajectory: the tree (GravityVisitor over the octree) or
 * the FP64 direct sum of the leaf() kernel over all pairs (validation). */
typedef enum { PT_FORCE_TREE = 0, PT_FORCE_DIRECT = 1 } pt_force_mode;

/* Where caller-supplied arrays live. */
typedef enum { PT_HOST = 0, PT_DEVICE = 1 } pt_memory;

/* The fields of paratreet::Configuration (src/Configuration.h:84-125) that
 * the gravity path reads.  CLI letters as registered in register_fields
 * (src/Configuration.h:132-153). */
typedef struct pt_config {
  int32_t min_n_subtrees;         /* '-n'; one GPU holds one subtree (tp_key 1) */
  int32_t min_n_partitions;       /* '-p'; one partition */
  int32_t max_particles_per_leaf; /* '-l'; bucket size B */
  int32_t decomp_type;            /* '-d'; pt_decomp_type, only PT_DECOMP_OCT */
  int32_t tree_type;              /* '-t'; pt_tree_type, only PT_TREE_OCT */
  int32_t num_iterations;         /* '-i'; used by pt_run */
  double dsoft;                   /* '-e' dSoft; applied to every particle if has_dsoft */
  int32_t has_dsoft;              /* Configuration::origin_of("dSoft") != Unknown */
  int32_t periodic;               /* bPeriodic: pt_compute_forces / pt_run sum the
                                     (2 n_replicas + 1)^3 replicas of the period box
                                     (examples/Gravity.C:35-46) */
  int32_t hexadecapole;           /* the -DHEXADECAPOLE build (examples/Makefile:5):
                                     moments to fourth order, GravityVisitor::open/node
                                     HEXADECAPOLE branches (examples/GravityVisitor.h:
                                     162-178,199-206); pt_compute_forces in FP64 or
                                     PT_FP32C, the list path FP64 only, no adjoint.
                                     0 = the quadrupole build */
  double period[3];               /* dxPeriod dyPeriod dzPeriod (used when periodic) */
  int32_t n_replicas;             /* nReplicas, 0..16 (used when periodic) */
  int32_t dual_tree;              /* the application's dual_tree flag ('-e' in
                                     examples/Main.C:69-71): pt_compute_forces runs
                                     Subtree::startDual's traversal (pt_start_dual from
                                     the root pair) before the down traversal, both adding
                                     to the accelerations and counters
                                     (examples/Gravity.C:30-33); not with periodic or
                                     hexadecapole */
} pt_config;

/* Interaction counters, src/ThreadStateHolder.h:10-13 (COUNT_INTERACTIONS). */
typedef struct pt_counters {
  uint64_t n_node_ints;  /* += target bucket size per node() call */
  uint64_t n_part_ints;  /* += n_source * n_target per leaf() call */
  uint64_t n_opens;      /* open() returned true */
  uint64_t n_closes;     /* open() returned false */
} pt_counters;

/* Stage timings in milliseconds (HIP events on the context stream). */
typedef enum {
  PT_T_UNIVERSE = 0, PT_T_KEYS, PT_T_SORT, PT_T_BUILD, PT_T_MOMENTS, PT_T_TRAVERSE, PT_T_FORCE,
  PT_T_INTEGRATE, PT_T_ADJOINT, PT_T_COUNT
} pt_timer;

/* Device pointers owned by the context (valid until the next call that
 * resizes it).  Particle arrays in the caller's original order. */
typedef struct pt_device_view {
  int64_t n;
  double *x, *y, *z, *vx, *vy, *vz, *mass, *soft;
  double *ax, *ay, *az, *pot;
  uint64_t* keys;        /* original order */
  uint64_t* sorted_keys; /* sorted order */
  uint32_t* perm;        /* sorted slot -> original index */
  void* stream;          /* hipStream_t */
} pt_device_view;

/* Reference defaults (examples/Main.C:26-41) restricted to this path:
 * max_particles_per_leaf = 12, octree, one subtree/partition. */
void pt_config_default(pt_config* cfg);

/* Creates a context on `device`; `stream` may be NULL (the context creates its
 * own default-priority stream) or an existing hipStream_t. */
pt_status pt_ctx_create(const pt_config* cfg, int device, void* stream, pt_ctx** out);
void pt_ctx_destroy(pt_ctx* ctx);
const char* pt_last_error(const pt_ctx* ctx);
pt_status pt_get_device_view(pt_ctx* ctx, pt_device_view* out);

/* Reader::load (src/Reader.C:11-108) + Reader::setSoft (src/Reader.C:110-116):
 * takes the particle set (SoA, original order).  vx/vy/vz may be NULL (zero
 * velocities). */
pt_status pt_set_particles(pt_ctx* ctx, int64_t n, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy,
                           const double* vz, const double* mass, const double* soft,
                           pt_memory where);

/* Driver::remakeUniverse (src/Driver.h:67-80) over the BoundingBox reduction
 * (src/BoundingBox.C:29-44,74-85).  box_out (may be NULL) = {lesser, greater}. */
pt_status pt_remake_universe(pt_ctx* ctx, double box_out[6]);

/* Decomposition::assignKeys (src/Decomposition.C:58-64) /
 * Particle::adjustNewUniverse (src/Particle.C:23-35): Morton key | 1<<63. */
pt_status pt_assign_keys(pt_ctx* ctx);

/* Subtree::buildTree's std::sort (src/Subtree.h:281), Reader::localSort
 * (src/Reader.C:237-241): ascending key, ties by original index. */
pt_status pt_sort(pt_ctx* ctx);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The sort of pt_sort over caller-owned device arrays:
 * keys (n, read only) -> sorted_keys (n, ascending, ties by input index) and
 * perm (n, sorted slot -> input index), enqueued on `stream` (a hipStream_t
 * of the current device, or NULL) and asynchronous.  Its temporaries live in
 * a fixed per-GPU device array held until the process ends
 * (26 MiB, n <= PT_SORT_NOALLOC_MAX_N; larger n returns PT_ERR_ARG); it calls
 * no allocator.  Calls that overlap on one GPU share that array and race.
 * Errors: pt_last_error(NULL) returns the calling thread's last message. */
#define PT_SORT_NOALLOC_MAX_N (1 << 21)
pt_status pt_sort_noalloc(const uint64_t* keys, int64_t n, uint64_t* sorted_keys, uint32_t* perm,
                          void* stream);

/* Subtree::buildTree / recursiveBuild (src/Subtree.h:275-349). */
pt_status pt_build_tree(pt_ctx* ctx);

/* CentroidData leaf constructor + Subtree::populateTree (src/Subtree.h:351-384,
 * examples/CentroidData.h:41-73). */
pt_status pt_upward(pt_ctx* ctx);

/* Partition::startDown<GravityVisitor> (src/Partition.h:129-134) traversal
 * part: builds per-bucket interaction lists for GravityVisitor(offset, theta)
 * (examples/GravityVisitor.h:16-20) and the counters of
 * src/Traverser.h:15-37. */
pt_status pt_traverse(pt_ctx* ctx, double theta, const double offset[3], pt_counters* counters);

/* The visitor calls of the same traversal: GravityVisitor::node/leaf
 * (examples/GravityVisitor.h:144-224) over the lists, writing acceleration and
 * potential per particle (the reference accumulates onto zeroed particles). */
pt_status pt_interact(pt_ctx* ctx, pt_precision precision);

/* Subtree::startDual<GravityVisitor> (src/Subtree.h:233-240) ->
 * DualTraverser::doTrav / runInvertedTraversal (src/Traverser.h:485-610) for
 * GravityVisitor(offset, theta) with its cell() (examples/GravityVisitor.h:
 * 226-230), from the node pair with keys (source_key, target_key): (1, 1) is
 * startDual's pair (the root against the root); another pair is a doTrav
 * resumed at that pair (DualTraverser::resumeTrav, :531-537).  Leaves the
 * leaf() and node() calls that move particles as the context's lists
 * (pt_get_lists), evaluates them in `precision` (accelerations and
 * potentials of this traversal alone, pt_get_accelerations) and returns the
 * counters, which add n_particles as the reference's nodes hold it (-1 for an
 * internal node) modulo 2^64.  Quadrupole, non-periodic contexts only (the
 * reference aborts on dual_tree with periodic, examples/Gravity.C:30). */
pt_status pt_start_dual(pt_ctx* ctx, double theta, const double offset[3], uint64_t source_key,
                        uint64_t target_key, pt_precision precision, pt_counters* counters);

/* One force evaluation: universe, keys, sort, build, upward, traverse,
 * interact -- paratreet::traversalFn with everything before it in one
 * Driver::run iteration (src/Driver.h:190-221, examples/Gravity.C:29-33). */
pt_status pt_compute_forces(pt_ctx* ctx, double theta, const double offset[3],
                            pt_precision precision, pt_counters* counters);

/* Partition::kick + Partition::perturb (src/Partition.h:329-379 ->
 * src/Particle.C:7-21) with the last computed accelerations. */
pt_status pt_kick_drift(pt_ctx* ctx, double dt);

/* The two halves of pt_kick_drift, as Driver::run calls them with the
 * iteration's output between them (src/Driver.h:222-247):
 * Partition::kick -> Particle::kick (src/Partition.h:329-336,
 * src/Particle.C:7-9), v += a dt / 2; and Partition::perturb ->
 * Particle::perturb (src/Partition.h:338-379, src/Particle.C:11-21),
 * v += a dt / 2 then x += v dt.  pt_kick then pt_perturb gives the same bits
 * as pt_kick_drift. */
pt_status pt_kick(pt_ctx* ctx, double dt);
pt_status pt_perturb(pt_ctx* ctx, double dt);

/* Driver::run (src/Driver.h:183-288) restricted to this path: `steps`
 * iterations of compute_forces + kick_drift with a fixed dt. */
pt_status pt_run(pt_ctx* ctx, int steps, double dt, double theta, pt_precision precision,
                 pt_counters* counters);

/* FP64 direct sum with the leaf() kernel for the given original-order target
 * indices, over all particles (error reference).  acc_out: nt x 3 host. */
pt_status pt_direct_sum(pt_ctx* ctx, const int64_t* targets, int64_t nt, double* acc_out);

/* ---- differentiation (no counterpart in the reference) ----
 * The reference has no adjoint; these entry points add reverse-mode
 * gradients of the Driver::run loop above.  Arrays are n x 3 row-major
 * (xyz interleaved) in original particle order, host or device per `where`.
 *
 * Vector-Jacobian product of one force evaluation at the current positions:
 * for w (n x 3) it returns grad_x = (da/dx)^T w (n x 3) and
 * grad_m = (da/dm)^T w (n), over the interactions pt_traverse lists (tree:
 * walked once per group of neighbouring buckets, each entry marked with the
 * buckets it serves) or over all pairs (direct); accelerations are those of
 * GravityVisitor::node/leaf (examples/GravityVisitor.h:144-224), opening
 * decisions held fixed. */
pt_status pt_force_vjp(pt_ctx* ctx, double theta, pt_force_mode mode, const double* w,
                       double* grad_x, double* grad_m, pt_memory where);

/* Arithmetic of the tree-mode adjoint in pt_force_vjp and pt_leapfrog_adjoint
 * (default PT_FP64).  PT_FP32C evaluates the pair expressions in FP32 on
 * coordinates shifted to the centre of each bucket group's box, with every
 * sum over pairs carried in FP64 (the form of the compensated forward).  The direct-sum
 * adjoint is always FP64. */
pt_status pt_set_adjoint_precision(pt_ctx* ctx, pt_precision precision);

/* pt_run (Driver::run, src/Driver.h:183-288, with Partition::kick/perturb,
 * src/Particle.C:7-21) from the current state, keeping a checkpoint of
 * positions and velocities every `checkpoint_every` steps (<= 0: one
 * checkpoint at step 0) for pt_leapfrog_adjoint.  The final state is left in
 * the context (pt_get_state). */
pt_status pt_leapfrog_forward(pt_ctx* ctx, int32_t steps, double dt, double theta,
                              pt_precision precision, pt_force_mode mode,
                              int32_t checkpoint_every, pt_counters* counters);

/* Reverse sweep of the last pt_leapfrog_forward: given the cotangents of the
 * final positions and velocities (lam_x, lam_v: n x 3, either may be NULL for
 * zero) returns the gradients with respect to the initial positions,
 * velocities (n x 3) and masses (n).  Steps between checkpoints are
 * recomputed with the forward's force model and precision; the adjoint
 * arithmetic is the context's adjoint precision (pt_set_adjoint_precision).
 * The final state is restored afterwards. */
pt_status pt_leapfrog_adjoint(pt_ctx* ctx, const double* lam_x, const double* lam_v,
                              double* grad_x, double* grad_v, double* grad_m, pt_memory where);

/* ---- read-back (host buffers, synchronous) ---- */

/* HEXADECAPOLE contexts: per node (level order, as pt_get_nodes) the mass and
 * the 21 reduced moments in FMOMR order (examples/moments.h: xx yy xy xz yz,
 * xxx xyy xxy yyy xxz yyz xyz, xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz),
 * unscaled; out: nn x 22. */
pt_status pt_get_node_hexadecapole(pt_ctx* ctx, double* out);
pt_status pt_get_universe(pt_ctx* ctx, double box[6]);
pt_status pt_get_keys(pt_ctx* ctx, uint64_t* keys);                 /* original order */
pt_status pt_get_sorted(pt_ctx* ctx, uint64_t* sorted_keys, uint32_t* perm);
pt_status pt_get_tree_size(pt_ctx* ctx, int64_t* n_nodes, int64_t* n_buckets, int32_t* max_depth);
/* Node table, level order.  Any output may be NULL. */
pt_status pt_get_nodes(pt_ctx* ctx, uint64_t* key, int32_t* type, int32_t* depth, int32_t* begin,
                       int32_t* count, int32_t* parent, int32_t* first_child);
/* CentroidData per node: 18 doubles per node in the order radius soft mass
 * cm[3] q[6](xx xy xz yy yz zz) lesser[3] greater[3], node-major. */
pt_status pt_get_node_data(pt_ctx* ctx, double* out);
pt_status pt_get_buckets(pt_ctx* ctx, int32_t* bucket_nodes);       /* DFS order */
pt_status pt_get_lists_size(pt_ctx* ctx, int64_t* n_pp, int64_t* n_pc);
/* Lists per bucket in canonical DFS order (ascending left-aligned key). */
pt_status pt_get_lists(pt_ctx* ctx, int64_t* pp_off, int32_t* pp, int64_t* pc_off, int32_t* pc);
/* The bucket groups and member-masked lists the tree adjoint walks (no
 * counterpart in the reference; read back for checking), built for the
 * current tree at opening angle theta.  sizes = {group slots, pp entries,
 * pc entries}; with gfirst non-NULL every array is also filled (offsets:
 * slots + 1 entries).  Slot s holds buckets [gfirst[s], gfirst[s] + gnb[s])
 * (gnb 0: unused); bit k of an entry's mask marks the slot's k-th bucket,
 * entries in walk order. */
pt_status pt_get_group_lists(pt_ctx* ctx, double theta, int64_t sizes[3], int32_t* gfirst,
                             int32_t* gnb, int64_t* pp_off, int32_t* pp, uint8_t* ppm,
                             int64_t* pc_off, int32_t* pc, uint8_t* pcm);
/* Accelerations / potentials in original particle order (n x 3 and n). */
pt_status pt_get_accelerations(pt_ctx* ctx, double* acc, double* pot);
pt_status pt_get_state(pt_ctx* ctx, double* pos, double* vel);       /* n x 3 each */
pt_status pt_get_counters(pt_ctx* ctx, pt_counters* out);
/* How pt_compute_forces walks the quadrupole build (no reference counterpart;
 * every mode gives every bucket exactly the open()/node()/leaf() calls of
 * TransposedDownTraverser, src/Traverser.h:154-217).  Grouped: runs of
 * consecutive buckets of at most `cap` particles walk once with their union
 * box and share the sources every member closes or reaches as leaves, then
 * sub-runs of at most 16 particles the same way, each bucket walking on its
 * own only below the nodes its run decided differently.  mode 0: one walk
 * per bucket; 1 (the default): automatic (the cap from the particle count
 * and the device's compute units, one walk per bucket for small sets);
 * 2..64: grouped with that cap. */
pt_status pt_set_fused_grouping(pt_ctx* ctx, int mode);
pt_status pt_set_timing(pt_ctx* ctx, int enabled);
pt_status pt_get_timings(pt_ctx* ctx, float ms[PT_T_COUNT]);

#ifdef __cplusplus
}
#endif

#endif /* PT_GRAVITY_H_ */

/*
 * pt_oracle.h -- CPU restatement of ParaTreeT's gravity hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (paratreet_amd, libpt_gravity.so) never links,
 * loads or falls back to it.
 *
 * Every function restates one piece of the reference (paratreet/paratreet) in
 * plain C with Real = double (the reference built with -DUSE_DOUBLE_FP, Morton
 * keys), without -DHEXADECAPOLE except the or_hexa_* and *_hexa functions.
 * Arithmetic follows the reference's operation order so results are
 * reproducible bit for bit when compiled with
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Pieces whose source is NOT in the reference tree (the unvendored
 * N-BodyShop/utility submodule: SFC::generateKey, Space::intersect,
 * OrientedBox) follow the explicit definitions written next to each function;
 * how each is pinned is stated in the header comment of pt_oracle.c.
 */
#ifndef PT_ORACLE_H_
#define PT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node types, numbered as in src/Node.h:82-95 restricted to the local tree. */
enum { OR_EMPTY = 0, OR_LEAF = 1, OR_INTERNAL = 2 };

/* One octree node (DFS pre-order table).  Mirrors Node<Data> fields
 * key/type/depth/parent and the particle range implied by recursiveBuild
 * (src/Subtree.h:313-349). child[] = -1 for leaves. */
typedef struct {
  uint64_t key;
  int32_t type;
  int32_t depth;
  int64_t begin;
  int64_t count;
  int32_t parent;
  int32_t child[8];
} or_node;

/* CentroidData (examples/CentroidData.h:24-91) with MultipoleMoments
 * (examples/MultipoleMoments.h:161-201, non-HEXADECAPOLE) flattened.
 * q = {xx, xy, xz, yy, yz, zz}.  Box lo/hi = OrientedBox lesser/greater. */
typedef struct {
  double radius, soft, mass;
  double cm[3];
  double q[6];
  double lo[3], hi[3];
  int64_t count;
} or_cdata;

/* Driver::remakeUniverse (src/Driver.h:67-80) over the particle bbox. */
void or_universe(const double* x, const double* y, const double* z, int64_t n,
                 double box_out[6]);

/* Decomposition::assignKeys (src/Decomposition.C:58-64): Morton key + bit 63. */
void or_keys(const double* x, const double* y, const double* z, int64_t n,
             const double box[6], uint64_t* keys);

/* Subtree::buildTree's sort (src/Subtree.h:281): ascending key, ties by
 * original index.  perm[i] = original index of sorted slot i. */
void or_sort(const uint64_t* keys, int64_t n, uint64_t* sorted_keys, int32_t* perm);

/* Subtree::buildTree / recursiveBuild (src/Subtree.h:275-349).  Writes nodes
 * in DFS pre-order; returns the node count, or -(needed) if cap is too small. */
int64_t or_build(const uint64_t* sorted_keys, int64_t n, int bucket,
                 or_node* nodes, int64_t cap);

/* Leaf CentroidData + Subtree::populateTree queue-order upward pass
 * (examples/CentroidData.h:41-73, examples/MultipoleMoments.h:204-353,458-482,
 * src/Subtree.h:351-384).  Inputs are in sorted order. */
void or_moments(const or_node* nodes, int64_t nn,
                const double* x, const double* y, const double* z,
                const double* m, const double* eps, or_cdata* out);

/* Number of buckets (non-empty leaves) and their node ids in DFS order. */
int64_t or_buckets(const or_node* nodes, int64_t nn, int32_t* bucket_nodes);

/* TransposedDownTraverser<GravityVisitor> (src/Traverser.h:154-217) per
 * bucket, with GravityVisitor::open/node/leaf (examples/GravityVisitor.h).
 * bucket_sel: NULL = all buckets, else the n_sel bucket ordinals to evaluate.
 * acc/pot are overwritten (zeroed) for the selected buckets' particles and
 * accumulate in DFS interaction order.  counters = {n_node_ints, n_part_ints,
 * n_opens, n_closes} (src/ThreadStateHolder.h:10-13), summed over the
 * selected buckets.  nthreads <= 0 uses the OpenMP default. */
void or_gravity(const or_node* nodes, int64_t nn, const or_cdata* d,
                const double* x, const double* y, const double* z,
                const double* m, const double* eps,
                double theta, const double offset[3],
                const int32_t* bucket_sel, int64_t n_sel,
                double* ax, double* ay, double* az, double* pot,
                uint64_t counters[4], int nthreads);

/* or_gravity over all buckets that also counts the P-C target pairs reaching
 * SPLINEQ's softened branch (examples/GravityVisitor.h:40-73: r^2 < twoh^2
 * after openSoftening was false): soft_pairs = {u < 1, u >= 1}. */
void or_gravity_soft_pairs(const or_node* nodes, int64_t nn, const or_cdata* d,
                           const double* x, const double* y, const double* z,
                           const double* m, const double* eps, double theta,
                           const double offset[3], double* ax, double* ay, double* az,
                           double* pot, uint64_t counters[4], uint64_t soft_pairs[2]);

/* The counters of or_gravity over all buckets (quadrupole build) from the
 * walk alone, without evaluating the forces. */
void or_traversal_counters(const or_node* nodes, int64_t nn, const or_cdata* d, double theta,
                           const double offset[3], uint64_t counters[4], int nthreads);

/* Interaction lists in DFS order, per bucket.  First call with pp/pc NULL
 * fills pp_count/pc_count (per bucket); second call with offsets (exclusive
 * prefix sums) fills node ids.  kind: leaf() sources go to pp, node() sources
 * to pc. */
void or_lists(const or_node* nodes, int64_t nn, const or_cdata* d,
              double theta, const double offset[3],
              int64_t* pp_count, int64_t* pc_count,
              const int64_t* pp_off, const int64_t* pc_off,
              int32_t* pp, int32_t* pc);

/* Direct O(N*M) sum with the leaf() kernel and SPLINE softening, 2h = eps_i +
 * eps_j, skipping r^2 == 0 (examples/GravityVisitor.h:87-120,144-160). */
void or_direct(const double* x, const double* y, const double* z,
               const double* m, const double* eps, int64_t n,
               const int64_t* targets, int64_t nt,
               double* ax, double* ay, double* az, int nthreads);

/* GravityVisitor::SPLINE / SPLINEQ exposed for closed-form tests. */
void or_spline(double r2, double twoh, double* a, double* b);
void or_splineq(double invr, double r2, double twoh,
                double* a, double* b, double* c, double* d);

/* Space::intersect(box, point, rsq) as defined in pt_oracle.c. */
int or_intersect_box_point(const double lo[3], const double hi[3],
                           const double p[3], double rsq);

/* ---- HEXADECAPOLE build (-DHEXADECAPOLE, examples/Makefile:5) ----
 * Reduced (trace-free) moments to fourth order per node, 22 values in FMOMR
 * order (examples/moments.h): m, xx yy xy xz yz, xxx xyy xxy yyy xxz yyz xyz,
 * xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz; unscaled.  raw (optional,
 * nn x 25): raw third and fourth moments about each node's cm. */
void or_hexa_moments(const or_node* nodes, int64_t nn, const or_cdata* d, const double* x,
                     const double* y, const double* z, const double* m, double* raw,
                     double* out);
/* raw third and fourth moments of n particles about cm (25 values) and their
 * reduction with (mass, q) to the 22 FMOMR-order values */
void or_hexa_raw_moments(const double* x, const double* y, const double* z, const double* m,
                         int64_t n, const double cm[3], double out[25]);
void or_hexa_reduce(double mass, const double q[6], const double r[25], double out[22]);
/* potential and acceleration of reduced moments h at r = target - cm */
void or_hexa_eval(const double h[22], double rx, double ry, double rz, double* pot,
                  double acc[3]);
/* or_gravity / or_lists of the HEXADECAPOLE build (open() with the softening
 * and monopole tests, node() with the reduced moments) */
void or_gravity_hexa(const or_node* nodes, int64_t nn, const or_cdata* d, const double* hexa,
                     const double* x, const double* y, const double* z, const double* m,
                     const double* eps, double theta, const double offset[3],
                     const int32_t* bucket_sel, int64_t n_sel, double* ax, double* ay,
                     double* az, double* pot, uint64_t counters[4], int nthreads);
void or_lists_hexa(const or_node* nodes, int64_t nn, const or_cdata* d, const double* hexa,
                   double theta, const double offset[3], int64_t* pp_count, int64_t* pc_count,
                   const int64_t* pp_off, const int64_t* pc_off, int32_t* pp, int32_t* pc);

/* Test knob: scales the real-space erf / erfc and exp values of both Ewald
 * forms by 1 + ulps * 2^-52 (0, 0 = unperturbed, the default). */
void or_set_ewald_libm_perturbation(double erf_ulps, double exp_ulps);

/* LeafEwaldFn + EwaldData::EwaldInit, quadrupole form (examples/Ewald.h,
 * examples/Ewald.C): the lattice of the root's monopole + quadrupole beyond
 * the explicit replicas, added to ax/ay/az/pot.  or_ewald uses the
 * reference's constants (alpha = 2/L, real cut 2.6 L, Fourier cut 2.8);
 * or_ewald_general takes them as parameters. */
void or_ewald(const double* x, const double* y, const double* z, int64_t n, double mass,
              const double cm[3], const double q[6], double L, int n_replicas, double* ax,
              double* ay, double* az, double* pot);
void or_ewald_general(const double* x, const double* y, const double* z, int64_t n,
                      double mass, const double cm[3], const double q[6], double L,
                      int n_replicas, double a_fac, double ew_cut, double h_cut, double* ax,
                      double* ay, double* az, double* pot);

/* DualTraverser<CentroidData, GravityVisitor> (src/Traverser.h:485-610):
 * doTrav from the pair (source, target) of node ids -- Subtree::startDual
 * (src/Subtree.h:233-240) is the pair (root, root) -- with
 * runInvertedTraversal, GravityVisitor::open/node/leaf/cell and its flags
 * TargetMustBeLeaf and ForceEvenDepth (examples/GravityVisitor.h:11-13,
 * 226-230).  leaf() and node() act only where the reference's n_particles
 * make them act (two leaves with particles; a node onto a leaf), while the
 * counters add n_particles as the reference holds it (-1 for an internal
 * node), modulo 2^64 as its unsigned long long counters do.  or_dual zeroes
 * ax/ay/az/pot (n particles, sorted order) and applies the visitor calls in
 * the reference's stack order; or_dual_lists returns the same calls as
 * per-bucket lists (two calls, as or_lists). */
void or_dual(const or_node* nodes, int64_t nn, const or_cdata* d, const double* x,
             const double* y, const double* z, const double* m, const double* eps, double theta,
             const double offset[3], int32_t source, int32_t target, double* ax, double* ay,
             double* az, double* pot, int64_t n, uint64_t counters[4]);
void or_dual_lists(const or_node* nodes, int64_t nn, const or_cdata* d, double theta,
                   const double offset[3], int32_t source, int32_t target, int64_t* pp_count,
                   int64_t* pc_count, const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                   int32_t* pc);

/* LeafEwaldFn + EwaldData::EwaldInit, HEXADECAPOLE form (examples/Ewald.h,
 * examples/Ewald.C): the lattice of the root's moments to fourth order (mass,
 * cm, 21 reduced moments in FMOMR order) beyond the explicit replicas, added
 * to ax/ay/az/pot.  Constants as or_ewald / or_ewald_general. */
void or_ewald_hexa(const double* x, const double* y, const double* z, int64_t n, double mass,
                   const double cm[3], const double h[21], double L, int n_replicas, double* ax,
                   double* ay, double* az, double* pot);
void or_ewald_hexa_general(const double* x, const double* y, const double* z, int64_t n,
                           double mass, const double cm[3], const double h[21], double L,
                           int n_replicas, double a_fac, double ew_cut, double h_cut, double* ax,
                           double* ay, double* az, double* pot);

/* Particle::kick then Particle::perturb (src/Particle.C:7-21). */
/* (da/dx)^T w and (da/dm)^T w of or_direct over all targets (test-only
 * reference for the adjoint). */
void or_direct_vjp(const double* x, const double* y, const double* z,
                   const double* m, const double* eps, int64_t n,
                   const double* wx, const double* wy, const double* wz,
                   double* gx, double* gy, double* gz, double* gm);

void or_kick_drift(double* x, double* y, double* z,
                   double* vx, double* vy, double* vz,
                   const double* ax, const double* ay, const double* az,
                   int64_t n, double dt);

#ifdef __cplusplus
}
#endif

#endif /* PT_ORACLE_H_ */

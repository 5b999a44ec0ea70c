/*
 * pt_oracle.c -- CPU restatement of ParaTreeT's gravity hot path (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product (libpt_gravity.so, paratreet_amd) links or
 * calls this file.
 *
 * Reference configuration restated: Real = double (-DUSE_DOUBLE_FP), Morton
 * keys (examples/Main.C:66-67 '-m'), TreeType::eOct with one subtree
 * (tp_key = 1), GravityVisitor with gconst = 1; the quadrupole build and,
 * in the or_hexa_* and *_hexa functions, the -DHEXADECAPOLE build.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_spline / or_direct   -- tests/gravity/direct.acc of the reference
 *                              (lambs.00200_subsamp_30K), plus SPLINE continuity
 *                              at u = 1 and u = 2 and the 1/r^2 far field.
 *   or_keys                 -- closed-form bit interleave, brute force over all
 *                              cells at 1-3 levels, monotonicity per axis.
 *   or_sort                 -- numpy lexsort (key, index).
 *   or_build                -- brute-force prefix counting on small N and the
 *                              structural invariants (ranges partition parents,
 *                              Internal <=> count > B, 8 children, depth).
 *   or_moments              -- closed forms M = sum m, cm = sum m x / M,
 *                              Q = sum m (x - cm)(x - cm)^T, exact bbox, radius
 *                              formulae, for every node.
 *   or_intersect_box_point  -- brute-force predicate cases with exact ties.
 *   or_direct_vjp           -- central finite differences of or_direct on
 *                              small sets (positions and masses, every
 *                              softening branch), and the symmetry identity
 *                              sum_i g_i = 0 for translation invariance.
 *   or_hexa_*, *_hexa       -- reduced moments against tensor algebra on
 *                              the particles of every node, the potential
 *                              against the Legendre series, acceleration
 *                              against its gradient, fifth-order truncation,
 *                              the hexadecapole open() by brute force, and
 *                              the reference's stated accuracy on lambs 30K
 *                              (tests/gravity/acc_test.sh:5).
 *   or_gravity / or_lists   -- interaction sets against a brute-force
 *                              enumeration of the opening predicate; forces
 *                              against or_direct (tree error ~1e-3 at
 *                              theta = 0.7) and against the closed-form
 *                              quadrupole expansion of a far cluster.
 *
 * Definitions of pieces that live in the unvendored N-BodyShop/utility
 * submodule (no source in the reference tree; parity there is against these
 * written definitions):
 *   SFC::generateKey (Morton): u = (x - lesser) / (greater - lesser) in FP64,
 *     q = clamp(floor(u * 2^21), 0, 2^21 - 1) per axis, bits interleaved with x
 *     the most significant bit of each triple.
 *   OrientedBox: empty box = (+inf, -inf); grow(point) per-axis min/max with
 *     strict compares; grow(box) = per-axis min of lessers / max of greaters
 *     (so an empty box is a no-op); center() = (lesser + greater) / 2.
 *   Space::intersect(box, point, rsq): per-axis clamp distance squared,
 *     accumulated x, y, z, then dsq <= rsq.  intersect(box, sphere) is the
 *     same with rsq = radius * radius.  intersect(sphere, sphere):
 *     |c1 - c2|^2 <= (r1 + r2)^2.
 *   Vector3D / scalar: per-component division.
 */
#include "pt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_KEY_BITS 64
#define OR_MAX_DEPTH 21 /* 63 key bits / 3 bits per level */

/* ------------------------------------------------------------------------- */
/* Universe and keys                                                           */
/* ------------------------------------------------------------------------- */

/* BoundingBox::grow over particles (src/Reader.C:97, src/BoundingBox.C:20-22)
 * then Driver::remakeUniverse (src/Driver.h:67-80). */
void or_universe(const double* x, const double* y, const double* z, int64_t n,
                 double box_out[6]) {
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; i++) {
    const double p[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; k++) {
      if (p[k] < lo[k]) lo[k] = p[k];
      if (p[k] > hi[k]) hi[k] = p[k];
    }
  }
  double bs[3], bc[3];
  for (int k = 0; k < 3; k++) {
    bs[k] = hi[k] - lo[k];
    bc[k] = (lo[k] + hi[k]) / 2.0;
  }
  double mx = (bs[0] > bs[1]) ? bs[0] : bs[1];
  mx = (mx > bs[2]) ? mx : bs[2];
  const double fEps = 1.0 + 1.91e-6;
  const double half = fEps * 0.5 * mx;
  for (int k = 0; k < 3; k++) {
    box_out[k] = bc[k] - half;
    box_out[3 + k] = bc[k] + half;
  }
}

static uint64_t quantise(double v, double lo, double hi) {
  const double u = (v - lo) / (hi - lo);
  const double s = floor(u * 2097152.0);
  if (!(s > 0.0)) return 0; /* also maps NaN to 0 */
  if (s >= 2097151.0) return 2097151u;
  return (uint64_t)s;
}

void or_keys(const double* x, const double* y, const double* z, int64_t n,
             const double box[6], uint64_t* keys) {
  for (int64_t i = 0; i < n; i++) {
    const uint64_t q[3] = {quantise(x[i], box[0], box[3]),
                           quantise(y[i], box[1], box[4]),
                           quantise(z[i], box[2], box[5])};
    uint64_t k = 0;
    for (int b = 20; b >= 0; b--) {
      k = (k << 3) | (((q[0] >> b) & 1u) << 2) | (((q[1] >> b) & 1u) << 1) |
          ((q[2] >> b) & 1u);
    }
    keys[i] = k | ((uint64_t)1 << (OR_KEY_BITS - 1));
  }
}

typedef struct {
  uint64_t key;
  int64_t idx;
} kv_t;

static int kv_cmp(const void* a, const void* b) {
  const kv_t* p = (const kv_t*)a;
  const kv_t* q = (const kv_t*)b;
  if (p->key != q->key) return p->key < q->key ? -1 : 1;
  return (p->idx > q->idx) - (p->idx < q->idx);
}

void or_sort(const uint64_t* keys, int64_t n, uint64_t* sorted_keys, int32_t* perm) {
  kv_t* a = (kv_t*)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) {
    a[i].key = keys[i];
    a[i].idx = i;
  }
  qsort(a, (size_t)n, sizeof(kv_t), kv_cmp);
  for (int64_t i = 0; i < n; i++) {
    sorted_keys[i] = a[i].key;
    perm[i] = (int32_t)a[i].idx;
  }
  free(a);
}

/* ------------------------------------------------------------------------- */
/* Tree build                                                                  */
/* ------------------------------------------------------------------------- */

/* Utility::mssb64_pos (src/Utility.h:9-21). */
static int mssb64_pos(uint64_t x) {
  int n;
  if (x == 0) return -1;
  n = 0;
  if (x > 0x00000000FFFFFFFFull) { n += 32; x >>= 32; }
  if (x > 0x000000000000FFFFull) { n += 16; x >>= 16; }
  if (x > 0x00000000000000FFull) { n += 8; x >>= 8; }
  if (x > 0x000000000000000Full) { n += 4; x >>= 4; }
  if (x > 0x0000000000000003ull) { n += 2; x >>= 2; }
  if (x > 0x0000000000000001ull) { n += 1; }
  return n;
}

/* Utility::removeLeadingZeros (src/Utility.h:91-93,105-108,127-130), lbf = 3. */
static uint64_t remove_leading_zeros(uint64_t k) {
  const int depth = mssb64_pos(k) / 3;
  return k << (OR_KEY_BITS - (3 * depth + 1));
}

/* OctTree::findChildsLastParticle (src/Modularization.C:60-65) with
 * Utility::binarySearchComp (src/Utility.h:34-50). */
static int64_t find_childs_last(const uint64_t* k, int64_t start, int64_t finish,
                                uint64_t child_key) {
  const uint64_t splitter = remove_leading_zeros(child_key + 1);
  int64_t lo = start, hi = finish;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (k[mid] >= splitter) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

/* Subtree::getType (src/Subtree.h:262-267). */
static int32_t node_type(int64_t n, int bucket, int depth) {
  if (n == 0) return OR_EMPTY;
  if (n <= bucket) return OR_LEAF;
  /* Keys hold 21 levels; a deeper split is impossible (all remaining keys
   * are equal).  Such a node is kept as an over-full leaf. */
  if (depth >= OR_MAX_DEPTH) return OR_LEAF;
  return OR_INTERNAL;
}

typedef struct {
  const uint64_t* k;
  int bucket;
  or_node* nodes;
  int64_t nn, cap;
} build_ctx;

static int64_t new_node(build_ctx* c, uint64_t key, int32_t type, int32_t depth,
                        int64_t begin, int64_t count, int32_t parent) {
  const int64_t id = c->nn++;
  if (id < c->cap) {
    or_node* nd = &c->nodes[id];
    nd->key = key;
    nd->type = type;
    nd->depth = depth;
    nd->begin = begin;
    nd->count = count;
    nd->parent = parent;
    for (int i = 0; i < 8; i++) nd->child[i] = -1;
  }
  return id;
}

/* Subtree::recursiveBuild (src/Subtree.h:313-349). */
static void recursive_build(build_ctx* c, int64_t id, uint64_t key, int32_t depth,
                            int64_t begin, int64_t count) {
  uint64_t child_key = key << 3;
  int64_t start = begin;
  const int64_t finish = begin + count;
  for (int i = 0; i < 8; i++) {
    int64_t first_ge = finish;
    if (i < 7) first_ge = find_childs_last(c->k, start, finish, child_key);
    const int64_t n = first_ge - start;
    const int32_t t = node_type(n, c->bucket, depth + 1);
    const int64_t cid = new_node(c, child_key, t, depth + 1, start, n, (int32_t)id);
    if (id < c->cap) c->nodes[id].child[i] = (int32_t)cid;
    if (t == OR_INTERNAL) recursive_build(c, cid, child_key, depth + 1, start, n);
    start = first_ge;
    child_key++;
  }
}

/* Subtree::buildTree (src/Subtree.h:275-311) after the sort, tp_key = 1. */
int64_t or_build(const uint64_t* sorted_keys, int64_t n, int bucket,
                 or_node* nodes, int64_t cap) {
  build_ctx c = {sorted_keys, bucket, nodes, 0, cap};
  const int32_t t = node_type(n, bucket, 0);
  const int64_t root = new_node(&c, 1, t, 0, 0, n, -1);
  if (t == OR_INTERNAL) recursive_build(&c, root, 1, 0, 0, n);
  return c.nn <= cap ? c.nn : -c.nn;
}

int64_t or_buckets(const or_node* nodes, int64_t nn, int32_t* bucket_nodes) {
  int64_t nb = 0;
  for (int64_t i = 0; i < nn; i++) {
    if (nodes[i].type == OR_LEAF) {
      if (bucket_nodes) bucket_nodes[nb] = (int32_t)i;
      nb++;
    }
  }
  return nb;
}

/* ------------------------------------------------------------------------- */
/* Moments                                                                     */
/* ------------------------------------------------------------------------- */

static void cdata_default(or_cdata* d) {
  memset(d, 0, sizeof(*d));
  for (int k = 0; k < 3; k++) {
    d->lo[k] = INFINITY;
    d->hi[k] = -INFINITY;
  }
}

/* MultipoleMoments::operator+=(particle) (examples/MultipoleMoments.h:275-353,
 * quadrupole branch :331-350). */
static void mom_add_particle(or_cdata* d, double px, double py, double pz,
                             double pm, double pe) {
  const double m1 = d->mass;
  d->mass += pm;
  if (d->mass == 0.0) {
    d->soft = 0.5 * (d->soft + pe);
    d->cm[0] = 0.5 * (d->cm[0] + px);
    d->cm[1] = 0.5 * (d->cm[1] + py);
    d->cm[2] = 0.5 * (d->cm[2] + pz);
    return;
  }
  d->soft = (m1 * d->soft + pm * pe) / d->mass;
  const double c1[3] = {d->cm[0], d->cm[1], d->cm[2]};
  const double p[3] = {px, py, pz};
  for (int k = 0; k < 3; k++) d->cm[k] = (m1 * d->cm[k] + pm * p[k]) / d->mass;
  double dr[3] = {c1[0] - d->cm[0], c1[1] - d->cm[1], c1[2] - d->cm[2]};
  d->q[0] += m1 * dr[0] * dr[0];
  d->q[3] += m1 * dr[1] * dr[1];
  d->q[5] += m1 * dr[2] * dr[2];
  d->q[1] += m1 * dr[0] * dr[1];
  d->q[2] += m1 * dr[0] * dr[2];
  d->q[4] += m1 * dr[1] * dr[2];
  for (int k = 0; k < 3; k++) dr[k] = p[k] - d->cm[k];
  d->q[0] += pm * dr[0] * dr[0];
  d->q[3] += pm * dr[1] * dr[1];
  d->q[5] += pm * dr[2] * dr[2];
  d->q[1] += pm * dr[0] * dr[1];
  d->q[2] += pm * dr[0] * dr[2];
  d->q[4] += pm * dr[1] * dr[2];
}

/* MultipoleMoments::operator+=(MultipoleMoments) (examples/MultipoleMoments.h:
 * 204-271, quadrupole branch :253-269). */
static void mom_merge(or_cdata* d, const or_cdata* m) {
  const double m1 = d->mass;
  d->mass += m->mass;
  if (d->mass == 0.0) {
    d->soft = 0.5 * (d->soft + m->soft);
    for (int k = 0; k < 3; k++) d->cm[k] = 0.5 * (d->cm[k] + m->cm[k]);
    return;
  }
  if (m1 == 0.0) { /* *this = m: every MultipoleMoments field */
    d->radius = m->radius;
    d->soft = m->soft;
    d->mass = m->mass;
    for (int k = 0; k < 3; k++) d->cm[k] = m->cm[k];
    for (int k = 0; k < 6; k++) d->q[k] = m->q[k];
    return;
  }
  if (m->mass == 0.0) return;
  d->soft = (m1 * d->soft + m->mass * m->soft) / d->mass;
  const double c1[3] = {d->cm[0], d->cm[1], d->cm[2]};
  for (int k = 0; k < 3; k++) d->cm[k] = (m1 * d->cm[k] + m->mass * m->cm[k]) / d->mass;
  double dr[3] = {c1[0] - d->cm[0], c1[1] - d->cm[1], c1[2] - d->cm[2]};
  d->q[0] += m1 * dr[0] * dr[0];
  d->q[3] += m1 * dr[1] * dr[1];
  d->q[5] += m1 * dr[2] * dr[2];
  d->q[1] += m1 * dr[0] * dr[1];
  d->q[2] += m1 * dr[0] * dr[2];
  d->q[4] += m1 * dr[1] * dr[2];
  for (int k = 0; k < 3; k++) dr[k] = m->cm[k] - d->cm[k];
  d->q[0] += m->q[0] + m->mass * dr[0] * dr[0];
  d->q[3] += m->q[3] + m->mass * dr[1] * dr[1];
  d->q[5] += m->q[5] + m->mass * dr[2] * dr[2];
  d->q[1] += m->q[1] + m->mass * dr[0] * dr[1];
  d->q[2] += m->q[2] + m->mass * dr[0] * dr[2];
  d->q[4] += m->q[4] + m->mass * dr[1] * dr[2];
}

static double len3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

/* calculateRadiusFarthestCorner (examples/MultipoleMoments.h:458-469). */
static void radius_farthest_corner(or_cdata* d) {
  double dl[3];
  for (int k = 0; k < 3; k++) {
    const double d1 = d->cm[k] - d->lo[k];
    const double d2 = d->hi[k] - d->cm[k];
    dl[k] = (d1 > d2 ? d1 : d2);
  }
  d->radius = len3(dl[0], dl[1], dl[2]);
}

/* CentroidData(const Particle*, n, depth) (examples/CentroidData.h:41-59). */
static void leaf_data(or_cdata* d, const double* x, const double* y, const double* z,
                      const double* m, const double* e, int64_t b, int64_t n) {
  cdata_default(d);
  for (int64_t i = b; i < b + n; i++) {
    mom_add_particle(d, x[i], y[i], z[i], m[i], e[i]);
    const double p[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; k++) {
      if (p[k] < d->lo[k]) d->lo[k] = p[k];
      if (p[k] > d->hi[k]) d->hi[k] = p[k];
    }
  }
  d->count = n;
  if (d->count > 1) {
    /* calculateRadiusBox (examples/MultipoleMoments.h:473-482) */
    d->radius = 0.5 * len3(d->hi[0] - d->lo[0], d->hi[1] - d->lo[1], d->hi[2] - d->lo[2]);
  } else {
    d->radius = 1.0;
  }
}

/* CentroidData::operator+= (examples/CentroidData.h:61-73). */
static void cdata_merge(or_cdata* d, const or_cdata* c) {
  for (int k = 0; k < 3; k++) {
    if (c->lo[k] < d->lo[k]) d->lo[k] = c->lo[k];
    if (c->hi[k] > d->hi[k]) d->hi[k] = c->hi[k];
  }
  mom_merge(d, c);
  d->count += c->count;
  if (d->count > 1) radius_farthest_corner(d);
  else d->radius = 1.0;
}

void or_moments(const or_node* nodes, int64_t nn,
                const double* x, const double* y, const double* z,
                const double* m, const double* eps, or_cdata* out) {
  if (nn <= 0) return;
  int64_t* queue = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
  int32_t* wait = (int32_t*)malloc(sizeof(int32_t) * (size_t)nn);
  int64_t qt = 0;
  for (int64_t i = 0; i < nn; i++) {
    wait[i] = nodes[i].type == OR_INTERNAL ? 8 : 0;
    if (nodes[i].type == OR_LEAF) {
      leaf_data(&out[i], x, y, z, m, eps, nodes[i].begin, nodes[i].count);
      queue[qt++] = i;
    } else {
      cdata_default(&out[i]);
    }
  }
  for (int64_t i = 0; i < nn; i++)
    if (nodes[i].type == OR_EMPTY) queue[qt++] = i;
  /* Subtree::populateTree (src/Subtree.h:351-384): FIFO queue, a parent is
   * enqueued once all 8 children have been merged into it. */
  int64_t qh = 0;
  while (qh < qt) {
    const int64_t id = queue[qh++];
    if (id == 0) continue; /* the root (tp_key) */
    const int32_t p = nodes[id].parent;
    cdata_merge(&out[p], &out[id]);
    if (--wait[p] == 0) queue[qt++] = p;
  }
  free(queue);
  free(wait);
}

/* ------------------------------------------------------------------------- */
/* HEXADECAPOLE moments (the -DHEXADECAPOLE build, examples/Makefile:5)        */
/* ------------------------------------------------------------------------- */
/*
 * The reference keeps pkdgrav-style reduced (symmetric trace-free) moments up
 * to fourth order, scaled by the node radius (FMOMR, examples/moments.h; made
 * per particle by momMakeFmomr and moved to a parent by momShiftFmomr and
 * momScaledAddFmomr, examples/MultipoleMoments.h:244-252,317-330).  As written
 * its leaf update runs only `if (radius > 0)`, and the radius is still 0 while
 * the leaf constructor adds its particles (CentroidData.h:41-59 sets it after
 * the loop), so every FMOMR stays zero.  This restatement builds the moments
 * the code intends: for every node the trace-free parts of
 *   R(a,b,c) = sum_j m_j dx^a dy^b dz^c,  d = x_j - cm,  a + b + c = 2, 3, 4,
 * unscaled (Real is double, so the radius scaling changes nothing but
 * rounding).  Leaves sum their particles in sorted order; a parent sums its
 * children's raw moments moved to its centre of mass with the binomial
 * expansion of (d + s)^a (d + s)^b (d + s)^c, s = cm_child - cm_parent, in
 * child order (a child's dipole about its own centre of mass is zero).  Second
 * moments are the quadrupole build's q (the same quantity, merged in the
 * reference's order).  Field order of the 22 reduced values (FMOMR,
 * examples/moments.h): m, xx yy xy xz yz, xxx xyy xxy yyy xxz yyz xyz,
 * xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz.
 */

/* index of component (a, b, .) among the (l+1)(l+2)/2 of order l, ordered by
 * a descending, then b descending */
static int hx_index(int l, int a, int b) { return (l - a) * (l - a + 1) / 2 + (l - a - b); }

static const double hx_binom[5][5] = {
    {1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};

/* raw moment of exponents (a, b, c) from (mass, q, R3 | R4) */
static double hx_raw(double mass, const double q[6], const double* r, int a, int b, int c) {
  const int l = a + b + c;
  if (l == 0) return mass;
  if (l == 1) return 0.0;
  if (l == 2) return q[hx_index(2, a, b)];
  if (l == 3) return r[hx_index(3, a, b)];
  return r[10 + hx_index(4, a, b)];
}

static void hx_pow(double v, double p[5]) {
  p[0] = 1.0;
  p[1] = v;
  p[2] = p[1] * v;
  p[3] = p[2] * v;
  p[4] = p[3] * v;
}

/* leaf: R3, R4 about cm from the particles [b, b + n) in order */
static void hx_leaf(double r[25], const double cm[3], const double* x, const double* y,
                    const double* z, const double* m, int64_t b, int64_t n) {
  for (int k = 0; k < 25; k++) r[k] = 0.0;
  for (int64_t i = b; i < b + n; i++) {
    double px[5], py[5], pz[5];
    hx_pow(x[i] - cm[0], px);
    hx_pow(y[i] - cm[1], py);
    hx_pow(z[i] - cm[2], pz);
    for (int l = 3; l <= 4; l++)
      for (int a = l; a >= 0; a--)
        for (int bb = l - a; bb >= 0; bb--) {
          const int c = l - a - bb;
          r[(l == 3 ? 0 : 10) + hx_index(l, a, bb)] += ((m[i] * px[a]) * py[bb]) * pz[c];
        }
  }
}

/* parent += child moved by s = cm_child - cm_parent */
static void hx_shift_add(double r[25], const double s[3], double mass, const double q[6],
                         const double rc[25]) {
  double px[5], py[5], pz[5];
  hx_pow(s[0], px);
  hx_pow(s[1], py);
  hx_pow(s[2], pz);
  for (int l = 3; l <= 4; l++)
    for (int a = l; a >= 0; a--)
      for (int b = l - a; b >= 0; b--) {
        const int c = l - a - b;
        double sum = 0.0;
        for (int i = 0; i <= a; i++)
          for (int j = 0; j <= b; j++)
            for (int k = 0; k <= c; k++) {
              if (i + j + k == 1) continue;
              const double coef = hx_binom[a][i] * hx_binom[b][j] * hx_binom[c][k];
              sum += (((coef * px[a - i]) * py[b - j]) * pz[c - k]) * hx_raw(mass, q, rc, i, j, k);
            }
        r[(l == 3 ? 0 : 10) + hx_index(l, a, b)] += sum;
      }
}

/* trace-free (reduced) moments from mass, q and R3 | R4 */
static void hx_reduce(double out[22], double mass, const double q[6], const double r[25]) {
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
  out[0] = mass;
  out[1] = q[0] - tr * (1.0 / 3.0);  /* xx */
  out[2] = q[3] - tr * (1.0 / 3.0);  /* yy */
  out[3] = q[1];                     /* xy */
  out[4] = q[2];                     /* xz */
  out[5] = q[4];                     /* yz */
  out[6] = R3(3, 0) - vx * (3.0 / 5.0);  /* xxx */
  out[7] = R3(1, 2) - vx * (1.0 / 5.0);  /* xyy */
  out[8] = R3(2, 1) - vy * (1.0 / 5.0);  /* xxy */
  out[9] = R3(0, 3) - vy * (3.0 / 5.0);  /* yyy */
  out[10] = R3(2, 0) - vz * (1.0 / 5.0); /* xxz */
  out[11] = R3(0, 2) - vz * (1.0 / 5.0); /* yyz */
  out[12] = R3(1, 1);                    /* xyz */
  out[13] = (R4(4, 0) - txx * (6.0 / 7.0)) + ss * (3.0 / 35.0);  /* xxxx */
  out[14] = R4(1, 3) - txy * (3.0 / 7.0);                        /* xyyy */
  out[15] = R4(3, 1) - txy * (3.0 / 7.0);                        /* xxxy */
  out[16] = (R4(0, 4) - tyy * (6.0 / 7.0)) + ss * (3.0 / 35.0);  /* yyyy */
  out[17] = R4(3, 0) - txz * (3.0 / 7.0);                        /* xxxz */
  out[18] = R4(0, 3) - tyz * (3.0 / 7.0);                        /* yyyz */
  out[19] = (R4(2, 2) - (txx + tyy) * (1.0 / 7.0)) + ss * (1.0 / 35.0);  /* xxyy */
  out[20] = R4(2, 1) - tyz * (1.0 / 7.0);                        /* xxyz */
  out[21] = R4(1, 2) - txz * (1.0 / 7.0);                        /* xyyz */
#undef R3
#undef R4
}

void or_hexa_moments(const or_node* nodes, int64_t nn, const or_cdata* d, const double* x,
                     const double* y, const double* z, const double* m, double* raw,
                     double* out) {
  if (nn <= 0) return;
  double* r = raw ? raw : (double*)malloc(sizeof(double) * 25 * (size_t)nn);
  /* pre-order: every child follows its parent, so a reverse sweep sees
   * children first */
  for (int64_t i = nn - 1; i >= 0; i--) {
    double* ri = r + 25 * i;
    if (nodes[i].type == OR_LEAF) {
      hx_leaf(ri, d[i].cm, x, y, z, m, nodes[i].begin, nodes[i].count);
    } else {
      for (int k = 0; k < 25; k++) ri[k] = 0.0;
      if (nodes[i].type == OR_INTERNAL && d[i].mass != 0.0) {
        for (int c = 0; c < 8; c++) {
          const int32_t ch = nodes[i].child[c];
          if (d[ch].mass == 0.0) continue;
          const double s[3] = {d[ch].cm[0] - d[i].cm[0], d[ch].cm[1] - d[i].cm[1],
                               d[ch].cm[2] - d[i].cm[2]};
          hx_shift_add(ri, s, d[ch].mass, d[ch].q, r + 25 * ch);
        }
      }
    }
    hx_reduce(out + 22 * i, d[i].mass, d[i].q, ri);
  }
  if (!raw) free(r);
}

void or_hexa_raw_moments(const double* x, const double* y, const double* z, const double* m,
                         int64_t n, const double cm[3], double out[25]) {
  hx_leaf(out, cm, x, y, z, m, 0, n);
}

void or_hexa_reduce(double mass, const double q[6], const double r[25], double out[22]) {
  hx_reduce(out, mass, q, r);
}

/* Potential and acceleration at r = target - cm of the reduced moments h
 * (GravityVisitor::node -> momEvalFmomrcm, examples/GravityVisitor.h:199-206,
 * examples/moments.C:1469-1526).  With c_l = (2l-1)!!/l! (3/2, 5/2, 35/8) and
 * v_l = M_l contracted with l-1 copies of r, s_l = v_l . r:
 *   pot = -(m / r + sum_l c_l s_l / r^(2l+1))
 *   acc = sum_l c_l (l v_l / r^(2l+1) - (2l+1) s_l r / r^(2l+3)) - m r / r^3. */
void or_hexa_eval(const double h[22], double rx, double ry, double rz, double* pot,
                  double acc[3]) {
  const double m = h[0];
  const double xx = h[1], yy = h[2], xy = h[3], xz = h[4], yz = h[5];
  const double zz = -(xx + yy);
  const double xxx = h[6], xyy = h[7], xxy = h[8], yyy = h[9], xxz = h[10], yyz = h[11],
               xyz = h[12];
  const double xzz = -(xxx + xyy), yzz = -(xxy + yyy), zzz = -(xxz + yyz);
  const double xxxx = h[13], xyyy = h[14], xxxy = h[15], yyyy = h[16], xxxz = h[17],
               yyyz = h[18], xxyy = h[19], xxyz = h[20], xyyz = h[21];
  const double xxzz = -(xxxx + xxyy), yyzz = -(xxyy + yyyy), xyzz = -(xxxy + xyyy);
  const double xzzz = -(xxxz + xyyz), yzzz = -(xxyz + yyyz), zzzz = -(xxzz + yyzz);
  const double rsq = rx * rx + ry * ry + rz * rz;
  const double ir = 1.0 / sqrt(rsq);
  const double ir2 = ir * ir;
  const double x2 = rx * rx, y2 = ry * ry, z2 = rz * rz;
  const double xyp = rx * ry, xzp = rx * rz, yzp = ry * rz;
  /* order 2 */
  const double v2x = xx * rx + xy * ry + xz * rz;
  const double v2y = xy * rx + yy * ry + yz * rz;
  const double v2z = xz * rx + yz * ry + zz * rz;
  const double s2 = v2x * rx + v2y * ry + v2z * rz;
  /* order 3 */
  const double v3x = xxx * x2 + xyy * y2 + xzz * z2 + 2.0 * (xxy * xyp + xxz * xzp + xyz * yzp);
  const double v3y = xxy * x2 + yyy * y2 + yzz * z2 + 2.0 * (xyy * xyp + xyz * xzp + yyz * yzp);
  const double v3z = xxz * x2 + yyz * y2 + zzz * z2 + 2.0 * (xyz * xyp + xzz * xzp + yzz * yzp);
  const double s3 = v3x * rx + v3y * ry + v3z * rz;
  /* order 4 */
  const double x3 = x2 * rx, y3 = y2 * ry, z3 = z2 * rz, xyz3 = xyp * rz;
  const double v4x = xxxx * x3 + xyyy * y3 + xzzz * z3 +
                     3.0 * (xxxy * x2 * ry + xxxz * x2 * rz + xxyy * rx * y2 + xxzz * rx * z2 +
                            xyyz * y2 * rz + xyzz * ry * z2) +
                     6.0 * xxyz * xyz3;
  const double v4y = xxxy * x3 + yyyy * y3 + yzzz * z3 +
                     3.0 * (xxyy * x2 * ry + xxyz * x2 * rz + xyyy * rx * y2 + xyzz * rx * z2 +
                            yyyz * y2 * rz + yyzz * ry * z2) +
                     6.0 * xyyz * xyz3;
  const double v4z = xxxz * x3 + yyyz * y3 + zzzz * z3 +
                     3.0 * (xxyz * x2 * ry + xxzz * x2 * rz + xyyz * rx * y2 + xzzz * rx * z2 +
                            yyzz * y2 * rz + yzzz * ry * z2) +
                     6.0 * xyzz * xyz3;
  const double s4 = v4x * rx + v4y * ry + v4z * rz;
  const double ir3 = ir * ir2, ir5 = ir3 * ir2, ir7 = ir5 * ir2, ir9 = ir7 * ir2,
               ir11 = ir9 * ir2;
  *pot = -(m * ir + 1.5 * s2 * ir5 + 2.5 * s3 * ir7 + 4.375 * s4 * ir9);
  const double g = m * ir3 + 7.5 * s2 * ir7 + 17.5 * s3 * ir9 + 39.375 * s4 * ir11;
  acc[0] = 3.0 * v2x * ir5 + 7.5 * v3x * ir7 + 17.5 * v4x * ir9 - rx * g;
  acc[1] = 3.0 * v2y * ir5 + 7.5 * v3y * ir7 + 17.5 * v4y * ir9 - ry * g;
  acc[2] = 3.0 * v2z * ir5 + 7.5 * v3z * ir7 + 17.5 * v4z * ir9 - rz * g;
}

/* ------------------------------------------------------------------------- */
/* GravityVisitor                                                              */
/* ------------------------------------------------------------------------- */

/* GravityVisitor::SPLINE (examples/GravityVisitor.h:87-120). */
void or_spline(double r2, double twoh, double* pa, double* pb) {
  const double r = sqrt(r2);
  double a, b;
  if (r < twoh) {
    const double dih = 2.0 / twoh;
    const double u = r * dih;
    if (u < 1.0) {
      a = dih * (7.0 / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u -
                 1.0 / 10.0 * u * u * u * u * u);
      b = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
    } else {
      const double dir = 1.0 / r;
      a = -1.0 / 15.0 * dir +
          dih * (8.0 / 5.0 - 4.0 / 3.0 * u * u + u * u * u - 3.0 / 10.0 * u * u * u * u +
                 1.0 / 30.0 * u * u * u * u * u);
      b = -1.0 / 15.0 * dir * dir * dir +
          dih * dih * dih * (8.0 / 3.0 - 3.0 * u + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
    }
  } else {
    a = 1.0 / r;
    b = a * a * a;
  }
  *pa = a;
  *pb = b;
}

/* GravityVisitor::SPLINEQ (examples/GravityVisitor.h:37-80), written exactly
 * as the reference has it (including its u < 1 polynomial grouping). */
void or_splineq(double invr, double r2, double twoh,
                double* pa, double* pb, double* pc, double* pd) {
  double u, dih, dir = invr;
  double a, b, c, d;
  if (r2 < twoh * twoh) {
    dih = 2.0 / twoh;
    u = dih / dir;
    if (u < 1.0) {
      a = dih * (7.0) / 5.0 - 2.0 / 3.0 * u * u + 3.0 / 10.0 * u * u * u * u -
          1.0 / 10.0 * u * u * u * u * u;
      b = dih * dih * dih * (4.0) / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u;
      c = dih * dih * dih * dih * dih * (12.0) / 5.0 - 3.0 / 2.0 * u;
      d = 3.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
    } else {
      a = -1.0 / 15.0 * dir + dih * (8.0 / 5.0) - 4.0 / 3.0 * u * u + u * u * u -
          3.0 / 10.0 * u * u * u * u + 1.0 / 30.0 * u * u * u * u * u;
      b = -1.0 / 15.0 * dir * dir * dir + dih * dih * dih * (8.0) / 3.0 - 3.0 * u +
          6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u;
      c = -1.0 / 5.0 * dir * dir * dir * dir * dir + 3.0 * dih * dih * dih * dih * dir +
          dih * dih * dih * dih * dih * (-12.0) / 5.0 + 1.0 / 2.0 * u;
      d = -dir * dir * dir * dir * dir * dir * dir + 3.0 * dih * dih * dih * dih * dir * dir * dir -
          1.0 / 2.0 * dih * dih * dih * dih * dih * dih * dir;
    }
  } else {
    a = dir;
    b = a * a * a;
    c = 3.0 * b * a * a;
    d = 5.0 * c * a * a;
  }
  *pa = a;
  *pb = b;
  *pc = c;
  *pd = d;
}

int or_intersect_box_point(const double lo[3], const double hi[3],
                           const double p[3], double rsq) {
  double dsq = 0.0;
  for (int k = 0; k < 3; k++) {
    if (p[k] < lo[k]) {
      const double t = lo[k] - p[k];
      dsq += t * t;
    } else if (p[k] > hi[k]) {
      const double t = p[k] - hi[k];
      dsq += t * t;
    }
  }
  return dsq <= rsq;
}

static int intersect_sphere_sphere(const double c1[3], double r1, const double c2[3],
                                   double r2) {
  const double dx = c1[0] - c2[0], dy = c1[1] - c2[1], dz = c1[2] - c2[2];
  const double dsq = dx * dx + dy * dy + dz * dz;
  const double rs = r1 + r2;
  return dsq <= rs * rs;
}

typedef struct {
  const or_node* nodes;
  const or_cdata* d;
  const double *x, *y, *z, *m, *e;
  double gravity_factor;
  double mono_factor;   /* GravityVisitor::monopole_gravity_factor */
  const double* hexa;   /* reduced moments per node (22 each); NULL = quadrupole build */
  double off[3];
  double *ax, *ay, *az, *pot;
  /* lists mode */
  int64_t npp, npc;
  int32_t *pp, *pc;
  uint64_t cnt[4];
  /* P-C target pairs that reach SPLINEQ's softened branch (r^2 < twoh^2
   * with openSoftening false), u < 1 and u >= 1: a coverage count for tests */
  uint64_t soft_pc[2];
} trav_ctx;

/* GravityVisitor::open, non-HEXADECAPOLE branch (examples/GravityVisitor.h:
 * 162-164,180). */
static int gv_open(const trav_ctx* t, const or_cdata* src, const or_cdata* tgt) {
  if (src->count <= 6) return 1; /* nMinParticleNode (:35) */
  const double dataRsq = src->radius * src->radius * t->gravity_factor;
  const double p[3] = {src->cm[0] + t->off[0], src->cm[1] + t->off[1], src->cm[2] + t->off[2]};
  return or_intersect_box_point(tgt->lo, tgt->hi, p, dataRsq);
}

/* GravityVisitor::openSoftening (examples/GravityVisitor.h:122-128). */
static int gv_open_softening(const trav_ctx* t, const or_cdata* src, const or_cdata* tgt) {
  const double sc[3] = {src->cm[0] + t->off[0], src->cm[1] + t->off[1], src->cm[2] + t->off[2]};
  const double sr = 2.0 * src->soft;
  if (intersect_sphere_sphere(sc, sr, tgt->cm, 2.0 * tgt->soft)) return 1;
  return or_intersect_box_point(tgt->lo, tgt->hi, sc, sr * sr);
}

/* GravityVisitor::open, HEXADECAPOLE branch (examples/GravityVisitor.h:
 * 162-178).  The softening test is called as written there, with the target
 * bucket's data in the source role: openSoftening(target.data, source.data). */
static int gv_open_hexa(const trav_ctx* t, const or_cdata* src, const or_cdata* tgt) {
  if (src->count <= 6) return 1;
  const double dataRsq = src->radius * src->radius * t->gravity_factor;
  const double p[3] = {src->cm[0] + t->off[0], src->cm[1] + t->off[1], src->cm[2] + t->off[2]};
  if (!or_intersect_box_point(tgt->lo, tgt->hi, p, dataRsq)) {
    if (!gv_open_softening(t, tgt, src)) return 0;
    const double monoRad = src->radius * t->mono_factor;
    return or_intersect_box_point(tgt->lo, tgt->hi, p, monoRad * monoRad);
  }
  return 1;
}

/* GravityVisitor::leaf (examples/GravityVisitor.h:144-160). */
static void gv_leaf(trav_ctx* t, const or_node* src, const or_node* tgt) {
  for (int64_t i = tgt->begin; i < tgt->begin + tgt->count; i++) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = src->begin; j < src->begin + src->count; j++) {
      const double dx = (t->x[j] + t->off[0]) - t->x[i];
      const double dy = (t->y[j] + t->off[1]) - t->y[i];
      const double dz = (t->z[j] + t->off[2]) - t->z[i];
      const double rsq = dx * dx + dy * dy + dz * dz;
      const double twoh = t->e[j] + t->e[i];
      if (rsq != 0) {
        double a, b;
        or_spline(rsq, twoh, &a, &b);
        const double f = b * t->m[j];
        ax += dx * f;
        ay += dy * f;
        az += dz * f;
      }
    }
    t->ax[i] += ax;
    t->ay[i] += ay;
    t->az[i] += az;
  }
}

/* GravityVisitor::node + addGravity (examples/GravityVisitor.h:130-139,184-224). */
static void gv_node(trav_ctx* t, const or_cdata* s, const or_node* tgt, const or_cdata* td) {
  if (s->count == 0) return;
  if (gv_open_softening(t, s, td)) {
    for (int64_t i = tgt->begin; i < tgt->begin + tgt->count; i++) {
      const double dx = (s->cm[0] + t->off[0]) - t->x[i];
      const double dy = (s->cm[1] + t->off[1]) - t->y[i];
      const double dz = (s->cm[2] + t->off[2]) - t->z[i];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq != 0) {
        const double f = s->mass / (rsq * sqrt(rsq));
        t->ax[i] += dx * f;
        t->ay[i] += dy * f;
        t->az[i] += dz * f;
      }
    }
    return;
  }
  for (int64_t i = tgt->begin; i < tgt->begin + tgt->count; i++) {
    const double rx = (t->x[i] - s->cm[0]) - t->off[0];
    const double ry = (t->y[i] - s->cm[1]) - t->off[1];
    const double rz = (t->z[i] - s->cm[2]) - t->off[2];
    const double rsq = rx * rx + ry * ry + rz * rz;
    const double dir = 1.0 / sqrt(rsq);
    const double twoh = s->soft + t->e[i];
    double a, b, c, d;
    if (rsq < twoh * twoh) t->soft_pc[(2.0 / twoh) / dir < 1.0 ? 0 : 1]++;
    or_splineq(dir, rsq, twoh, &a, &b, &c, &d);
    const double qx = s->q[0] * rx + s->q[1] * ry + s->q[2] * rz;
    const double qy = s->q[1] * rx + s->q[3] * ry + s->q[4] * rz;
    const double qz = s->q[2] * rx + s->q[4] * ry + s->q[5] * rz;
    const double qir = 0.5 * (qx * rx + qy * ry + qz * rz);
    const double tr = 0.5 * (s->q[0] + s->q[3] + s->q[5]);
    const double qir3 = b * s->mass + d * qir - c * tr;
    t->pot[i] += -s->mass * a - c * qir + b * tr;
    t->ax[i] += (-qir3 * rx) + (c * qx);
    t->ay[i] += (-qir3 * ry) + (c * qy);
    t->az[i] += (-qir3 * rz) + (c * qz);
  }
}

/* GravityVisitor::node, HEXADECAPOLE branch (examples/GravityVisitor.h:
 * 184-206): addGravity when the softening spheres meet, else the reduced
 * moments evaluated at every target (potential and acceleration). */
static void gv_node_hexa(trav_ctx* t, int32_t si, const or_node* tgt, const or_cdata* td) {
  const or_cdata* s = &t->d[si];
  if (s->count == 0) return;
  if (gv_open_softening(t, s, td)) {
    for (int64_t i = tgt->begin; i < tgt->begin + tgt->count; i++) {
      const double dx = (s->cm[0] + t->off[0]) - t->x[i];
      const double dy = (s->cm[1] + t->off[1]) - t->y[i];
      const double dz = (s->cm[2] + t->off[2]) - t->z[i];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq != 0) {
        const double f = s->mass / (rsq * sqrt(rsq));
        t->ax[i] += dx * f;
        t->ay[i] += dy * f;
        t->az[i] += dz * f;
      }
    }
    return;
  }
  const double* h = t->hexa + 22 * (int64_t)si;
  for (int64_t i = tgt->begin; i < tgt->begin + tgt->count; i++) {
    const double rx = (t->x[i] - s->cm[0]) - t->off[0];
    const double ry = (t->y[i] - s->cm[1]) - t->off[1];
    const double rz = (t->z[i] - s->cm[2]) - t->off[2];
    double pot, acc[3];
    or_hexa_eval(h, rx, ry, rz, &pot, acc);
    t->pot[i] += pot;
    t->ax[i] += acc[0];
    t->ay[i] += acc[1];
    t->az[i] += acc[2];
  }
}

/* TransposedDownTraverser::recurse (src/Traverser.h:154-217) seen from one
 * target bucket: the sequence of visitor calls this bucket receives. */
static void visit(trav_ctx* t, int32_t ni, int32_t bn, int lists) {
  const or_node* nd = &t->nodes[ni];
  const or_node* tg = &t->nodes[bn];
  switch (nd->type) {
    case OR_LEAF:
      if (lists) {
        if (t->pp) t->pp[t->npp] = ni;
        t->npp++;
      } else {
        gv_leaf(t, nd, tg);
      }
      t->cnt[1] += (uint64_t)(nd->count * tg->count);
      break;
    case OR_INTERNAL: {
      const int open = t->hexa ? gv_open_hexa(t, &t->d[ni], &t->d[bn])
                               : gv_open(t, &t->d[ni], &t->d[bn]);
      t->cnt[open ? 2 : 3]++;
      if (open) {
        for (int c = 0; c < 8; c++) visit(t, nd->child[c], bn, lists);
      } else {
        if (lists) {
          if (t->pc) t->pc[t->npc] = ni;
          t->npc++;
        } else if (t->hexa) {
          gv_node_hexa(t, ni, tg, &t->d[bn]);
        } else {
          gv_node(t, &t->d[ni], tg, &t->d[bn]);
        }
        t->cnt[0] += (uint64_t)tg->count;
      }
      break;
    }
    default:
      break;
  }
}

static double gravity_factor(double theta) {
  /* GravityVisitor ctor (examples/GravityVisitor.h:14-19) */
  const double opening_geometry_factor_squared = 4.0 / 3.0;
  return opening_geometry_factor_squared / (theta * theta);
}

static double mono_factor(double theta) {
  /* GravityVisitor ctor (examples/GravityVisitor.h:16-19) */
  const double opening_geometry_factor_squared = 4.0 / 3.0;
  return sqrt(opening_geometry_factor_squared) / (theta * theta * theta * theta);
}

static void gravity_impl(const or_node* nodes, int64_t nn, const or_cdata* d,
                         const double* hexa, const double* x, const double* y, const double* z,
                         const double* m, const double* eps, double theta,
                         const double offset[3], const int32_t* bucket_sel, int64_t n_sel,
                         double* ax, double* ay, double* az, double* pot, uint64_t counters[4],
                         int nthreads, uint64_t* soft_pairs) {
  const int64_t nb = or_buckets(nodes, nn, NULL);
  int32_t* bn = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  or_buckets(nodes, nn, bn);
  const int64_t ns = bucket_sel ? n_sel : nb;
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, s0 = 0, s1 = 0;
  const double gf = gravity_factor(theta);
  const double gm = mono_factor(theta);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads) \
    reduction(+ : c0, c1, c2, c3, s0, s1)
#endif
  for (int64_t s = 0; s < ns; s++) {
    const int32_t b = bn[bucket_sel ? bucket_sel[s] : s];
    trav_ctx t;
    memset(&t, 0, sizeof(t));
    t.nodes = nodes;
    t.d = d;
    t.x = x; t.y = y; t.z = z; t.m = m; t.e = eps;
    t.gravity_factor = gf;
    t.mono_factor = gm;
    t.hexa = hexa;
    for (int k = 0; k < 3; k++) t.off[k] = offset[k];
    t.ax = ax; t.ay = ay; t.az = az; t.pot = pot;
    for (int64_t i = nodes[b].begin; i < nodes[b].begin + nodes[b].count; i++) {
      ax[i] = ay[i] = az[i] = pot[i] = 0.0;
    }
    visit(&t, 0, b, 0);
    c0 += t.cnt[0];
    c1 += t.cnt[1];
    c2 += t.cnt[2];
    c3 += t.cnt[3];
    s0 += t.soft_pc[0];
    s1 += t.soft_pc[1];
  }
  counters[0] = c0;
  counters[1] = c1;
  counters[2] = c2;
  counters[3] = c3;
  if (soft_pairs) {
    soft_pairs[0] = s0;
    soft_pairs[1] = s1;
  }
  free(bn);
  (void)nthreads;
}

void or_gravity(const or_node* nodes, int64_t nn, const or_cdata* d,
                const double* x, const double* y, const double* z,
                const double* m, const double* eps,
                double theta, const double offset[3],
                const int32_t* bucket_sel, int64_t n_sel,
                double* ax, double* ay, double* az, double* pot,
                uint64_t counters[4], int nthreads) {
  gravity_impl(nodes, nn, d, NULL, x, y, z, m, eps, theta, offset, bucket_sel, n_sel, ax, ay, az,
               pot, counters, nthreads, NULL);
}

void or_gravity_soft_pairs(const or_node* nodes, int64_t nn, const or_cdata* d,
                           const double* x, const double* y, const double* z,
                           const double* m, const double* eps, double theta,
                           const double offset[3], double* ax, double* ay, double* az,
                           double* pot, uint64_t counters[4], uint64_t soft_pairs[2]) {
  gravity_impl(nodes, nn, d, NULL, x, y, z, m, eps, theta, offset, NULL, 0, ax, ay, az, pot,
               counters, 0, soft_pairs);
}

void or_gravity_hexa(const or_node* nodes, int64_t nn, const or_cdata* d, const double* hexa,
                     const double* x, const double* y, const double* z, const double* m,
                     const double* eps, double theta, const double offset[3],
                     const int32_t* bucket_sel, int64_t n_sel, double* ax, double* ay,
                     double* az, double* pot, uint64_t counters[4], int nthreads) {
  gravity_impl(nodes, nn, d, hexa, x, y, z, m, eps, theta, offset, bucket_sel, n_sel, ax, ay, az,
               pot, counters, nthreads, NULL);
}

static void lists_impl(const or_node* nodes, int64_t nn, const or_cdata* d, const double* hexa,
                       double theta, const double offset[3], int64_t* pp_count,
                       int64_t* pc_count, const int64_t* pp_off, const int64_t* pc_off,
                       int32_t* pp, int32_t* pc) {
  const int64_t nb = or_buckets(nodes, nn, NULL);
  int32_t* bn = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  or_buckets(nodes, nn, bn);
  const double gf = gravity_factor(theta);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8)
#endif
  for (int64_t s = 0; s < nb; s++) {
    trav_ctx t;
    memset(&t, 0, sizeof(t));
    t.nodes = nodes;
    t.d = d;
    t.gravity_factor = gf;
    t.mono_factor = mono_factor(theta);
    t.hexa = hexa;
    for (int k = 0; k < 3; k++) t.off[k] = offset[k];
    if (pp && pc) {
      t.pp = pp + pp_off[s];
      t.pc = pc + pc_off[s];
    }
    visit(&t, 0, bn[s], 1);
    pp_count[s] = t.npp;
    pc_count[s] = t.npc;
  }
  free(bn);
}

void or_traversal_counters(const or_node* nodes, int64_t nn, const or_cdata* d, double theta,
                           const double offset[3], uint64_t counters[4], int nthreads) {
  const int64_t nb = or_buckets(nodes, nn, NULL);
  int32_t* bn = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  or_buckets(nodes, nn, bn);
  const double gf = gravity_factor(theta);
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : c0, c1, c2, c3)
#endif
  for (int64_t s = 0; s < nb; s++) {
    trav_ctx t;
    memset(&t, 0, sizeof(t));
    t.nodes = nodes;
    t.d = d;
    t.gravity_factor = gf;
    for (int k = 0; k < 3; k++) t.off[k] = offset[k];
    visit(&t, 0, bn[s], 1); /* lists mode with no arrays: the walk and counts only */
    c0 += t.cnt[0];
    c1 += t.cnt[1];
    c2 += t.cnt[2];
    c3 += t.cnt[3];
  }
  counters[0] = c0;
  counters[1] = c1;
  counters[2] = c2;
  counters[3] = c3;
  free(bn);
  (void)nthreads;
}

void or_lists(const or_node* nodes, int64_t nn, const or_cdata* d,
              double theta, const double offset[3],
              int64_t* pp_count, int64_t* pc_count,
              const int64_t* pp_off, const int64_t* pc_off,
              int32_t* pp, int32_t* pc) {
  lists_impl(nodes, nn, d, NULL, theta, offset, pp_count, pc_count, pp_off, pc_off, pp, pc);
}

void or_lists_hexa(const or_node* nodes, int64_t nn, const or_cdata* d, const double* hexa,
                   double theta, const double offset[3], int64_t* pp_count, int64_t* pc_count,
                   const int64_t* pp_off, const int64_t* pc_off, int32_t* pp, int32_t* pc) {
  lists_impl(nodes, nn, d, hexa, theta, offset, pp_count, pc_count, pp_off, pc_off, pp, pc);
}

void or_direct(const double* x, const double* y, const double* z,
               const double* m, const double* eps, int64_t n,
               const int64_t* targets, int64_t nt,
               double* ax, double* ay, double* az, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t s = 0; s < nt; s++) {
    const int64_t i = targets[s];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int64_t j = 0; j < n; j++) {
      const double dx = x[j] - x[i];
      const double dy = y[j] - y[i];
      const double dz = z[j] - z[i];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq != 0) {
        double a, b;
        or_spline(rsq, eps[j] + eps[i], &a, &b);
        const double f = b * m[j];
        sx += dx * f;
        sy += dy * f;
        sz += dz * f;
      }
    }
    ax[s] = sx;
    ay[s] = sy;
    az[s] = sz;
  }
  (void)nthreads;
}

/* d b / d r of GravityVisitor::SPLINE's force factor b
 * (examples/GravityVisitor.h:82-120), differentiated term by term with
 * u = 2 r / twoh. */
static double spline_db_dr(double r, double twoh) {
  if (r < twoh) {
    const double dih = 2.0 / twoh;
    const double u = r * dih;
    if (u < 1.0) return pow(dih, 4) * (-2.4 * u + 1.5 * u * u);
    return 0.2 / pow(r, 4) + pow(dih, 4) * (-3.0 + 2.4 * u - 0.5 * u * u);
  }
  return -3.0 / pow(r, 4);
}

/* Vector-Jacobian product of the direct sum (no counterpart in the
 * reference): for cotangents w on the accelerations of or_direct (all
 * targets), g = (da/dx)^T w and gm = (da/dm)^T w.  Serial loop over ordered
 * pairs with the explicit 3x3 block da_i/dd_ij = m_j (b I + b'/r d d^T),
 * d = x_j - x_i. */
void or_direct_vjp(const double* x, const double* y, const double* z,
                   const double* m, const double* eps, int64_t n,
                   const double* wx, const double* wy, const double* wz,
                   double* gx, double* gy, double* gz, double* gm) {
  for (int64_t i = 0; i < n; i++) gx[i] = gy[i] = gz[i] = gm[i] = 0.0;
  for (int64_t i = 0; i < n; i++) {
    const double w[3] = {wx[i], wy[i], wz[i]};
    for (int64_t j = 0; j < n; j++) {
      const double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
      const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq == 0) continue;
      const double r = sqrt(rsq);
      double a, b;
      or_spline(rsq, eps[j] + eps[i], &a, &b);
      const double bp = spline_db_dr(r, eps[j] + eps[i]);
      double J[3][3];
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) J[p][q] = m[j] * ((p == q ? b : 0.0) + bp / r * d[p] * d[q]);
      double jt[3] = {0, 0, 0};
      for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++) jt[q] += J[p][q] * w[p];
      gx[i] -= jt[0];
      gy[i] -= jt[1];
      gz[i] -= jt[2];
      gx[j] += jt[0];
      gy[j] += jt[1];
      gz[j] += jt[2];
      gm[j] += b * (d[0] * w[0] + d[1] * w[1] + d[2] * w[2]);
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Ewald sum of the periodic build (quadrupole form)                           */
/* ------------------------------------------------------------------------- */
/* Relative perturbation of the erf / erfc and exp values in the real-space
 * Ewald terms (both forms), 1.0 = none.  The hole terms g1..g5 cancel
 * catastrophically just outside the series radius, so a one-ulp change of
 * these values moves the result by far more than one ulp; the parity tests
 * set this to measure that sensitivity of the restated algorithm. */
static double or_ew_erf_scale = 1.0, or_ew_exp_scale = 1.0;

void or_set_ewald_libm_perturbation(double erf_ulps, double exp_ulps) {
  or_ew_erf_scale = 1.0 + erf_ulps * 0x1p-52;
  or_ew_exp_scale = 1.0 + exp_ulps * 0x1p-52;
}

/*
 * LeafEwaldFn with EwaldData::EwaldInit (examples/Ewald.h, examples/Ewald.C,
 * non-HEXADECAPOLE branches): after the explicit replica traversals, every
 * particle gets the rest of the infinite lattice of the root's monopole +
 * quadrupole (mass M, centre cm, raw second moments q) by the Ewald split
 * with alpha = a_fac / L: real-space erfc terms over the lattice vectors
 * within ew_cut * L (those of the explicit replicas -- the "hole" --
 * contribute -erf instead, removing the bare 1/r part the tree already
 * summed; a Taylor series replaces the closed forms below r^2 = 1.2e-3 L^2),
 * and the Fourier terms of the integer vectors 0 < |h| <= h_cut.  The
 * reference uses a_fac = 2, ew_cut = 2.6, h_cut = 2.8 and L = dxPeriod.
 * Adds to ax/ay/az/pot.
 */
void or_ewald_general(const double* x, const double* y, const double* z, int64_t n,
                      double mass, const double cm[3], const double q[6], double L,
                      int n_replicas, double a_fac, double ew_cut, double h_cut, double* ax,
                      double* ay, double* az, double* pot) {
  const double pi = 3.14159265358979323846;
  const double alpha = a_fac / L, alpha2 = alpha * alpha;
  const double k1 = pi / (alpha2 * L * L * L);
  const double ka = 2.0 * alpha / sqrt(pi);
  const double q2 = 0.5 * (q[0] + q[3] + q[5]);
  int nrep = (int)ceil(ew_cut);
  if (n_replicas > nrep) nrep = n_replicas;
  const double cut2 = ew_cut * ew_cut * L * L, inner2 = 1.2e-3 * L * L;
  /* Fourier table: C(h) = -(e^{-k4 h^2} / (pi h^2 L)) (M - (2 pi / L)^2 h.q.h / 2); the sine
   * coefficients vanish at quadrupole order */
  const int hr = (int)ceil(h_cut);
  const double k4 = pi * pi / (alpha2 * L * L);
  int nh = 0;
  double* hv = (double*)malloc(sizeof(double) * 4 * (size_t)(2 * hr + 1) * (2 * hr + 1) * (2 * hr + 1));
  for (int hx = -hr; hx <= hr; hx++)
    for (int hy = -hr; hy <= hr; hy++)
      for (int hz = -hr; hz <= hr; hz++) {
        const int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > h_cut * h_cut) continue;
        const double g0 = exp(-k4 * h2) / (pi * h2 * L);
        const double w = 2.0 * pi / L;
        const double hqh = 0.5 * (hx * (q[0] * hx + q[1] * hy + q[2] * hz) +
                                  hy * (q[1] * hx + q[3] * hy + q[4] * hz) +
                                  hz * (q[2] * hx + q[4] * hy + q[5] * hz));
        hv[4 * nh + 0] = w * hx;
        hv[4 * nh + 1] = w * hy;
        hv[4 * nh + 2] = w * hz;
        hv[4 * nh + 3] = -(g0 * mass - w * w * g0 * hqh);
        nh++;
      }
  for (int64_t i = 0; i < n; i++) {
    double p = mass * k1, fx = 0.0, fy = 0.0, fz = 0.0;
    const double dx = x[i] - cm[0], dy = y[i] - cm[1], dz = z[i] - cm[2];
    for (int ix = -nrep; ix <= nrep; ix++) {
      const int hole_x = ix >= -n_replicas && ix <= n_replicas;
      const double rx = dx + ix * L;
      for (int iy = -nrep; iy <= nrep; iy++) {
        const int hole_xy = hole_x && iy >= -n_replicas && iy <= n_replicas;
        const double ry = dy + iy * L;
        for (int iz = -nrep; iz <= nrep; iz++) {
          const int hole = hole_xy && iz >= -n_replicas && iz <= n_replicas;
          const double rz = dz + iz * L;
          double r2 = rx * rx + ry * ry + rz * rz;
          if (r2 > cut2 && !hole) continue;
          double g0, g1, g2, g3;
          if (r2 < inner2) {
            /* series of -erf(alpha r)/r and its derivatives about r = 0 */
            const double s2 = r2 * alpha2;
            double an = ka;
            g0 = an * ((1.0 / 3.0) * s2 - 1.0);
            an *= 2.0 * alpha2;
            g1 = an * ((1.0 / 5.0) * s2 - (1.0 / 3.0));
            an *= 2.0 * alpha2;
            g2 = an * ((1.0 / 7.0) * s2 - (1.0 / 5.0));
            an *= 2.0 * alpha2;
            g3 = an * ((1.0 / 9.0) * s2 - (1.0 / 7.0));
          } else {
            const double r = sqrt(r2);
            const double ir2 = 1.0 / r2;
            const double a = (exp(-r2 * alpha2) * or_ew_exp_scale) * ka * ir2;
            g0 = ((hole ? -erf(alpha * r) : erfc(alpha * r)) * or_ew_erf_scale) / r;
            g1 = g0 * ir2 + a;
            g2 = 3.0 * g1 * ir2 + 2.0 * alpha2 * a;
            g3 = 5.0 * g2 * ir2 + 4.0 * alpha2 * alpha2 * a;
          }
          const double qx = q[0] * rx + q[1] * ry + q[2] * rz;
          const double qy = q[1] * rx + q[3] * ry + q[4] * rz;
          const double qz = q[2] * rx + q[4] * ry + q[5] * rz;
          const double qrr = 0.5 * (qx * rx + qy * ry + qz * rz);
          const double qta = g1 * mass - g2 * q2 + g3 * qrr;
          p -= g0 * mass - g1 * q2 + g2 * qrr;
          fx += g2 * qx - rx * qta;
          fy += g2 * qy - ry * qta;
          fz += g2 * qz - rz * qta;
        }
      }
    }
    for (int k = 0; k < nh; k++) {
      const double hd = hv[4 * k] * dx + hv[4 * k + 1] * dy + hv[4 * k + 2] * dz;
      const double c = cos(hd), sn = sin(hd), cf = hv[4 * k + 3];
      p += cf * c;
      fx += hv[4 * k] * (cf * sn);
      fy += hv[4 * k + 1] * (cf * sn);
      fz += hv[4 * k + 2] * (cf * sn);
    }
    ax[i] += fx;
    ay[i] += fy;
    az[i] += fz;
    pot[i] += p;
  }
  free(hv);
}

void or_ewald(const double* x, const double* y, const double* z, int64_t n, double mass,
              const double cm[3], const double q[6], double L, int n_replicas, double* ax,
              double* ay, double* az, double* pot) {
  or_ewald_general(x, y, z, n, mass, cm, q, L, n_replicas, 2.0, 2.6, 2.8, ax, ay, az, pot);
}

void or_kick_drift(double* x, double* y, double* z,
                   double* vx, double* vy, double* vz,
                   const double* ax, const double* ay, const double* az,
                   int64_t n, double dt) {
  for (int64_t i = 0; i < n; i++) {
    /* Particle::kick (src/Particle.C:7-9) */
    vx[i] += ax[i] * dt / 2;
    vy[i] += ay[i] * dt / 2;
    vz[i] += az[i] * dt / 2;
    /* Particle::perturb (src/Particle.C:11-15) */
    vx[i] += (ax[i] * dt / 2);
    vy[i] += (ay[i] * dt / 2);
    vz[i] += (az[i] * dt / 2);
    x[i] += (vx[i] * dt);
    y[i] += (vy[i] * dt);
    z[i] += (vz[i] * dt);
  }
}

/* ------------------------------------------------------------------------- */
/* Dual-tree traversal                                                         */
/* ------------------------------------------------------------------------- */

/* Space::enclose(A, B) as defined in the header comment: A's box contains
 * B's box, per-axis lesser_A <= lesser_B and greater_B <= greater_A. */
static int enclose_box(const or_cdata* a, const or_cdata* b) {
  for (int k = 0; k < 3; k++)
    if (!(a->lo[k] <= b->lo[k]) || !(b->hi[k] <= a->hi[k])) return 0;
  return 1;
}

/* SpatialNode::n_particles as the reference's nodes hold it: -1 for an
 * internal node (src/Node.h:97-107), 0 for an empty leaf, the count for a
 * leaf. */
static int64_t n_particles(const or_node* nd) {
  return nd->type == OR_INTERNAL ? -1 : (nd->type == OR_LEAF ? nd->count : 0);
}

typedef struct {
  trav_ctx t;
  const int32_t* bucket_of; /* node -> bucket ordinal, -1 if not a bucket */
  int64_t *pp_count, *pc_count;
  const int64_t *pp_off, *pc_off;
  int32_t *pp, *pc;
} dual_ctx;

typedef struct {
  int32_t s, t;
} dual_pair;

typedef struct {
  dual_pair* v;
  int64_t n, cap;
} dual_stack;

static void ds_push(dual_stack* st, int32_t s, int32_t t) {
  if (st->n == st->cap) {
    st->cap = st->cap ? 2 * st->cap : 256;
    st->v = (dual_pair*)realloc(st->v, sizeof(dual_pair) * (size_t)st->cap);
  }
  st->v[st->n].s = s;
  st->v[st->n].t = t;
  st->n++;
}

static void add_count(uint64_t* c, int64_t v) { *c += (uint64_t)v; }

/* doLeaf (src/Traverser.h:23-29): GravityVisitor::leaf plus the count
 * n_source * n_target with the reference's n_particles values. */
static void dual_leaf(dual_ctx* c, int32_t s, int32_t t) {
  const or_node* sn = &c->t.nodes[s];
  const or_node* tn = &c->t.nodes[t];
  add_count(&c->t.cnt[1], n_particles(sn) * n_particles(tn));
  /* leaf() loops over the target's and the source's n_particles: only two
   * leaves with particles interact */
  if (sn->type != OR_LEAF || tn->type != OR_LEAF) return;
  const int32_t b = c->bucket_of[t];
  if (c->pp_count) {
    if (c->pp) c->pp[c->pp_off[b] + c->pp_count[b]] = s;
    c->pp_count[b]++;
  } else {
    gv_leaf(&c->t, sn, tn);
  }
}

/* doNode (src/Traverser.h:31-37): GravityVisitor::node plus the count
 * n_target. */
static void dual_node(dual_ctx* c, int32_t s, int32_t t) {
  const or_node* tn = &c->t.nodes[t];
  add_count(&c->t.cnt[0], n_particles(tn));
  /* node() loops over the target's n_particles (-1 for an internal node) and
   * returns at once for a source without particles */
  if (tn->type != OR_LEAF || c->t.d[s].count == 0) return;
  const int32_t b = c->bucket_of[t];
  if (c->pc_count) {
    if (c->pc) c->pc[c->pc_off[b] + c->pc_count[b]] = s;
    c->pc_count[b]++;
  } else {
    gv_node(&c->t, &c->t.d[s], tn, &c->t.d[t]);
  }
}

/* doOpen / doCell (src/Traverser.h:15-21,39-45) with the counts. */
static int dual_open(dual_ctx* c, int32_t s, int32_t t) {
  const int o = gv_open(&c->t, &c->t.d[s], &c->t.d[t]);
  c->t.cnt[o ? 2 : 3]++;
  return o;
}

static int dual_cell(dual_ctx* c, int32_t s, int32_t t) {
  /* GravityVisitor::cell (examples/GravityVisitor.h:226-230) */
  const int o = !enclose_box(&c->t.d[s], &c->t.d[t]);
  c->t.cnt[o ? 2 : 3]++;
  return o;
}

/* DualTraverser::runInvertedTraversal (src/Traverser.h:509-529). */
static void dual_inverted(dual_ctx* c, int32_t source, int32_t target) {
  dual_stack st = {0, 0, 0};
  ds_push(&st, source, target);
  while (st.n > 0) {
    const int32_t node = st.v[--st.n].t;
    const or_node* nd = &c->t.nodes[node];
    if (nd->type != OR_INTERNAL) {          /* isLeaf(): Leaf or EmptyLeaf */
      dual_leaf(c, source, node);
    } else if (dual_open(c, source, node)) {
      for (int j = 0; j < 8; j++) ds_push(&st, source, nd->child[j]);
    } else {
      dual_node(c, source, node);
    }
  }
  free(st.v);
}

/* DualTraverser::doTrav (src/Traverser.h:539-608) from one (source, target)
 * pair, GravityVisitor's flags TargetMustBeLeaf and ForceEvenDepth
 * (examples/GravityVisitor.h:12-13).  The pairs are held on a LIFO stack as
 * the reference's std::stack. */
static void dual_trav(dual_ctx* c, int32_t source, int32_t target) {
  dual_stack st = {0, 0, 0};
  ds_push(&st, source, target);
  while (st.n > 0) {
    st.n--;
    const int32_t s = st.v[st.n].s, t = st.v[st.n].t;
    const or_node* sn = &c->t.nodes[s];
    const or_node* tn = &c->t.nodes[t];
    if (tn->type == OR_EMPTY) continue;
    if (sn->type == OR_LEAF) {
      if (tn->type != OR_INTERNAL) dual_leaf(c, s, t);
      else dual_inverted(c, s, t);
    } else if (sn->type == OR_INTERNAL) {
      if (tn->type != OR_INTERNAL) {
        if (dual_open(c, s, t)) {
          for (int i = 0; i < 8; i++) ds_push(&st, sn->child[i], t);
        } else {
          dual_node(c, s, t);
        }
      } else if (!dual_cell(c, s, t)) {
        dual_inverted(c, s, t);
      } else {
        for (int i = 0; i < 8; i++) {
          if (tn->depth <= sn->depth) {
            for (int j = 0; j < 8; j++) ds_push(&st, sn->child[i], tn->child[j]);
          } else {
            ds_push(&st, sn->child[i], t);
          }
        }
      }
    }
    /* an EmptyLeaf source: the switch's default, nothing */
  }
  free(st.v);
}

static void dual_setup(dual_ctx* c, const or_node* nodes, int64_t nn, const or_cdata* d,
                       double theta, const double offset[3], int32_t* bucket_of) {
  memset(c, 0, sizeof(*c));
  c->t.nodes = nodes;
  c->t.d = d;
  c->t.gravity_factor = gravity_factor(theta);
  c->t.mono_factor = mono_factor(theta);
  for (int k = 0; k < 3; k++) c->t.off[k] = offset[k];
  for (int64_t i = 0; i < nn; i++) bucket_of[i] = -1;
  int64_t b = 0;
  for (int64_t i = 0; i < nn; i++)
    if (nodes[i].type == OR_LEAF) bucket_of[i] = (int32_t)b++;
  c->bucket_of = bucket_of;
}

void or_dual(const or_node* nodes, int64_t nn, const or_cdata* d, const double* x,
             const double* y, const double* z, const double* m, const double* eps, double theta,
             const double offset[3], int32_t source, int32_t target, double* ax, double* ay,
             double* az, double* pot, int64_t n, uint64_t counters[4]) {
  int32_t* bo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn > 0 ? nn : 1));
  dual_ctx c;
  dual_setup(&c, nodes, nn, d, theta, offset, bo);
  c.t.x = x; c.t.y = y; c.t.z = z; c.t.m = m; c.t.e = eps;
  c.t.ax = ax; c.t.ay = ay; c.t.az = az; c.t.pot = pot;
  for (int64_t i = 0; i < n; i++) ax[i] = ay[i] = az[i] = pot[i] = 0.0;
  dual_trav(&c, source, target);
  for (int k = 0; k < 4; k++) counters[k] = c.t.cnt[k];
  free(bo);
}

void or_dual_lists(const or_node* nodes, int64_t nn, const or_cdata* d, double theta,
                   const double offset[3], int32_t source, int32_t target, int64_t* pp_count,
                   int64_t* pc_count, const int64_t* pp_off, const int64_t* pc_off, int32_t* pp,
                   int32_t* pc) {
  int32_t* bo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn > 0 ? nn : 1));
  dual_ctx c;
  dual_setup(&c, nodes, nn, d, theta, offset, bo);
  const int64_t nb = or_buckets(nodes, nn, NULL);
  for (int64_t b = 0; b < nb; b++) pp_count[b] = pc_count[b] = 0;
  c.pp_count = pp_count;
  c.pc_count = pc_count;
  c.pp_off = pp_off;
  c.pc_off = pc_off;
  c.pp = pp;
  c.pc = pc;
  dual_trav(&c, source, target);
  free(bo);
}

/* ------------------------------------------------------------------------- */
/* Ewald, HEXADECAPOLE form                                                    */
/* ------------------------------------------------------------------------- */

/* The root's reduced moments (FMOMR order: xx yy xy xz yz, xxx xyy xxy yyy
 * xxz yyz xyz, xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz) completed with
 * their z components by tracelessness, as momFmomr2Momc does
 * (examples/moments.C:1563-1597). */
typedef struct {
  double m;
  double xx, yy, xy, xz, yz, zz;
  double xxx, xyy, xxy, yyy, xxz, yyz, xyz, xzz, yzz, zzz;
  double xxxx, xyyy, xxxy, yyyy, xxxz, yyyz, xxyy, xxyz, xyyz, xxzz, xyzz, xzzz, yyzz, yzzz,
      zzzz;
} or_momc;

static void momc_from_reduced(double mass, const double h[21], or_momc* c) {
  c->m = mass;
  c->xx = h[0]; c->yy = h[1]; c->xy = h[2]; c->xz = h[3]; c->yz = h[4];
  c->xxx = h[5]; c->xyy = h[6]; c->xxy = h[7]; c->yyy = h[8]; c->xxz = h[9]; c->yyz = h[10];
  c->xyz = h[11];
  c->xxxx = h[12]; c->xyyy = h[13]; c->xxxy = h[14]; c->yyyy = h[15]; c->xxxz = h[16];
  c->yyyz = h[17]; c->xxyy = h[18]; c->xxyz = h[19]; c->xyyz = h[20];
  c->zz = -(c->xx + c->yy);
  c->xzz = -(c->xxx + c->xyy);
  c->yzz = -(c->xxy + c->yyy);
  c->zzz = -(c->xxz + c->yyz);
  c->xxzz = -(c->xxxx + c->xxyy);
  c->xyzz = -(c->xxxy + c->xyyy);
  c->xzzz = -(c->xxxz + c->xyyz);
  c->yyzz = -(c->xxyy + c->yyyy);
  c->yzzz = -(c->xxyz + c->yyyz);
  c->zzzz = -(c->xxzz + c->yyzz);
}

/* M2.v, M3[., v, v] / 2 and M4[., v, v, v] / 6 of the complete moments. */
static void momc_contract(const or_momc* M, double x, double y, double z, double q2[3],
                          double q3[3], double q4[3]) {
  q2[0] = M->xx * x + M->xy * y + M->xz * z;
  q2[1] = M->xy * x + M->yy * y + M->yz * z;
  q2[2] = M->xz * x + M->yz * y + M->zz * z;
  const double xx = 0.5 * x * x, yy = 0.5 * y * y, zz = 0.5 * z * z;
  const double xy = x * y, xz = x * z, yz = y * z;
  q3[0] = M->xxx * xx + M->xxy * xy + M->xxz * xz + M->xyy * yy + M->xyz * yz + M->xzz * zz;
  q3[1] = M->xxy * xx + M->xyy * xy + M->xyz * xz + M->yyy * yy + M->yyz * yz + M->yzz * zz;
  q3[2] = M->xxz * xx + M->xyz * xy + M->xzz * xz + M->yyz * yy + M->yzz * yz + M->zzz * zz;
  const double onethird = 1.0 / 3.0;
  const double xxx = onethird * xx * x, xxy = xx * y, xxz = xx * z, xyy = yy * x, yyy = onethird * yy * y;
  const double yyz = yy * z, xzz = zz * x, yzz = zz * y, zzz = onethird * zz * z, xyz = xy * z;
  q4[0] = M->xxxx * xxx + M->xxxy * xxy + M->xxxz * xxz + M->xxyy * xyy + M->xxyz * xyz +
          M->xxzz * xzz + M->xyyy * yyy + M->xyyz * yyz + M->xyzz * yzz + M->xzzz * zzz;
  q4[1] = M->xxxy * xxx + M->xxyy * xxy + M->xxyz * xxz + M->xyyy * xyy + M->xyyz * xyz +
          M->xyzz * xzz + M->yyyy * yyy + M->yyyz * yyz + M->yyzz * yzz + M->yzzz * zzz;
  q4[2] = M->xxxz * xxx + M->xxyz * xxy + M->xxzz * xxz + M->xyyz * xyy + M->xyzz * xyz +
          M->xzzz * xzz + M->yyyz * yyy + M->yyzz * yyz + M->yzzz * yzz + M->zzzz * zzz;
}

/* LeafEwaldFn + EwaldData::EwaldInit, HEXADECAPOLE branches
 * (examples/Ewald.h:6-13,31-42,57-67,127-166; examples/Ewald.C:11-51,89-158):
 * the lattice of the root's moments to fourth order (mass, centre of mass,
 * reduced moments h in FMOMR order) beyond the explicit replicas.  Real
 * space: g0..g5 of -erf / erfc over the lattice vectors within ew_cut L (the
 * replicas' hole with -erf; the series below r^2 = 1.2e-3 L^2); Fourier
 * space: for the integer vectors 0 < |h| <= h_cut, with w = 2 pi / L and
 * g = e^{-k4 h^2} / (pi h^2 L), the cosine coefficient
 * -g (m - w^2 M2[h^2] / 2 + w^4 M4[h^4] / 24) and the sine coefficient
 * g w^3 M3[h^3] / 6 (QEVAL of the complete moments with the odd, then the
 * even, gam[] zeroed).  alpha = a_fac / L. */
void or_ewald_hexa_general(const double* x, const double* y, const double* z, int64_t n,
                           double mass, const double cm[3], const double h[21], double L,
                           int n_replicas, double a_fac, double ew_cut, double h_cut, double* ax,
                           double* ay, double* az, double* pot) {
  const double pi = 3.14159265358979323846;
  or_momc M;
  momc_from_reduced(mass, h, &M);
  /* traces of the complete moments (Ewald.h:32-43) */
  const double Q4xx = 0.5 * (M.xxxx + M.xxyy + M.xxzz), Q4xy = 0.5 * (M.xxxy + M.xyyy + M.xyzz);
  const double Q4xz = 0.5 * (M.xxxz + M.xyyz + M.xzzz), Q4yy = 0.5 * (M.xxyy + M.yyyy + M.yyzz);
  const double Q4yz = 0.5 * (M.xxyz + M.yyyz + M.yzzz), Q4zz = 0.5 * (M.xxzz + M.yyzz + M.zzzz);
  const double Q4 = 0.25 * (Q4xx + Q4yy + Q4zz);
  const double Q3x = 0.5 * (M.xxx + M.xyy + M.xzz), Q3y = 0.5 * (M.xxy + M.yyy + M.yzz);
  const double Q3z = 0.5 * (M.xxz + M.yyz + M.zzz);
  const double Q2 = 0.5 * (M.xx + M.yy + M.zz);
  const double alpha = a_fac / L, alpha2 = alpha * alpha;
  const double k1 = pi / (alpha2 * L * L * L);
  const double ka = 2.0 * alpha / sqrt(pi);
  int nrep = (int)ceil(ew_cut);
  if (n_replicas > nrep) nrep = n_replicas;
  const double cut2 = ew_cut * ew_cut * L * L, inner2 = 1.2e-3 * L * L;
  /* Fourier table {w h, C(h), S(h)} */
  const int hr = (int)ceil(h_cut);
  const double k4 = pi * pi / (alpha2 * L * L);
  const double w = 2.0 * pi / L;
  int nh = 0;
  double* hv = (double*)malloc(sizeof(double) * 5 * (size_t)(2 * hr + 1) * (2 * hr + 1) * (2 * hr + 1));
  for (int hx = -hr; hx <= hr; hx++)
    for (int hy = -hr; hy <= hr; hy++)
      for (int hz = -hr; hz <= hr; hz++) {
        const int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > h_cut * h_cut) continue;
        const double g = exp(-k4 * h2) / (pi * h2 * L);
        double q2[3], q3[3], q4[3];
        momc_contract(&M, hx, hy, hz, q2, q3, q4);
        const double m2 = 0.5 * (q2[0] * hx + q2[1] * hy + q2[2] * hz);
        const double m3 = (1.0 / 3.0) * (q3[0] * hx + q3[1] * hy + q3[2] * hz);
        const double m4 = 0.25 * (q4[0] * hx + q4[1] * hy + q4[2] * hz);
        hv[5 * nh + 0] = w * hx;
        hv[5 * nh + 1] = w * hy;
        hv[5 * nh + 2] = w * hz;
        hv[5 * nh + 3] = -g * (mass - w * w * m2 + w * w * w * w * m4);
        hv[5 * nh + 4] = g * w * w * w * m3;
        nh++;
      }
  for (int64_t i = 0; i < n; i++) {
    double p = mass * k1, fx = 0.0, fy = 0.0, fz = 0.0;
    const double dx = x[i] - cm[0], dy = y[i] - cm[1], dz = z[i] - cm[2];
    for (int ix = -nrep; ix <= nrep; ix++) {
      const int hole_x = ix >= -n_replicas && ix <= n_replicas;
      const double rx = dx + ix * L;
      for (int iy = -nrep; iy <= nrep; iy++) {
        const int hole_xy = hole_x && iy >= -n_replicas && iy <= n_replicas;
        const double ry = dy + iy * L;
        for (int iz = -nrep; iz <= nrep; iz++) {
          const int hole = hole_xy && iz >= -n_replicas && iz <= n_replicas;
          const double rz = dz + iz * L;
          double r2 = rx * rx + ry * ry + rz * rz;
          if (r2 > cut2 && !hole) continue;
          double g0, g1, g2, g3, g4, g5;
          if (r2 < inner2) {
            const double s2 = r2 * alpha2;
            double an = ka;
            g0 = an * ((1.0 / 3.0) * s2 - 1.0);
            an *= 2.0 * alpha2;
            g1 = an * ((1.0 / 5.0) * s2 - (1.0 / 3.0));
            an *= 2.0 * alpha2;
            g2 = an * ((1.0 / 7.0) * s2 - (1.0 / 5.0));
            an *= 2.0 * alpha2;
            g3 = an * ((1.0 / 9.0) * s2 - (1.0 / 7.0));
            an *= 2.0 * alpha2;
            g4 = an * ((1.0 / 11.0) * s2 - (1.0 / 9.0));
            an *= 2.0 * alpha2;
            g5 = an * ((1.0 / 13.0) * s2 - (1.0 / 11.0));
          } else {
            const double r = sqrt(r2);
            const double ir2 = 1.0 / r2;
            const double a = (exp(-r2 * alpha2) * or_ew_exp_scale) * ka * ir2;
            g0 = ((hole ? -erf(alpha * r) : erfc(alpha * r)) * or_ew_erf_scale) / r;
            g1 = g0 * ir2 + a;
            double an = 2.0 * alpha2;
            g2 = 3.0 * g1 * ir2 + an * a;
            an *= 2.0 * alpha2;
            g3 = 5.0 * g2 * ir2 + an * a;
            an *= 2.0 * alpha2;
            g4 = 7.0 * g3 * ir2 + an * a;
            an *= 2.0 * alpha2;
            g5 = 9.0 * g4 * ir2 + an * a;
          }
          double q2[3], q3[3], q4[3];
          momc_contract(&M, rx, ry, rz, q2, q3, q4);
          const double Q4vx = Q4xx * rx + Q4xy * ry + Q4xz * rz;
          const double Q4vy = Q4xy * rx + Q4yy * ry + Q4yz * rz;
          const double Q4vz = Q4xz * rx + Q4yz * ry + Q4zz * rz;
          const double Q2mir = 0.5 * (q2[0] * rx + q2[1] * ry + q2[2] * rz) -
                               (Q3x * rx + Q3y * ry + Q3z * rz) + Q4;
          const double Q3mir = (1.0 / 3.0) * (q3[0] * rx + q3[1] * ry + q3[2] * rz) -
                               0.5 * (Q4vx * rx + Q4vy * ry + Q4vz * rz);
          const double Q4mir = 0.25 * (q4[0] * rx + q4[1] * ry + q4[2] * rz);
          const double Qta = g1 * mass - g2 * Q2 + g3 * Q2mir + g4 * Q3mir + g5 * Q4mir;
          p -= g0 * mass - g1 * Q2 + g2 * Q2mir + g3 * Q3mir + g4 * Q4mir;
          fx += g2 * (q2[0] - Q3x) + g3 * (q3[0] - Q4vx) + g4 * q4[0] - rx * Qta;
          fy += g2 * (q2[1] - Q3y) + g3 * (q3[1] - Q4vy) + g4 * q4[1] - ry * Qta;
          fz += g2 * (q2[2] - Q3z) + g3 * (q3[2] - Q4vz) + g4 * q4[2] - rz * Qta;
        }
      }
    }
    for (int k = 0; k < nh; k++) {
      const double hd = hv[5 * k] * dx + hv[5 * k + 1] * dy + hv[5 * k + 2] * dz;
      const double c = cos(hd), sn = sin(hd), cc = hv[5 * k + 3], cs = hv[5 * k + 4];
      p += cc * c + cs * sn;
      const double f = cc * sn - cs * c;
      fx += hv[5 * k] * f;
      fy += hv[5 * k + 1] * f;
      fz += hv[5 * k + 2] * f;
    }
    ax[i] += fx;
    ay[i] += fy;
    az[i] += fz;
    pot[i] += p;
  }
  free(hv);
}

void or_ewald_hexa(const double* x, const double* y, const double* z, int64_t n, double mass,
                   const double cm[3], const double h[21], double L, int n_replicas, double* ax,
                   double* ay, double* az, double* pot) {
  or_ewald_hexa_general(x, y, z, n, mass, cm, h, L, n_replicas, 2.0, 2.6, 2.8, ax, ay, az, pot);
}

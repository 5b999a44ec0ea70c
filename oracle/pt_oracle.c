/*
 * pt_oracle.c -- CPU restatement of ParaTreeT's gravity hot path (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product (libpt_gravity.so, paratreet_amd) links or
 * calls this file.
 *
 * Reference configuration restated: Real = double (-DUSE_DOUBLE_FP), no
 * -DHEXADECAPOLE, Morton keys (examples/Main.C:66-67 '-m'), TreeType::eOct with
 * one subtree (tp_key = 1), GravityVisitor with gconst = 1.
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
  double off[3];
  double *ax, *ay, *az, *pot;
  /* lists mode */
  int64_t npp, npc;
  int32_t *pp, *pc;
  uint64_t cnt[4];
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
      const int open = gv_open(t, &t->d[ni], &t->d[bn]);
      t->cnt[open ? 2 : 3]++;
      if (open) {
        for (int c = 0; c < 8; c++) visit(t, nd->child[c], bn, lists);
      } else {
        if (lists) {
          if (t->pc) t->pc[t->npc] = ni;
          t->npc++;
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

void or_gravity(const or_node* nodes, int64_t nn, const or_cdata* d,
                const double* x, const double* y, const double* z,
                const double* m, const double* eps,
                double theta, const double offset[3],
                const int32_t* bucket_sel, int64_t n_sel,
                double* ax, double* ay, double* az, double* pot,
                uint64_t counters[4], int nthreads) {
  const int64_t nb = or_buckets(nodes, nn, NULL);
  int32_t* bn = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  or_buckets(nodes, nn, bn);
  const int64_t ns = bucket_sel ? n_sel : nb;
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  const double gf = gravity_factor(theta);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads) reduction(+ : c0, c1, c2, c3)
#endif
  for (int64_t s = 0; s < ns; s++) {
    const int32_t b = bn[bucket_sel ? bucket_sel[s] : s];
    trav_ctx t;
    memset(&t, 0, sizeof(t));
    t.nodes = nodes;
    t.d = d;
    t.x = x; t.y = y; t.z = z; t.m = m; t.e = eps;
    t.gravity_factor = gf;
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

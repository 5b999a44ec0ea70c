// Adjoint (vector-Jacobian products) of the gravity interactions, FP64 and
// FP32 (shifted coordinates, FP64 sums).
//
// For a cotangent w_i on every particle's acceleration these kernels compute
//   g_j  = sum_i (da_i/dx_j)^T w_i      (positions)
//   gm_j = sum_i (da_i/dm_j)^T w_i      (masses)
// for the tree force of GravityVisitor::node / leaf
// (examples/GravityVisitor.h:37-160,184-224) on the interactions the forward
// pass evaluates (the lists of pt_launch_group_traverse: per group of
// consecutive buckets, the union of the buckets' lists with member masks), and
// for the FP64 direct sum.
//
// leaf() pair (target t, source j, d = x_j + off - x_t):  a_t += m_j b(|d|) d:
//   da_t/dd = m_j K,  K = b I + (b'/|d|) d d^T  (symmetric), so the target
//   gets -m_j K w_t, the source +m_j K w_t and the source mass b d.w_t.
// node() pair (target t, node s, r = x_t - cm_s - off), quadrupole form
//   a = -qir3 r + c q,  q = Q r,  qir = r.q / 2,  qir3 = b M + d qir - c tr:
//   (da/dr)^T w = -qir3 w - k1 r (r.w) - d q (r.w) + c Q w + (c'/|r|) r (q.w),
//   k1 = (M b' + qir d' - tr c') / |r|;
//   the target gets it, the node centre gets minus it, the node mass
//   -b (r.w), each stored quadrupole component d(a.w)/dQ, and the node
//   softening d(a.w)/dtwoh (softened branch only).
// addGravity (the openSoftening branch): a = M dd/|dd|^3, dd = cm + off - x_t,
//   T = M (I/|dd|^3 - 3 dd dd^T/|dd|^5); the target gets -T w, the centre
//   +T w, the mass dd.w/|dd|^3.
// Node cotangents then flow to member particles through the closed forms of
// the moments (MultipoleMoments::operator+=, examples/MultipoleMoments.h:
// 204-270, is the parallel-axis merge of M = sum m, cm = sum m x / M,
// Q = sum m u u^T with u = x - cm, soft = sum m e / M):
//   dcm/dx_j = m_j/M I,  dQ_kl/dx_j = m_j (e_k u_l + u_k e_l),
//   dM/dm_j = 1, dcm/dm_j = u_j/M, dQ/dm_j = u_j u_j^T, dsoft/dm_j = (e_j - soft)/M.
// Opening decisions (open, openSoftening) are piecewise constant and held
// fixed.
#include "pt_common.h"
#include "pt_kernels.h"
#include "pt_walk.h"

#include <algorithm>

namespace {

constexpr int kAdjWaves = 4;
constexpr int kAdjThreads = kAdjWaves * PT_WAVE;
constexpr int kANT = PT_GROUP_TARGETS;  // targets per pass over a group's lists
constexpr int kNodeAdj = 11;  // node cotangent: cm(3), M, Q(6: xx xy xz yy yz zz), soft
// Nodes of the top levels (ids < kHotNodes in level order) are in the lists
// of most buckets; their cotangents are accumulated into kHotParts private
// copies (chosen by work-group) and summed afterwards, so concurrent atomic
// adds to one address come from 1/kHotParts of the work-groups.
constexpr int64_t kHotNodes = 16384;
constexpr int kHotParts = 32;

// GravityVisitor::SPLINE force factor b and db/dr (examples/GravityVisitor.h:
// 82-120); r is |d|, dir = 1 / r, twoh the summed softening.
__device__ __forceinline__ void spline_b_db(double r, double dir, double twoh, double& b,
                                            double& db) {
  if (r < twoh) {
    const double dih = 2.0 / twoh;
    const double u = r * dih;
    if (u < 1.0) {
      b = dih * dih * dih * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u);
      db = dih * dih * dih * dih * (-12.0 / 5.0 * u + 3.0 / 2.0 * u * u);
    } else {
      b = -1.0 / 15.0 * dir * dir * dir +
          dih * dih * dih * (8.0 / 3.0 - 3.0 * u + 6.0 / 5.0 * u * u - 1.0 / 6.0 * u * u * u);
      db = 1.0 / 5.0 * dir * dir * dir * dir +
           dih * dih * dih * dih * (-3.0 + 12.0 / 5.0 * u - 1.0 / 2.0 * u * u);
    }
  } else {
    const double a = dir;
    b = a * a * a;
    db = -3.0 * b * a;
  }
}

// GravityVisitor::SPLINEQ b, c, d (examples/GravityVisitor.h:37-78, the
// expressions as written there), their r-derivatives (br, cr, dr_) and their
// twoh-derivatives at fixed r (bt, ct, dt_).
struct SplineQ {
  double b, c, d, br, cr, dr_, bt, ct, dt_;
};

__device__ __forceinline__ SplineQ splineq_grad(double r, double dir, double r2, double twoh) {
  double b, c, d, br, cr, dr_, bt, ct, dt_;
  if (r2 < twoh * twoh) {
    const double dih = 2.0 / twoh;
    const double u = dih * r;
    const double dih2 = dih * dih, dih3 = dih2 * dih, dih4 = dih3 * dih, dih5 = dih4 * dih,
                 dih6 = dih5 * dih;
    double bD, cD, dD;  // derivatives with respect to dih at fixed r
    if (u < 1.0) {
      b = dih3 * 4.0 / 3.0 - 6.0 / 5.0 * u * u + 1.0 / 2.0 * u * u * u;
      c = dih5 * 12.0 / 5.0 - 3.0 / 2.0 * u;
      d = 3.0 / 2.0 * dih6 * dir;
      br = dih * (-12.0 / 5.0 * u + 3.0 / 2.0 * u * u);
      cr = -3.0 / 2.0 * dih;
      dr_ = -3.0 / 2.0 * dih6 * dir * dir;
      bD = 4.0 * dih2 + r * (-12.0 / 5.0 * u + 3.0 / 2.0 * u * u);
      cD = 12.0 * dih4 - 3.0 / 2.0 * r;
      dD = 9.0 * dih5 * dir;
    } else {
      const double dir2 = dir * dir, dir3 = dir2 * dir;
      b = -1.0 / 15.0 * dir3 + dih3 * 8.0 / 3.0 - 3.0 * u + 6.0 / 5.0 * u * u -
            1.0 / 6.0 * u * u * u;
      c = -1.0 / 5.0 * dir3 * dir2 + 3.0 * dih4 * dir - dih5 * 12.0 / 5.0 + 1.0 / 2.0 * u;
      d = -dir3 * dir3 * dir + 3.0 * dih4 * dir3 - 1.0 / 2.0 * dih6 * dir;
      br = 1.0 / 5.0 * dir2 * dir2 + dih * (-3.0 + 12.0 / 5.0 * u - 1.0 / 2.0 * u * u);
      cr = dir3 * dir3 - 3.0 * dih4 * dir2 + 1.0 / 2.0 * dih;
      dr_ = 7.0 * dir3 * dir3 * dir2 - 9.0 * dih4 * dir2 * dir2 + 1.0 / 2.0 * dih6 * dir2;
      bD = 8.0 * dih2 + r * (-3.0 + 12.0 / 5.0 * u - 1.0 / 2.0 * u * u);
      cD = 12.0 * dih3 * dir - 12.0 * dih4 + 1.0 / 2.0 * r;
      dD = 12.0 * dih3 * dir3 - 3.0 * dih5 * dir;
    }
    const double dd_dt = -0.5 * dih2;  // d(dih)/d(twoh)
    bt = bD * dd_dt;
    ct = cD * dd_dt;
    dt_ = dD * dd_dt;
  } else {
    const double a = dir;
    b = a * a * a;
    c = 3.0 * b * a * a;
    d = 5.0 * c * a * a;
    br = -3.0 * b * a;
    cr = -5.0 * c * a;
    dr_ = -7.0 * d * a;
    bt = ct = dt_ = 0.0;
  }
  // built once from scalars: field stores in the branches can be merged into
  // one store through a pointer select, which keeps the struct in scratch memory
  return SplineQ{b, c, d, br, cr, dr_, bt, ct, dt_};
}

struct TgtA {
  double x, y, z, e, wx, wy, wz;
  int32_t bt, pad;  // bt: the target's bucket within its group
};

// Squared distance from the box [lo, hi] to the point p.
__device__ __forceinline__ double box_dsq(const double lo[3], const double hi[3], double px,
                                          double py, double pz) {
  const double p[3] = {px, py, pz};
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double t = p[k] < lo[k] ? lo[k] - p[k] : (p[k] > hi[k] ? p[k] - hi[k] : 0.0);
    d += t * t;
  }
  return d;
}

// The far classification of a pass: a closed node whose centre lies farther
// from the box around the group's buckets than twice its softening plus the
// largest bucket softening fails both openSoftening tests for every member
// (each member's box and centre lie inside that box), and one farther than
// its softening plus the largest target softening is beyond the softened
// range of every pair; a source particle likewise beyond its softening plus
// the largest target softening, and a leaf when its centre is farther than
// its reach plus those softenings.  Margins of 2e-4 on the reaches and 1e-4
// on the squares also cover the FP32 form's shifted coordinates.
// Per wave, in LDS: lo[3], hi[3] of the box around the group's buckets, the
// largest bucket softening, the largest target softening of the pass.
constexpr int kPassFar = 8;

__device__ __forceinline__ bool node_far_pass(const double* __restrict__ f, double cx, double cy,
                                              double cz, double soft) {
  const double dsq = box_dsq(f, f + 3, cx, cy, cz);
  const double mono = 2.0 * (soft + f[6]);
  const double reach = (soft + f[7]) * 1.0002;
  const double m = mono > reach ? mono : reach;
  return dsq > 1.0001 * m * m;
}

__device__ __forceinline__ bool leaf_far_pass(const double* __restrict__ f, double cx, double cy,
                                              double cz, double lreach, double lsoft) {
  const double dsq = box_dsq(f, f + 3, cx, cy, cz);
  const double reach = (lreach + lsoft + f[7]) * 1.0002;
  return dsq > 1.0001 * reach * reach;
}

// One node() list entry against the pass's targets (FP64): the target
// gradients into tg, the node cotangent into lam; mono: the members for which
// openSoftening sends the node to addGravity.
__device__ __forceinline__ void node_entry_pairs(const GravNode& s, unsigned mask, unsigned mono,
                                                 const TgtA* __restrict__ st, int ng, double ox,
                                                 double oy, double oz, double (&tg)[kANT][3],
                                                 double (&lam)[kNodeAdj]) {
  const double cx = s.cx + ox, cy = s.cy + oy, cz = s.cz + oz;
#pragma unroll
  for (int t = 0; t < kANT; t++) {
    if (t < ng && ((mask >> st[t].bt) & 1u)) {
      const TgtA tt = st[t];
      if ((mono >> tt.bt) & 1u) {
        const double dx = cx - tt.x, dy = cy - tt.y, dz = cz - tt.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 != 0) {
          const double ir = pt_rsqrt64(r2);
          const double ir3 = ir * ir * ir;
          const double dw = dx * tt.wx + dy * tt.wy + dz * tt.wz;
          const double f = -3.0 * s.mass * ir3 * ir * ir * dw;
          const double Tx = s.mass * ir3 * tt.wx + f * dx;
          const double Ty = s.mass * ir3 * tt.wy + f * dy;
          const double Tz = s.mass * ir3 * tt.wz + f * dz;
          tg[t][0] -= Tx;
          tg[t][1] -= Ty;
          tg[t][2] -= Tz;
          lam[0] += Tx;
          lam[1] += Ty;
          lam[2] += Tz;
          lam[3] += dw * ir3;
        }
      } else {
        const double rx = (tt.x - s.cx) - ox, ry = (tt.y - s.cy) - oy, rz = (tt.z - s.cz) - oz;
        const double r2 = fma(rx, rx, fma(ry, ry, rz * rz));
        const double ir = pt_rsqrt64(r2);
        const double r = r2 * ir;
        const SplineQ q = splineq_grad(r, ir, r2, s.soft + tt.e);
        const double qx = fma(s.qxx, rx, fma(s.qxy, ry, s.qxz * rz));
        const double qy = fma(s.qxy, rx, fma(s.qyy, ry, s.qyz * rz));
        const double qz = fma(s.qxz, rx, fma(s.qyz, ry, s.qzz * rz));
        const double qir = 0.5 * fma(qx, rx, fma(qy, ry, qz * rz));
        const double qir3 = fma(q.b, s.mass, fma(q.d, qir, -q.c * s.tr));
        const double rw = fma(rx, tt.wx, fma(ry, tt.wy, rz * tt.wz));
        const double qw = fma(qx, tt.wx, fma(qy, tt.wy, qz * tt.wz));
        const double Qwx = fma(s.qxx, tt.wx, fma(s.qxy, tt.wy, s.qxz * tt.wz));
        const double Qwy = fma(s.qxy, tt.wx, fma(s.qyy, tt.wy, s.qyz * tt.wz));
        const double Qwz = fma(s.qxz, tt.wx, fma(s.qyz, tt.wy, s.qzz * tt.wz));
        const double k1 = fma(s.mass, q.br, fma(qir, q.dr_, -s.tr * q.cr)) * ir;
        const double k2 = q.cr * ir;
        // J = -qir3 w + (k2 (q.w) - k1 (r.w)) r - d (r.w) q + c Q w
        const double s1 = fma(k2, qw, -k1 * rw);
        const double drw = q.d * rw;
        const double Jx = fma(-qir3, tt.wx, fma(s1, rx, fma(-drw, qx, q.c * Qwx)));
        const double Jy = fma(-qir3, tt.wy, fma(s1, ry, fma(-drw, qy, q.c * Qwy)));
        const double Jz = fma(-qir3, tt.wz, fma(s1, rz, fma(-drw, qz, q.c * Qwz)));
        tg[t][0] += Jx;
        tg[t][1] += Jy;
        tg[t][2] += Jz;
        lam[0] -= Jx;
        lam[1] -= Jy;
        lam[2] -= Jz;
        lam[3] = fma(-q.b, rw, lam[3]);
        // d(a.w)/dQ per stored component: (-d (r.w) / 2) r r^T + (c (r.w) / 2) I
        // + c sym(w r^T), written as r_k p_l + r_l p_k (+ the trace part) with
        // p = -d (r.w) r / 2 + c w
        const double hd = -0.5 * drw, hc = 0.5 * q.c * rw;
        const double qpx = fma(hd, rx, q.c * tt.wx), qpy = fma(hd, ry, q.c * tt.wy),
                     qpz = fma(hd, rz, q.c * tt.wz);
        lam[4] += fma(rx, qpx, hc);
        lam[5] += fma(ry, qpx, rx * qpy);
        lam[6] += fma(rz, qpx, rx * qpz);
        lam[7] += fma(ry, qpy, hc);
        lam[8] += fma(rz, qpy, ry * qpz);
        lam[9] += fma(rz, qpz, hc);
        lam[10] += fma(-fma(q.bt, s.mass, fma(q.dt_, qir, -q.ct * s.tr)), rw, q.ct * qw);
      }
    }
  }
}

// One source particle against the pass's targets (FP64), the softened form.
__device__ __forceinline__ void particle_pairs(double jx, double jy, double jz, double jm,
                                               double je, unsigned jmask,
                                               const TgtA* __restrict__ st, int ng,
                                               double (&tg)[kANT][3], double& sgx, double& sgy,
                                               double& sgz, double& sgm) {
#pragma unroll
  for (int t = 0; t < kANT; t++) {
    if (t < ng && ((jmask >> st[t].bt) & 1u)) {
      const TgtA tt = st[t];
      const double dx = jx - tt.x, dy = jy - tt.y, dz = jz - tt.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 != 0) {
        const double ir = pt_rsqrt64(r2);
        double bb, db;
        spline_b_db(r2 * ir, ir, je + tt.e, bb, db);
        const double dw = dx * tt.wx + dy * tt.wy + dz * tt.wz;
        const double f = db * ir * dw;
        const double Kx = jm * (bb * tt.wx + f * dx);
        const double Ky = jm * (bb * tt.wy + f * dy);
        const double Kz = jm * (bb * tt.wz + f * dz);
        tg[t][0] -= Kx;
        tg[t][1] -= Ky;
        tg[t][2] -= Kz;
        sgx += Kx;
        sgy += Ky;
        sgz += Kz;
        sgm += bb * dw;
      }
    }
  }
}

// One source particle of a far window (every pair beyond its softening, so
// not the target itself) against all kANT targets of the pass (FP64), without
// branches: a pair whose target bucket is not in the mask, and the padding
// targets past the pass's count (bucket index 31, the position of a real
// target, so r > 0), is evaluated with a zero cotangent w, which zeroes each
// of its contributions (all are linear in w); consecutive targets' pairs then
// interleave.  The source's position gradient is summed before its mass
// factor.
__device__ __forceinline__ void particle_pairs_far(double jx, double jy, double jz, double jm,
                                                   unsigned jmask, const TgtA* __restrict__ st,
                                                   double (&tg)[kANT][3], double& sgx,
                                                   double& sgy, double& sgz, double& sgm) {
  double kx = 0.0, ky = 0.0, kz = 0.0;
#pragma unroll
  for (int t = 0; t < kANT; t++) {
    const TgtA tt = st[t];
    const bool act = (jmask >> tt.bt) & 1u;
    const double wx = act ? tt.wx : 0.0, wy = act ? tt.wy : 0.0, wz = act ? tt.wz : 0.0;
    const double dx = jx - tt.x, dy = jy - tt.y, dz = jz - tt.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir = pt_rsqrt64(r2);
    const double a2 = ir * ir;
    const double bb = ir * a2;
    const double dw = fma(dx, wx, fma(dy, wy, dz * wz));
    const double f = -3.0 * (bb * a2) * dw;  // b'(r) (d.w) / r
    const double Kx = fma(bb, wx, f * dx);
    const double Ky = fma(bb, wy, f * dy);
    const double Kz = fma(bb, wz, f * dz);
    tg[t][0] = fma(-jm, Kx, tg[t][0]);
    tg[t][1] = fma(-jm, Ky, tg[t][1]);
    tg[t][2] = fma(-jm, Kz, tg[t][2]);
    kx += Kx;
    ky += Ky;
    kz += Kz;
    sgm = fma(bb, dw, sgm);
  }
  sgx += jm * kx;
  sgy += jm * ky;
  sgz += jm * kz;
}

// One wave per bucket group (consecutive buckets, at most PT_GROUP_TARGETS
// particles unless one bucket is larger; neighbouring buckets share most of
// their sources, so a group's union list is far shorter than the sum of its
// buckets' lists).  Node sources: one list entry per lane, looping over the
// group's targets (PT_GROUP_TARGETS per pass) from LDS and evaluating the
// pairs whose target bucket is in the entry's member mask; the node cotangent
// is summed over the targets of a pass before one atomic add per field.  Leaf
// sources: flattened to one source particle per lane (the leaf's mask
// carried along); the source's gradient is summed over the targets of a pass
// before its atomic adds.  Target gradients are wave-reduced per pass.
__global__ void __launch_bounds__(kAdjThreads)
adjoint_group_kernel(const Bucket* __restrict__ bk, const int32_t* __restrict__ gfirst,
                     const int32_t* __restrict__ gnb, int64_t ns,
                     const TreeNode* __restrict__ nodes,
                    const GravNode* __restrict__ gn, const double* __restrict__ px,
                    const double* __restrict__ py, const double* __restrict__ pz,
                    const double* __restrict__ pm, const double* __restrict__ pe,
                    const int64_t* __restrict__ pp_off, const int32_t* __restrict__ pp,
                    const uint8_t* __restrict__ ppm, const int64_t* __restrict__ pc_off,
                    const int32_t* __restrict__ pc, const uint8_t* __restrict__ pcm,
                    double ox, double oy, double oz, const double* __restrict__ wx,
                    const double* __restrict__ wy, const double* __restrict__ wz,
                    double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
                    double* __restrict__ gmass, double* __restrict__ lam_node, int64_t nn,
                    double* __restrict__ hot, int64_t nhot) {
  __shared__ TgtA s_t[kAdjWaves][kANT];
  __shared__ Bucket s_bk[kAdjWaves][PT_GROUP_BUCKETS];
  __shared__ int32_t s_beg[kAdjWaves][PT_WAVE];
  __shared__ double s_far[kAdjWaves][kPassFar];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t gi = (int64_t)blockIdx.x * kAdjWaves + wave;
  if (gi >= ns) return;
  const int nbk = gnb[gi];
  if (nbk <= 0) return;
  const int64_t b0 = gfirst[gi];
  if (lane < nbk) s_bk[wave][lane] = bk[b0 + lane];
  __builtin_amdgcn_wave_barrier();
  // the group's targets are consecutive: its buckets' ranges in order
  const int64_t first = s_bk[wave][0].begin;
  const int nt = (int)((int64_t)s_bk[wave][nbk - 1].begin + s_bk[wave][nbk - 1].count - first);
  const int64_t pc0 = pc_off[gi], pc1 = pc_off[gi + 1];
  const int64_t pp0 = pp_off[gi], pp1 = pp_off[gi + 1];
  if (lane < 7) {
    // the group's box and largest bucket softening
    double v = lane < 3 ? INFINITY : (lane < 6 ? -INFINITY : 0.0);
    for (int k = 0; k < nbk; k++) {
      const Bucket& bb = s_bk[wave][k];
      const double c[7] = {bb.lox, bb.loy, bb.loz, bb.hix, bb.hiy, bb.hiz, bb.soft};
      v = lane < 3 ? fmin(v, c[lane]) : fmax(v, c[lane]);
    }
    s_far[wave][lane] = v;
  }
  for (int g0 = 0; g0 < nt; g0 += kANT) {
    const int ng = (nt - g0) < kANT ? (nt - g0) : kANT;
    // member buckets with a target in this pass
    unsigned pass = 0;
    for (int k = 0; k < nbk; k++) {
      const int64_t lo = s_bk[wave][k].begin, hi = lo + s_bk[wave][k].count;
      if (hi > first + g0 && lo < first + g0 + ng) pass |= 1u << k;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < ng) {
      const int64_t i = first + g0 + lane;
      int bt = 0;
      for (int k = 1; k < nbk; k++)
        if (i >= s_bk[wave][k].begin) bt = k;
      s_t[wave][lane] = TgtA{px[i], py[i], pz[i], pe[i], wx[i], wy[i], wz[i], bt, 0};
    } else if (lane < kANT) {
      // padding for the branch-free far-window form: a real target's
      // position, no bucket, zero cotangent
      const int64_t i = first + g0;
      s_t[wave][lane] = TgtA{px[i], py[i], pz[i], 0.0, 0.0, 0.0, 0.0, 31, 0};
    }
    __builtin_amdgcn_wave_barrier();
    double tg[kANT][3];
#pragma unroll
    for (int t = 0; t < kANT; t++) tg[t][0] = tg[t][1] = tg[t][2] = 0.0;
    if (lane == 0) {
      double emax = 0.0;
      for (int t = 0; t < ng; t++) emax = s_t[wave][t].e > emax ? s_t[wave][t].e : emax;
      s_far[wave][7] = emax;
    }
    __builtin_amdgcn_wave_barrier();

    // ---- node() sources: in chunks in which every entry is far no member's
    // openSoftening can hold, so its tests are skipped
    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int64_t e = e0 + lane;
      const unsigned mask = e < pc1 ? (unsigned)pcm[e] & pass : 0u;
      int32_t sid = 0;
      if (mask != 0u) {
        sid = pc[e];
        const GravNode s = gn[sid];
        const bool far = node_far_pass(s_far[wave], s.cx + ox, s.cy + oy, s.cz + oz, s.soft);
        // over the lanes holding an entry (the others are inactive here)
        const bool all_far = __ballot(!far) == 0ull;
        double lam[kNodeAdj];
#pragma unroll
        for (int k = 0; k < kNodeAdj; k++) lam[k] = 0.0;
        // members for which openSoftening sends it to addGravity (none when
        // every entry of the chunk is far)
        unsigned mono = 0;
        if (!all_far) {
          for (unsigned mm = mask; mm != 0u; mm &= mm - 1u) {
            const int k = __builtin_ctz(mm);
            if (pt_open_softening(s.cx + ox, s.cy + oy, s.cz + oz, s.soft, s_bk[wave][k]))
              mono |= 1u << k;
          }
        }
        node_entry_pairs(s, mask, mono, s_t[wave], ng, ox, oy, oz, tg, lam);
        double* dst = sid < nhot
                          ? hot + (int64_t)(blockIdx.x % kHotParts) * kNodeAdj * nhot + sid
                          : lam_node + sid;
        const int64_t stride = sid < nhot ? nhot : nn;
#pragma unroll
        for (int k = 0; k < kNodeAdj; k++)
          if (lam[k] != 0.0) atomicAdd(dst + (int64_t)k * stride, lam[k]);
      }
    }

    // ---- leaf() sources, flattened to particles, the leaves that may hold a
    // softened pair first: windows past them take the branch-free form
    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      const int64_t l = l0 + lane;
      int32_t cnt = 0, beg = 0;
      bool isnear = false;
      const unsigned msk = l < pp1 ? (unsigned)ppm[l] & pass : 0u;
      if (msk != 0u) {
        const TreeNode nd = nodes[pp[l]];
        cnt = nd.count;
        beg = nd.begin;
        isnear = !leaf_far_pass(s_far[wave], nd.cx + ox, nd.cy + oy, nd.cz + oz, nd.leaf_reach,
                                nd.leaf_soft);
      }
      const int cn = isnear ? cnt : 0, cf = isnear ? 0 : cnt;
      const int incl_n = pt_wave_incl_scan(cn);
      const int incl_f = pt_wave_incl_scan(cf);
      const int near_total = __builtin_amdgcn_readlane(incl_n, PT_WAVE - 1);
      const int total = near_total + __builtin_amdgcn_readlane(incl_f, PT_WAVE - 1);
      // each entry's first slot, near entries first
      const int exc = isnear ? incl_n - cn : near_total + incl_f - cf;
      int carry = 0;
      for (int s0 = 0; s0 < total; s0 += PT_WAVE) {
        const bool wfar = s0 >= near_total;
        // a slot's entry: the lane of the last entry start at or below it in
        // this window (marks by start slot, a ballot of the marks), or the
        // previous window's last entry
        int32_t* mark = s_beg[wave];
        mark[lane] = -1;
        __builtin_amdgcn_wave_barrier();
        if (cnt > 0 && exc >= s0 && exc < s0 + PT_WAVE) mark[exc - s0] = lane;
        __builtin_amdgcn_wave_barrier();
        const unsigned long long starts = __ballot(mark[lane] >= 0);
        const unsigned long long upto = starts & (~0ull >> (PT_WAVE - 1 - lane));
        const int leaf = upto ? mark[PT_WAVE - 1 - __clzll(upto)] : carry;
        carry = __builtin_amdgcn_readlane(leaf, PT_WAVE - 1);
        const int lbeg = __shfl(beg, leaf, PT_WAVE);
        const int lexc = __shfl(exc, leaf, PT_WAVE);
        const unsigned jmask = (unsigned)__shfl((int)msk, leaf, PT_WAVE);
        const int slot = s0 + lane;
        if (slot < total) {
          const int64_t j = (int64_t)lbeg + (slot - lexc);
          const double jx = px[j] + ox, jy = py[j] + oy, jz = pz[j] + oz;
          const double jm = pm[j], je = pe[j];
          double sgx = 0.0, sgy = 0.0, sgz = 0.0, sgm = 0.0;
          if (wfar)
            particle_pairs_far(jx, jy, jz, jm, jmask, s_t[wave], tg, sgx, sgy, sgz, sgm);
          else
            particle_pairs(jx, jy, jz, jm, je, jmask, s_t[wave], ng, tg, sgx, sgy, sgz, sgm);
          atomicAdd(&gx[j], sgx);
          atomicAdd(&gy[j], sgy);
          atomicAdd(&gz[j], sgz);
          atomicAdd(&gmass[j], sgm);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }

#pragma unroll
    for (int t = 0; t < kANT; t++) {
      if (t < ng) {
        const double rx = pt_wave_sum(tg[t][0]);
        const double ry = pt_wave_sum(tg[t][1]);
        const double rz = pt_wave_sum(tg[t][2]);
        if (lane == t) {
          const int64_t i = first + g0 + t;
          atomicAdd(&gx[i], rx);
          atomicAdd(&gy[i], ry);
          atomicAdd(&gz[i], rz);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 form: the same pair expressions in FP32 on coordinates shifted to the
// centre of the group's box (in FP64, then rounded), with the hardware reciprocal
// and reciprocal square root; every sum that spans pairs (target gradients,
// node and source-particle cotangents) is carried in FP64.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// spline_b_db in FP32; ir = 1 / r.
__device__ __forceinline__ void spline_b_db32(float r, float ir, float twoh, float& b, float& db) {
  if (r < twoh) {
    const float dih = 2.0f * rcp32(twoh);
    const float u = r * dih;
    const float dih3 = dih * dih * dih;
    if (u < 1.0f) {
      b = dih3 * (4.0f / 3.0f - 6.0f / 5.0f * u * u + 1.0f / 2.0f * u * u * u);
      db = dih3 * dih * (-12.0f / 5.0f * u + 3.0f / 2.0f * u * u);
    } else {
      b = -1.0f / 15.0f * ir * ir * ir +
          dih3 * (8.0f / 3.0f - 3.0f * u + 6.0f / 5.0f * u * u - 1.0f / 6.0f * u * u * u);
      db = 1.0f / 5.0f * ir * ir * ir * ir +
           dih3 * dih * (-3.0f + 12.0f / 5.0f * u - 1.0f / 2.0f * u * u);
    }
  } else {
    b = ir * ir * ir;
    db = -3.0f * b * ir;
  }
}

struct SplineQ32 {
  float b, c, d, br, cr, dr_, bt, ct, dt_;
};

// splineq_grad in FP32; ir = 1 / r.
__device__ __forceinline__ SplineQ32 splineq_grad32(float r, float ir, float r2, float twoh) {
  float b, c, d, br, cr, dr_, bt, ct, dt_;
  const float dir = ir;
  if (r2 < twoh * twoh) {
    const float dih = 2.0f * rcp32(twoh);
    const float u = dih * r;
    const float dih2 = dih * dih, dih3 = dih2 * dih, dih4 = dih3 * dih, dih5 = dih4 * dih,
                dih6 = dih5 * dih;
    float bD, cD, dD;
    if (u < 1.0f) {
      b = dih3 * 4.0f / 3.0f - 6.0f / 5.0f * u * u + 1.0f / 2.0f * u * u * u;
      c = dih5 * 12.0f / 5.0f - 3.0f / 2.0f * u;
      d = 3.0f / 2.0f * dih6 * dir;
      br = dih * (-12.0f / 5.0f * u + 3.0f / 2.0f * u * u);
      cr = -3.0f / 2.0f * dih;
      dr_ = -3.0f / 2.0f * dih6 * dir * dir;
      bD = 4.0f * dih2 + r * (-12.0f / 5.0f * u + 3.0f / 2.0f * u * u);
      cD = 12.0f * dih4 - 3.0f / 2.0f * r;
      dD = 9.0f * dih5 * dir;
    } else {
      const float dir2 = dir * dir, dir3 = dir2 * dir;
      b = -1.0f / 15.0f * dir3 + dih3 * 8.0f / 3.0f - 3.0f * u + 6.0f / 5.0f * u * u -
            1.0f / 6.0f * u * u * u;
      c = -1.0f / 5.0f * dir3 * dir2 + 3.0f * dih4 * dir - dih5 * 12.0f / 5.0f + 1.0f / 2.0f * u;
      d = -dir3 * dir3 * dir + 3.0f * dih4 * dir3 - 1.0f / 2.0f * dih6 * dir;
      br = 1.0f / 5.0f * dir2 * dir2 + dih * (-3.0f + 12.0f / 5.0f * u - 1.0f / 2.0f * u * u);
      cr = dir3 * dir3 - 3.0f * dih4 * dir2 + 1.0f / 2.0f * dih;
      dr_ = 7.0f * dir3 * dir3 * dir2 - 9.0f * dih4 * dir2 * dir2 + 1.0f / 2.0f * dih6 * dir2;
      bD = 8.0f * dih2 + r * (-3.0f + 12.0f / 5.0f * u - 1.0f / 2.0f * u * u);
      cD = 12.0f * dih3 * dir - 12.0f * dih4 + 1.0f / 2.0f * r;
      dD = 12.0f * dih3 * dir3 - 3.0f * dih5 * dir;
    }
    const float dd_dt = -0.5f * dih2;
    bt = bD * dd_dt;
    ct = cD * dd_dt;
    dt_ = dD * dd_dt;
  } else {
    const float a = dir;
    b = a * a * a;
    c = 3.0f * b * a * a;
    d = 5.0f * c * a * a;
    br = -3.0f * b * a;
    cr = -5.0f * c * a;
    dr_ = -7.0f * d * a;
    bt = ct = dt_ = 0.0f;
  }
  // built once from scalars: field stores in the branches can be merged into
  // one store through a pointer select, which keeps the struct in scratch memory
  return SplineQ32{b, c, d, br, cr, dr_, bt, ct, dt_};
}

struct TgtA32 {
  float x, y, z, e, wx, wy, wz;
  int32_t bt;
};

// One source particle against the pass's targets (FP32 pairs, target
// gradients summed in FP64).  far (wave-uniform): every pair is beyond its
// softening (and so not the target itself).
__device__ __forceinline__ void particle_pairs32(bool far, float jx, float jy, float jz, float jm,
                                                 float je, unsigned jmask,
                                                 const TgtA32* __restrict__ st, int ng,
                                                 double (&tg)[kANT][3], float& sgx, float& sgy,
                                                 float& sgz, float& sgm) {
#pragma unroll
  for (int t = 0; t < kANT; t++) {
    if (t < ng && ((jmask >> st[t].bt) & 1u)) {
      const TgtA32 tt = st[t];
      const float dx = jx - tt.x, dy = jy - tt.y, dz = jz - tt.z;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      if (far || r2 != 0.0f) {
        const float ir = __builtin_amdgcn_rsqf(r2);
        float bb, db;
        if (far) {
          bb = ir * ir * ir;
          db = -3.0f * bb * ir;
        } else {
          spline_b_db32(r2 * ir, ir, je + tt.e, bb, db);
        }
        const float dw = fmaf(dx, tt.wx, fmaf(dy, tt.wy, dz * tt.wz));
        const float f = db * ir * dw;
        const float Kx = jm * fmaf(bb, tt.wx, f * dx);
        const float Ky = jm * fmaf(bb, tt.wy, f * dy);
        const float Kz = jm * fmaf(bb, tt.wz, f * dz);
        tg[t][0] -= (double)Kx;
        tg[t][1] -= (double)Ky;
        tg[t][2] -= (double)Kz;
        sgx += Kx;
        sgy += Ky;
        sgz += Kz;
        sgm += bb * dw;
      }
    }
  }
}

// adjoint_group_kernel in the FP32 form (same work mapping; node moments from
// the FP32 copies, node centres and softening decisions from the FP64 fields
// of the same records).
__global__ void __launch_bounds__(kAdjThreads)
adjoint_group32_kernel(const Bucket* __restrict__ bk, const int32_t* __restrict__ gfirst,
                       const int32_t* __restrict__ gnb, int64_t ns,
                      const TreeNode* __restrict__ nodes, const GravNode32* __restrict__ gn32,
                      const double* __restrict__ px,
                      const double* __restrict__ py, const double* __restrict__ pz,
                      const double* __restrict__ pm, const double* __restrict__ pe,
                      const int64_t* __restrict__ pp_off, const int32_t* __restrict__ pp,
                      const uint8_t* __restrict__ ppm, const int64_t* __restrict__ pc_off,
                      const int32_t* __restrict__ pc, const uint8_t* __restrict__ pcm,
                      double ox, double oy, double oz, const double* __restrict__ wx,
                      const double* __restrict__ wy, const double* __restrict__ wz,
                      double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
                      double* __restrict__ gmass, double* __restrict__ lam_node, int64_t nn,
                      double* __restrict__ hot, int64_t nhot) {
  __shared__ TgtA32 s_t[kAdjWaves][kANT];
  __shared__ Bucket s_bk[kAdjWaves][PT_GROUP_BUCKETS];
  __shared__ int32_t s_beg[kAdjWaves][PT_WAVE];
  __shared__ double s_far[kAdjWaves][kPassFar];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PT_WAVE);
  const int lane = threadIdx.x & (PT_WAVE - 1);
  const int64_t gi = (int64_t)blockIdx.x * kAdjWaves + wave;
  if (gi >= ns) return;
  const int nbk = gnb[gi];
  if (nbk <= 0) return;
  const int64_t b0 = gfirst[gi];
  if (lane < nbk) s_bk[wave][lane] = bk[b0 + lane];
  __builtin_amdgcn_wave_barrier();
  // the group's targets are consecutive: its buckets' ranges in order
  const int64_t first = s_bk[wave][0].begin;
  const int nt = (int)((int64_t)s_bk[wave][nbk - 1].begin + s_bk[wave][nbk - 1].count - first);
  // FP32 frame: the centre of the box around the group's bucket boxes
  double lx = s_bk[wave][0].lox, ly = s_bk[wave][0].loy, lz = s_bk[wave][0].loz;
  double hx = s_bk[wave][0].hix, hy = s_bk[wave][0].hiy, hz = s_bk[wave][0].hiz;
  for (int k = 1; k < nbk; k++) {
    lx = fmin(lx, s_bk[wave][k].lox);
    ly = fmin(ly, s_bk[wave][k].loy);
    lz = fmin(lz, s_bk[wave][k].loz);
    hx = fmax(hx, s_bk[wave][k].hix);
    hy = fmax(hy, s_bk[wave][k].hiy);
    hz = fmax(hz, s_bk[wave][k].hiz);
  }
  const double ccx = (lx + hx) * 0.5, ccy = (ly + hy) * 0.5, ccz = (lz + hz) * 0.5;
  if (lane < 7) {
    // the group's box and largest bucket softening
    double v = lane < 3 ? INFINITY : (lane < 6 ? -INFINITY : 0.0);
    for (int k = 0; k < nbk; k++) {
      const Bucket& bb = s_bk[wave][k];
      const double c[7] = {bb.lox, bb.loy, bb.loz, bb.hix, bb.hiy, bb.hiz, bb.soft};
      v = lane < 3 ? fmin(v, c[lane]) : fmax(v, c[lane]);
    }
    s_far[wave][lane] = v;
  }
  const int64_t pc0 = pc_off[gi], pc1 = pc_off[gi + 1];
  const int64_t pp0 = pp_off[gi], pp1 = pp_off[gi + 1];
  for (int g0 = 0; g0 < nt; g0 += kANT) {
    const int ng = (nt - g0) < kANT ? (nt - g0) : kANT;
    // member buckets with a target in this pass
    unsigned pass = 0;
    for (int k = 0; k < nbk; k++) {
      const int64_t lo = s_bk[wave][k].begin, hi = lo + s_bk[wave][k].count;
      if (hi > first + g0 && lo < first + g0 + ng) pass |= 1u << k;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < ng) {
      const int64_t i = first + g0 + lane;
      int bt = 0;
      for (int k = 1; k < nbk; k++)
        if (i >= s_bk[wave][k].begin) bt = k;
      s_t[wave][lane] = TgtA32{(float)(px[i] - ccx), (float)(py[i] - ccy), (float)(pz[i] - ccz),
                               (float)pe[i], (float)wx[i], (float)wy[i], (float)wz[i], bt};
    }
    __builtin_amdgcn_wave_barrier();
    double tg[kANT][3];
#pragma unroll
    for (int t = 0; t < kANT; t++) tg[t][0] = tg[t][1] = tg[t][2] = 0.0;
    if (lane == 0) {
      double emax = 0.0;
      for (int t = 0; t < ng; t++) emax = fmax(emax, (double)s_t[wave][t].e);
      s_far[wave][7] = emax;
    }
    __builtin_amdgcn_wave_barrier();

    // ---- node() sources
    for (int64_t e0 = pc0; e0 < pc1; e0 += PT_WAVE) {
      const int64_t e = e0 + lane;
      const unsigned mask = e < pc1 ? (unsigned)pcm[e] & pass : 0u;
      if (mask != 0u) {
        const int32_t sid = pc[e];
        const GravNode32 s = gn32[sid];
        const double cxd = s.cx + ox, cyd = s.cy + oy, czd = s.cz + oz;
        // over the lanes holding an entry (the others are inactive here)
        const bool all_far = __ballot(!node_far_pass(s_far[wave], cxd, cyd, czd, s.soft64)) == 0ull;
        // members for which openSoftening sends it to addGravity (none when
        // every entry of the chunk is far)
        unsigned mono = 0;
        if (!all_far) {
          for (unsigned mm = mask; mm != 0u; mm &= mm - 1u) {
            const int k = __builtin_ctz(mm);
            if (pt_open_softening(cxd, cyd, czd, s.soft64, s_bk[wave][k])) mono |= 1u << k;
          }
        }
        const float cx = (float)(cxd - ccx), cy = (float)(cyd - ccy), cz = (float)(czd - ccz);
        const float tr = 0.5f * (s.qxx + s.qyy + s.qzz);
        float lam[kNodeAdj];
#pragma unroll
        for (int k = 0; k < kNodeAdj; k++) lam[k] = 0.0f;
#pragma unroll
        for (int t = 0; t < kANT; t++) {
          if (t < ng && ((mask >> s_t[wave][t].bt) & 1u)) {
            const TgtA32 tt = s_t[wave][t];
            if ((mono >> tt.bt) & 1u) {
              const float dx = cx - tt.x, dy = cy - tt.y, dz = cz - tt.z;
              const float r2 = dx * dx + dy * dy + dz * dz;
              if (r2 != 0.0f) {
                const float ir = __builtin_amdgcn_rsqf(r2);
                const float ir3 = ir * ir * ir;
                const float dw = dx * tt.wx + dy * tt.wy + dz * tt.wz;
                const float f = -3.0f * s.mass * ir3 * ir * ir * dw;
                const float Tx = s.mass * ir3 * tt.wx + f * dx;
                const float Ty = s.mass * ir3 * tt.wy + f * dy;
                const float Tz = s.mass * ir3 * tt.wz + f * dz;
                tg[t][0] -= (double)Tx;
                tg[t][1] -= (double)Ty;
                tg[t][2] -= (double)Tz;
                lam[0] += Tx;
                lam[1] += Ty;
                lam[2] += Tz;
                lam[3] += dw * ir3;
              }
            } else {
              const float rx = tt.x - cx, ry = tt.y - cy, rz = tt.z - cz;
              const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
              const float ir = __builtin_amdgcn_rsqf(r2);
              const float r = r2 * ir;
              const SplineQ32 q = splineq_grad32(r, ir, r2, s.soft + tt.e);
              const float qx = fmaf(s.qxx, rx, fmaf(s.qxy, ry, s.qxz * rz));
              const float qy = fmaf(s.qxy, rx, fmaf(s.qyy, ry, s.qyz * rz));
              const float qz = fmaf(s.qxz, rx, fmaf(s.qyz, ry, s.qzz * rz));
              const float qir = 0.5f * fmaf(qx, rx, fmaf(qy, ry, qz * rz));
              const float qir3 = fmaf(q.b, s.mass, fmaf(q.d, qir, -q.c * tr));
              const float rw = fmaf(rx, tt.wx, fmaf(ry, tt.wy, rz * tt.wz));
              const float qw = fmaf(qx, tt.wx, fmaf(qy, tt.wy, qz * tt.wz));
              const float Qwx = fmaf(s.qxx, tt.wx, fmaf(s.qxy, tt.wy, s.qxz * tt.wz));
              const float Qwy = fmaf(s.qxy, tt.wx, fmaf(s.qyy, tt.wy, s.qyz * tt.wz));
              const float Qwz = fmaf(s.qxz, tt.wx, fmaf(s.qyz, tt.wy, s.qzz * tt.wz));
              const float k1 = fmaf(s.mass, q.br, fmaf(qir, q.dr_, -tr * q.cr)) * ir;
              const float k2 = q.cr * ir;
              // J = -qir3 w + (k2 (q.w) - k1 (r.w)) r - d (r.w) q + c Q w
              const float s1 = fmaf(k2, qw, -k1 * rw);
              const float drw = q.d * rw;
              const float Jx = fmaf(-qir3, tt.wx, fmaf(s1, rx, fmaf(-drw, qx, q.c * Qwx)));
              const float Jy = fmaf(-qir3, tt.wy, fmaf(s1, ry, fmaf(-drw, qy, q.c * Qwy)));
              const float Jz = fmaf(-qir3, tt.wz, fmaf(s1, rz, fmaf(-drw, qz, q.c * Qwz)));
              tg[t][0] += (double)Jx;
              tg[t][1] += (double)Jy;
              tg[t][2] += (double)Jz;
              lam[0] -= Jx;
              lam[1] -= Jy;
              lam[2] -= Jz;
              lam[3] = fmaf(-q.b, rw, lam[3]);
              // d(a.w)/dQ as r_k p_l + r_l p_k (+ the trace part), p = -d (r.w) r / 2 + c w
              const float hd = -0.5f * drw, hc = 0.5f * q.c * rw;
              const float qpx = fmaf(hd, rx, q.c * tt.wx), qpy = fmaf(hd, ry, q.c * tt.wy),
                          qpz = fmaf(hd, rz, q.c * tt.wz);
              lam[4] += fmaf(rx, qpx, hc);
              lam[5] += fmaf(ry, qpx, rx * qpy);
              lam[6] += fmaf(rz, qpx, rx * qpz);
              lam[7] += fmaf(ry, qpy, hc);
              lam[8] += fmaf(rz, qpy, ry * qpz);
              lam[9] += fmaf(rz, qpz, hc);
              lam[10] += fmaf(-fmaf(q.bt, s.mass, fmaf(q.dt_, qir, -q.ct * tr)), rw, q.ct * qw);
            }
          }
        }
        double* dst = sid < nhot
                          ? hot + (int64_t)(blockIdx.x % kHotParts) * kNodeAdj * nhot + sid
                          : lam_node + sid;
        const int64_t stride = sid < nhot ? nhot : nn;
#pragma unroll
        for (int k = 0; k < kNodeAdj; k++)
          if (lam[k] != 0.0f) atomicAdd(dst + (int64_t)k * stride, (double)lam[k]);
      }
    }

    // ---- leaf() sources, flattened to particles, the leaves that may hold a
    // softened pair first: windows past them take the branch-free form
    for (int64_t l0 = pp0; l0 < pp1; l0 += PT_WAVE) {
      const int64_t l = l0 + lane;
      int32_t cnt = 0, beg = 0;
      bool isnear = false;
      const unsigned msk = l < pp1 ? (unsigned)ppm[l] & pass : 0u;
      if (msk != 0u) {
        const TreeNode nd = nodes[pp[l]];
        cnt = nd.count;
        beg = nd.begin;
        isnear = !leaf_far_pass(s_far[wave], nd.cx + ox, nd.cy + oy, nd.cz + oz, nd.leaf_reach,
                                nd.leaf_soft);
      }
      const int cn = isnear ? cnt : 0, cf = isnear ? 0 : cnt;
      const int incl_n = pt_wave_incl_scan(cn);
      const int incl_f = pt_wave_incl_scan(cf);
      const int near_total = __builtin_amdgcn_readlane(incl_n, PT_WAVE - 1);
      const int total = near_total + __builtin_amdgcn_readlane(incl_f, PT_WAVE - 1);
      // each entry's first slot, near entries first
      const int exc = isnear ? incl_n - cn : near_total + incl_f - cf;
      int carry = 0;
      for (int s0 = 0; s0 < total; s0 += PT_WAVE) {
        const bool wfar = s0 >= near_total;
        // a slot's entry: the lane of the last entry start at or below it in
        // this window (marks by start slot, a ballot of the marks), or the
        // previous window's last entry
        int32_t* mark = s_beg[wave];
        mark[lane] = -1;
        __builtin_amdgcn_wave_barrier();
        if (cnt > 0 && exc >= s0 && exc < s0 + PT_WAVE) mark[exc - s0] = lane;
        __builtin_amdgcn_wave_barrier();
        const unsigned long long starts = __ballot(mark[lane] >= 0);
        const unsigned long long upto = starts & (~0ull >> (PT_WAVE - 1 - lane));
        const int leaf = upto ? mark[PT_WAVE - 1 - __clzll(upto)] : carry;
        carry = __builtin_amdgcn_readlane(leaf, PT_WAVE - 1);
        const int lbeg = __shfl(beg, leaf, PT_WAVE);
        const int lexc = __shfl(exc, leaf, PT_WAVE);
        const unsigned jmask = (unsigned)__shfl((int)msk, leaf, PT_WAVE);
        const int slot = s0 + lane;
        if (slot < total) {
          const int64_t j = (int64_t)lbeg + (slot - lexc);
          const float jx = (float)((px[j] + ox) - ccx), jy = (float)((py[j] + oy) - ccy),
                      jz = (float)((pz[j] + oz) - ccz);
          const float jm = (float)pm[j], je = (float)pe[j];
          float sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, sgm = 0.0f;
          particle_pairs32(wfar, jx, jy, jz, jm, je, jmask, s_t[wave], ng, tg, sgx, sgy, sgz, sgm);
          atomicAdd(&gx[j], (double)sgx);
          atomicAdd(&gy[j], (double)sgy);
          atomicAdd(&gz[j], (double)sgz);
          atomicAdd(&gmass[j], (double)sgm);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }

#pragma unroll
    for (int t = 0; t < kANT; t++) {
      if (t < ng) {
        const double rx = pt_wave_sum(tg[t][0]);
        const double ry = pt_wave_sum(tg[t][1]);
        const double rz = pt_wave_sum(tg[t][2]);
        if (lane == t) {
          const int64_t i = first + g0 + t;
          atomicAdd(&gx[i], rx);
          atomicAdd(&gy[i], ry);
          atomicAdd(&gz[i], rz);
        }
      }
    }
  }
}

// Pushes node cotangents to every particle through the closed-form moments of
// its leaf and each ancestor (parent links up to the root).  head/head_pos
// give the particle's bucket (head_pos[j] + head[j] - 1).
__global__ void __launch_bounds__(256)
adjoint_down_kernel(int64_t n, const uint32_t* __restrict__ head,
                    const uint32_t* __restrict__ head_pos, const int32_t* __restrict__ bucket_node,
                    const int32_t* __restrict__ parent, const GravNode* __restrict__ gn,
                    const double* __restrict__ lam_node, int64_t nn,
                    const double* __restrict__ px, const double* __restrict__ py,
                    const double* __restrict__ pz, const double* __restrict__ pm,
                    const double* __restrict__ pe, double* __restrict__ gx,
                    double* __restrict__ gy, double* __restrict__ gz,
                    double* __restrict__ gmass) {
  const int64_t j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= n) return;
  const int64_t bucket = (int64_t)head_pos[j] + head[j] - 1;
  int32_t s = bucket_node[bucket];
  const double x = px[j], y = py[j], z = pz[j], m = pm[j], e = pe[j];
  double ax = 0.0, ay = 0.0, az = 0.0, am = 0.0;
  while (s >= 0) {
    const GravNode g = gn[s];
    if (g.mass != 0.0) {
      double l[kNodeAdj];
#pragma unroll
      for (int k = 0; k < kNodeAdj; k++) l[k] = lam_node[(int64_t)k * nn + s];
      const double ux = x - g.cx, uy = y - g.cy, uz = z - g.cz;
      const double im = 1.0 / g.mass;
      ax += m * (im * l[0] + 2.0 * l[4] * ux + l[5] * uy + l[6] * uz);
      ay += m * (im * l[1] + l[5] * ux + 2.0 * l[7] * uy + l[8] * uz);
      az += m * (im * l[2] + l[6] * ux + l[8] * uy + 2.0 * l[9] * uz);
      am += l[3] + (l[0] * ux + l[1] * uy + l[2] * uz) * im + l[4] * ux * ux + l[5] * ux * uy +
            l[6] * ux * uz + l[7] * uy * uy + l[8] * uy * uz + l[9] * uz * uz +
            l[10] * (e - g.soft) * im;
    }
    s = parent[s];
  }
  gx[j] += ax;
  gy[j] += ay;
  gz[j] += az;
  gmass[j] += am;
}

// FP64 direct sum of the leaf() kernel over every pair, one block per target.
__global__ void __launch_bounds__(256)
direct_all_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                  const double* __restrict__ z, const double* __restrict__ m,
                  const double* __restrict__ e, double* __restrict__ ax, double* __restrict__ ay,
                  double* __restrict__ az, double* __restrict__ pot) {
  __shared__ double red[3][256 / PT_WAVE];
  const int64_t i = blockIdx.x;
  const double tx = x[i], ty = y[i], tz = z[i], te = e[i];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    const double dx = x[j] - tx, dy = y[j] - ty, dz = z[j] - tz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 != 0) {
      double b, db;
      const double r = sqrt(r2);
      spline_b_db(r, 1.0 / r, e[j] + te, b, db);
      const double f = b * m[j];
      sx += dx * f;
      sy += dy * f;
      sz += dz * f;
    }
  }
  sx = pt_wave_sum(sx);
  sy = pt_wave_sum(sy);
  sz = pt_wave_sum(sz);
  const int wave = threadIdx.x / PT_WAVE;
  if ((threadIdx.x & (PT_WAVE - 1)) == 0) {
    red[0][wave] = sx;
    red[1][wave] = sy;
    red[2][wave] = sz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[3] = {0, 0, 0};
    for (int w = 0; w < 256 / PT_WAVE; w++)
      for (int k = 0; k < 3; k++) r[k] += red[k][w];
    ax[i] = r[0];
    ay[i] = r[1];
    az[i] = r[2];
    pot[i] = 0.0;
  }
}

// Exact VJP of the direct sum, one block per particle i:
//   g_i += sum_j K_ij (m_i w_j - m_j w_i),  gm_i += sum_j b_ij (x_i - x_j).w_j.
__global__ void __launch_bounds__(256)
direct_adjoint_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                      const double* __restrict__ z, const double* __restrict__ m,
                      const double* __restrict__ e, const double* __restrict__ wx,
                      const double* __restrict__ wy, const double* __restrict__ wz,
                      double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
                      double* __restrict__ gmass) {
  __shared__ double red[4][256 / PT_WAVE];
  const int64_t i = blockIdx.x;
  const double tx = x[i], ty = y[i], tz = z[i], te = e[i], tm = m[i];
  const double twx = wx[i], twy = wy[i], twz = wz[i];
  double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    const double dx = x[j] - tx, dy = y[j] - ty, dz = z[j] - tz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 != 0) {
      const double r = sqrt(r2);
      double b, db;
      spline_b_db(r, 1.0 / r, e[j] + te, b, db);
      const double vx = tm * wx[j] - m[j] * twx;
      const double vy = tm * wy[j] - m[j] * twy;
      const double vz = tm * wz[j] - m[j] * twz;
      const double dv = dx * vx + dy * vy + dz * vz;
      const double f = db / r * dv;
      sx += b * vx + f * dx;
      sy += b * vy + f * dy;
      sz += b * vz + f * dz;
      sm -= b * (dx * wx[j] + dy * wy[j] + dz * wz[j]);
    }
  }
  sx = pt_wave_sum(sx);
  sy = pt_wave_sum(sy);
  sz = pt_wave_sum(sz);
  sm = pt_wave_sum(sm);
  const int wave = threadIdx.x / PT_WAVE;
  if ((threadIdx.x & (PT_WAVE - 1)) == 0) {
    red[0][wave] = sx;
    red[1][wave] = sy;
    red[2][wave] = sz;
    red[3][wave] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[4] = {0, 0, 0, 0};
    for (int w = 0; w < 256 / PT_WAVE; w++)
      for (int k = 0; k < 4; k++) r[k] += red[k][w];
    gx[i] += r[0];
    gy[i] += r[1];
    gz[i] += r[2];
    gmass[i] += r[3];
  }
}

// Reverse of one kick-drift step (v' = v + dt a(x), x' = x + dt v'):
//   w = dt^2 lx' + dt lv'  (cotangent of a(x)),  lv = lv' + dt lx'.
// lx itself becomes lx' + (da/dx)^T w once the VJP has been added.
__global__ void __launch_bounds__(256)
adjoint_w_kernel(int64_t n, double dt, const double* __restrict__ lx,
                 const double* __restrict__ ly, const double* __restrict__ lz,
                 double* __restrict__ lvx, double* __restrict__ lvy, double* __restrict__ lvz,
                 double* __restrict__ wx, double* __restrict__ wy, double* __restrict__ wz) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const double dt2 = dt * dt;
  wx[i] = dt2 * lx[i] + dt * lvx[i];
  wy[i] = dt2 * ly[i] + dt * lvy[i];
  wz[i] = dt2 * lz[i] + dt * lvz[i];
  lvx[i] += dt * lx[i];
  lvy[i] += dt * ly[i];
  lvz[i] += dt * lz[i];
}

__global__ void __launch_bounds__(256)
gather3_kernel(const uint32_t* __restrict__ perm, int64_t n, const double* __restrict__ a,
               const double* __restrict__ b, const double* __restrict__ c,
               double* __restrict__ as, double* __restrict__ bs, double* __restrict__ cs) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  as[i] = a[j];
  bs[i] = b[j];
  cs[i] = c[j];
}

// out[perm[i]] += in[i] for four arrays (sorted -> caller order).
__global__ void __launch_bounds__(256)
scatter_add4_kernel(const uint32_t* __restrict__ perm, int64_t n, const double* __restrict__ a,
                    const double* __restrict__ b, const double* __restrict__ c,
                    const double* __restrict__ d, double* __restrict__ oa,
                    double* __restrict__ ob, double* __restrict__ oc, double* __restrict__ od) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  oa[j] += a[i];
  ob[j] += b[i];
  oc[j] += c[i];
  od[j] += d[i];
}

// n x 3 interleaved <-> three arrays.
__global__ void __launch_bounds__(256)
split3_kernel(int64_t n, const double* __restrict__ v, double* __restrict__ a,
              double* __restrict__ b, double* __restrict__ c) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  a[i] = v[3 * i];
  b[i] = v[3 * i + 1];
  c[i] = v[3 * i + 2];
}

__global__ void __launch_bounds__(256)
join3_kernel(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
             const double* __restrict__ c, double* __restrict__ v) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  v[3 * i] = a[i];
  v[3 * i + 1] = b[i];
  v[3 * i + 2] = c[i];
}

// lam_node[k][i] = sum over the private copies, for the hot nodes i < nhot.
__global__ void __launch_bounds__(256)
hot_reduce_kernel(const double* __restrict__ hot, int64_t nhot, double* __restrict__ lam_node,
                  int64_t nn) {
  const int64_t q = blockIdx.x * 256ll + threadIdx.x;
  if (q >= nhot * kNodeAdj) return;
  const int64_t k = q / nhot, i = q - k * nhot;
  double v = 0.0;
  for (int p = 0; p < kHotParts; p++) v += hot[((int64_t)p * kNodeAdj + k) * nhot + i];
  lam_node[k * nn + i] = v;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int64_t pt_adjoint_node_scratch(int64_t nn) {
  return (int64_t)kNodeAdj * nn + (int64_t)kHotParts * kNodeAdj * std::min<int64_t>(nn, kHotNodes);
}

int pt_node_adjoint_fields() { return kNodeAdj; }

int pt_launch_adjoint_groups(int precision, const Bucket* bk, const int32_t* gfirst,
                             const int32_t* gnb, int64_t ns, const TreeNode* nodes,
                             const GravNode* gn, const GravNode32* gn32, const double* x,
                             const double* y, const double* z, const double* m, const double* e,
                             const int64_t* pp_off, const int32_t* pp, const uint8_t* ppm,
                             const int64_t* pc_off, const int32_t* pc, const uint8_t* pcm,
                             const double off[3], const double* wx, const double* wy,
                             const double* wz, double* gx, double* gy, double* gz, double* gm,
                             double* lam_node, int64_t nn, hipStream_t s) {
  if (ns <= 0) return 0;
  // lam_node holds pt_adjoint_node_scratch(nn) doubles, all zero on entry:
  // the node cotangents, then the private copies of the hot ones
  const int64_t nhot = std::min<int64_t>(nn, kHotNodes);
  double* hot = lam_node + (int64_t)kNodeAdj * nn;
  const unsigned grid = (unsigned)((ns + kAdjWaves - 1) / kAdjWaves);
  if (precision != 0)  // the compensated-FP32 precision
    adjoint_group32_kernel<<<grid, kAdjThreads, 0, s>>>(
        bk, gfirst, gnb, ns, nodes, gn32, x, y, z, m, e, pp_off, pp, ppm, pc_off, pc, pcm, off[0],
        off[1], off[2], wx, wy, wz, gx, gy, gz, gm, lam_node, nn, hot, nhot);
  else
    adjoint_group_kernel<<<grid, kAdjThreads, 0, s>>>(
        bk, gfirst, gnb, ns, nodes, gn, x, y, z, m, e, pp_off, pp, ppm, pc_off, pc, pcm, off[0],
        off[1], off[2], wx, wy, wz, gx, gy, gz, gm, lam_node, nn, hot, nhot);
  hot_reduce_kernel<<<grid_for(nhot * kNodeAdj), 256, 0, s>>>(hot, nhot, lam_node, nn);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_adjoint_down(int64_t n, const uint32_t* head, const uint32_t* head_pos,
                           const int32_t* bucket_node, const int32_t* parent, const GravNode* gn,
                           const double* lam_node, int64_t nn, const double* x, const double* y,
                           const double* z, const double* m, const double* e, double* gx,
                           double* gy, double* gz, double* gm, hipStream_t s) {
  if (n <= 0) return 0;
  adjoint_down_kernel<<<grid_for(n), 256, 0, s>>>(n, head, head_pos, bucket_node, parent, gn,
                                                  lam_node, nn, x, y, z, m, e, gx, gy, gz, gm);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_direct_all(int64_t n, const double* x, const double* y, const double* z,
                         const double* m, const double* e, double* ax, double* ay, double* az,
                         double* pot, hipStream_t s) {
  if (n <= 0) return 0;
  direct_all_kernel<<<(unsigned)n, 256, 0, s>>>(n, x, y, z, m, e, ax, ay, az, pot);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_direct_adjoint(int64_t n, const double* x, const double* y, const double* z,
                             const double* m, const double* e, const double* wx,
                             const double* wy, const double* wz, double* gx, double* gy,
                             double* gz, double* gm, hipStream_t s) {
  if (n <= 0) return 0;
  direct_adjoint_kernel<<<(unsigned)n, 256, 0, s>>>(n, x, y, z, m, e, wx, wy, wz, gx, gy, gz, gm);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_adjoint_w(int64_t n, double dt, const double* lx, const double* ly,
                        const double* lz, double* lvx, double* lvy, double* lvz, double* wx,
                        double* wy, double* wz, hipStream_t s) {
  if (n <= 0) return 0;
  adjoint_w_kernel<<<grid_for(n), 256, 0, s>>>(n, dt, lx, ly, lz, lvx, lvy, lvz, wx, wy, wz);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_gather3(const uint32_t* perm, int64_t n, const double* a, const double* b,
                      const double* c, double* as, double* bs, double* cs, hipStream_t s) {
  if (n <= 0) return 0;
  gather3_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, a, b, c, as, bs, cs);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_scatter_add4(const uint32_t* perm, int64_t n, const double* a, const double* b,
                           const double* c, const double* d, double* oa, double* ob, double* oc,
                           double* od, hipStream_t s) {
  if (n <= 0) return 0;
  scatter_add4_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, a, b, c, d, oa, ob, oc, od);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_split3(int64_t n, const double* v, double* a, double* b, double* c, hipStream_t s) {
  if (n <= 0) return 0;
  split3_kernel<<<grid_for(n), 256, 0, s>>>(n, v, a, b, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pt_launch_join3(int64_t n, const double* a, const double* b, const double* c, double* v,
                    hipStream_t s) {
  if (n <= 0) return 0;
  join3_kernel<<<grid_for(n), 256, 0, s>>>(n, a, b, c, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

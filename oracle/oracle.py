"""ctypes binding of the CPU oracle (oracle/pt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(paratreet_amd) never imports this module.

Everything here takes numpy arrays; the C side is single-source-of-truth for the
arithmetic (see the header comment of pt_oracle.c for how each function is
pinned).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpt_oracle.so")

OR_EMPTY, OR_LEAF, OR_INTERNAL = 0, 1, 2

NODE_DTYPE = np.dtype([
    ("key", np.uint64), ("type", np.int32), ("depth", np.int32),
    ("begin", np.int64), ("count", np.int64), ("parent", np.int32),
    ("child", np.int32, (8,)),
], align=True)

CDATA_DTYPE = np.dtype([
    ("radius", np.float64), ("soft", np.float64), ("mass", np.float64),
    ("cm", np.float64, (3,)), ("q", np.float64, (6,)),
    ("lo", np.float64, (3,)), ("hi", np.float64, (3,)), ("count", np.int64),
], align=True)

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double


def _declare(L):
    L.or_universe.argtypes = [_P, _P, _P, _I64, _P]
    L.or_keys.argtypes = [_P, _P, _P, _I64, _P, _P]
    L.or_sort.argtypes = [_P, _I64, _P, _P]
    L.or_build.argtypes = [_P, _I64, ctypes.c_int, _P, _I64]
    L.or_build.restype = _I64
    L.or_moments.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P]
    L.or_buckets.argtypes = [_P, _I64, _P]
    L.or_buckets.restype = _I64
    L.or_gravity.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _D, _P, _P, _I64,
                             _P, _P, _P, _P, _P, ctypes.c_int]
    L.or_gravity_soft_pairs.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P,
                                        _P, _P, _P]
    L.or_lists.argtypes = [_P, _I64, _P, _D, _P, _P, _P, _P, _P, _P, _P]
    L.or_traversal_counters.argtypes = [_P, _I64, _P, _D, _P, _P, ctypes.c_int]
    L.or_dual.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _D, _P, ctypes.c_int32,
                          ctypes.c_int32, _P, _P, _P, _P, _I64, _P]
    L.or_dual_lists.argtypes = [_P, _I64, _P, _D, _P, ctypes.c_int32, ctypes.c_int32, _P, _P,
                                _P, _P, _P, _P]
    L.or_direct.argtypes = [_P, _P, _P, _P, _P, _I64, _P, _I64, _P, _P, _P, ctypes.c_int]
    L.or_spline.argtypes = [_D, _D, _P, _P]
    L.or_splineq.argtypes = [_D, _D, _D, _P, _P, _P, _P]
    L.or_intersect_box_point.argtypes = [_P, _P, _P, _D]
    L.or_intersect_box_point.restype = ctypes.c_int
    L.or_kick_drift.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _D]
    L.or_direct_vjp.argtypes = [_P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P]
    L.or_hexa_moments.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _P]
    L.or_hexa_raw_moments.argtypes = [_P, _P, _P, _P, _I64, _P, _P]
    L.or_hexa_reduce.argtypes = [_D, _P, _P, _P]
    L.or_hexa_eval.argtypes = [_P, _D, _D, _D, _P, _P]
    L.or_gravity_hexa.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _I64,
                                  _P, _P, _P, _P, _P, ctypes.c_int]
    L.or_lists_hexa.argtypes = [_P, _I64, _P, _P, _D, _P, _P, _P, _P, _P, _P, _P]
    L.or_set_ewald_libm_perturbation.argtypes = [_D, _D]
    L.or_ewald.argtypes = [_P, _P, _P, _I64, _D, _P, _P, _D, ctypes.c_int, _P, _P, _P, _P]
    L.or_ewald_hexa.argtypes = [_P, _P, _P, _I64, _D, _P, _P, _D, ctypes.c_int, _P, _P, _P, _P]
    L.or_ewald_hexa_general.argtypes = [_P, _P, _P, _I64, _D, _P, _P, _D, ctypes.c_int, _D, _D,
                                        _D, _P, _P, _P, _P]
    L.or_ewald_general.argtypes = [_P, _P, _P, _I64, _D, _P, _P, _D, ctypes.c_int, _D, _D, _D,
                                   _P, _P, _P, _P]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def universe(x, y, z):
    x, y, z = _f64(x), _f64(y), _f64(z)
    box = np.zeros(6)
    lib().or_universe(_p(x), _p(y), _p(z), len(x), _p(box))
    return box


def keys(x, y, z, box):
    x, y, z = _f64(x), _f64(y), _f64(z)
    box = _f64(box)
    k = np.zeros(len(x), dtype=np.uint64)
    lib().or_keys(_p(x), _p(y), _p(z), len(x), _p(box), _p(k))
    return k


def sort(k):
    k = np.ascontiguousarray(k, dtype=np.uint64)
    sk = np.zeros_like(k)
    perm = np.zeros(len(k), dtype=np.int32)
    lib().or_sort(_p(k), len(k), _p(sk), _p(perm))
    return sk, perm


def build_tree(sorted_keys, bucket):
    sk = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    cap = max(64, 4 * len(sk) // max(bucket, 1) + 64)
    while True:
        nodes = np.zeros(cap, dtype=NODE_DTYPE)
        nn = lib().or_build(_p(sk), len(sk), bucket, _p(nodes), cap)
        if nn >= 0:
            return nodes[:nn].copy()
        cap = -nn


def moments(nodes, x, y, z, m, eps):
    nodes = np.ascontiguousarray(nodes)
    out = np.zeros(len(nodes), dtype=CDATA_DTYPE)
    x, y, z, m, eps = (_f64(a) for a in (x, y, z, m, eps))
    lib().or_moments(_p(nodes), len(nodes), _p(x), _p(y), _p(z), _p(m), _p(eps), _p(out))
    return out


def buckets(nodes):
    nodes = np.ascontiguousarray(nodes)
    nb = lib().or_buckets(_p(nodes), len(nodes), None)
    out = np.zeros(nb, dtype=np.int32)
    lib().or_buckets(_p(nodes), len(nodes), _p(out))
    return out


HEXA_FIELDS = ("m", "xx", "yy", "xy", "xz", "yz", "xxx", "xyy", "xxy", "yyy", "xxz", "yyz",
               "xyz", "xxxx", "xyyy", "xxxy", "yyyy", "xxxz", "yyyz", "xxyy", "xxyz", "xyyz")


def hexa_moments(nodes, data, x, y, z, m, raw=False):
    """Reduced moments (nn x 22, HEXA_FIELDS order) of every node; with
    raw=True also the raw third/fourth moments about each cm (nn x 25)."""
    nodes = np.ascontiguousarray(nodes)
    data = np.ascontiguousarray(data)
    x, y, z, m = (_f64(a) for a in (x, y, z, m))
    out = np.zeros((len(nodes), 22))
    r = np.zeros((len(nodes), 25))
    lib().or_hexa_moments(_p(nodes), len(nodes), _p(data), _p(x), _p(y), _p(z), _p(m), _p(r),
                          _p(out))
    return (out, r) if raw else out


def hexa_raw_moments(pos, m, cm):
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (_f64(pos[:, k]) for k in range(3))
    out = np.zeros(25)
    lib().or_hexa_raw_moments(_p(x), _p(y), _p(z), _p(_f64(m)), len(x), _p(_f64(cm)), _p(out))
    return out


def hexa_reduce(mass, q, raw):
    out = np.zeros(22)
    lib().or_hexa_reduce(float(mass), _p(_f64(q)), _p(_f64(raw)), _p(out))
    return out


def hexa_eval(h, r):
    """(potential, acceleration) of reduced moments h at r = target - cm."""
    h = _f64(h)
    pot = ctypes.c_double()
    acc = np.zeros(3)
    lib().or_hexa_eval(_p(h), float(r[0]), float(r[1]), float(r[2]), ctypes.byref(pot), _p(acc))
    return pot.value, acc


def gravity(nodes, data, x, y, z, m, eps, theta, offset=(0.0, 0.0, 0.0),
            bucket_sel=None, nthreads=0, hexa=None):
    """Accelerations and potentials in sorted order + counters
    {n_node_ints, n_part_ints, n_opens, n_closes}.  hexa: the nodes' reduced
    moments (hexa_moments) for the HEXADECAPOLE build, None for quadrupole."""
    n = len(x)
    x, y, z, m, eps = (_f64(a) for a in (x, y, z, m, eps))
    off = _f64(offset)
    ax, ay, az, pot = (np.zeros(n) for _ in range(4))
    cnt = np.zeros(4, dtype=np.uint64)
    sel = None if bucket_sel is None else np.ascontiguousarray(bucket_sel, dtype=np.int32)
    if hexa is None:
        lib().or_gravity(_p(nodes), len(nodes), _p(data), _p(x), _p(y), _p(z), _p(m), _p(eps),
                         float(theta), _p(off), _p(sel), 0 if sel is None else len(sel),
                         _p(ax), _p(ay), _p(az), _p(pot), _p(cnt), int(nthreads))
    else:
        h = _f64(hexa)
        lib().or_gravity_hexa(_p(nodes), len(nodes), _p(data), _p(h), _p(x), _p(y), _p(z),
                              _p(m), _p(eps), float(theta), _p(off), _p(sel),
                              0 if sel is None else len(sel), _p(ax), _p(ay), _p(az), _p(pot),
                              _p(cnt), int(nthreads))
    return np.stack([ax, ay, az], axis=1), pot, cnt


def gravity_soft_pairs(nodes, data, x, y, z, m, eps, theta, offset=(0.0, 0.0, 0.0)):
    """gravity() over all buckets (quadrupole build) plus the number of P-C
    target pairs that reach SPLINEQ's softened branch: (acc, pot, counters,
    (pairs with u < 1, pairs with u >= 1))."""
    n = len(x)
    x, y, z, m, eps = (_f64(a) for a in (x, y, z, m, eps))
    ax, ay, az, pot = (np.zeros(n) for _ in range(4))
    cnt = np.zeros(4, dtype=np.uint64)
    sp = np.zeros(2, dtype=np.uint64)
    lib().or_gravity_soft_pairs(_p(nodes), len(nodes), _p(data), _p(x), _p(y), _p(z), _p(m),
                                _p(eps), float(theta), _p(_f64(offset)), _p(ax), _p(ay), _p(az),
                                _p(pot), _p(cnt), _p(sp))
    return np.stack([ax, ay, az], axis=1), pot, cnt, (int(sp[0]), int(sp[1]))


def traversal_counters(nodes, data, theta, offset=(0.0, 0.0, 0.0), nthreads=0):
    """gravity()'s counters (quadrupole build) from the walk alone."""
    cnt = np.zeros(4, dtype=np.uint64)
    lib().or_traversal_counters(_p(nodes), len(nodes), _p(data), float(theta), _p(_f64(offset)),
                                _p(cnt), int(nthreads))
    return cnt


def lists(nodes, data, theta, offset=(0.0, 0.0, 0.0), hexa=None):
    """Per-bucket interaction lists in DFS order: (pp_off, pp, pc_off, pc) with
    node ids into `nodes` (HEXADECAPOLE open() when hexa is given)."""
    nb = len(buckets(nodes))
    off = _f64(offset)
    ppc = np.zeros(nb, dtype=np.int64)
    pcc = np.zeros(nb, dtype=np.int64)
    L = lib()
    if hexa is not None:
        h = _f64(hexa)

        def call(*a):
            L.or_lists_hexa(_p(nodes), len(nodes), _p(data), _p(h), float(theta), _p(off), *a)
    else:
        def call(*a):
            L.or_lists(_p(nodes), len(nodes), _p(data), float(theta), _p(off), *a)
    call(_p(ppc), _p(pcc), None, None, None, None)
    pp_off = np.zeros(nb + 1, dtype=np.int64)
    pc_off = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(ppc, out=pp_off[1:])
    np.cumsum(pcc, out=pc_off[1:])
    pp = np.zeros(max(1, pp_off[-1]), dtype=np.int32)
    pc = np.zeros(max(1, pc_off[-1]), dtype=np.int32)
    call(_p(ppc), _p(pcc), _p(pp_off), _p(pc_off), _p(pp), _p(pc))
    return pp_off, pp[:pp_off[-1]], pc_off, pc[:pc_off[-1]]


def dual(nodes, data, x, y, z, m, eps, theta, source, target, offset=(0.0, 0.0, 0.0)):
    """DualTraverser<GravityVisitor> from the node pair (source, target) (ids
    into `nodes`): accelerations and potentials in sorted order and the
    counters {n_node_ints, n_part_ints, n_opens, n_closes} (uint64, wrapping
    as the reference's do)."""
    n = len(x)
    x, y, z, m, eps = (_f64(a) for a in (x, y, z, m, eps))
    ax, ay, az, pot = (np.zeros(n) for _ in range(4))
    cnt = np.zeros(4, dtype=np.uint64)
    lib().or_dual(_p(nodes), len(nodes), _p(data), _p(x), _p(y), _p(z), _p(m), _p(eps),
                  float(theta), _p(_f64(offset)), int(source), int(target), _p(ax), _p(ay),
                  _p(az), _p(pot), n, _p(cnt))
    return np.stack([ax, ay, az], axis=1), pot, cnt


def dual_lists(nodes, data, theta, source, target, offset=(0.0, 0.0, 0.0)):
    """The leaf() and node() calls of dual() as per-bucket lists (pp_off, pp,
    pc_off, pc), each bucket's entries in the traversal's order."""
    nb = len(buckets(nodes))
    off = _f64(offset)
    ppc = np.zeros(nb, dtype=np.int64)
    pcc = np.zeros(nb, dtype=np.int64)
    L = lib()
    L.or_dual_lists(_p(nodes), len(nodes), _p(data), float(theta), _p(off), int(source),
                    int(target), _p(ppc), _p(pcc), None, None, None, None)
    pp_off = np.zeros(nb + 1, dtype=np.int64)
    pc_off = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(ppc, out=pp_off[1:])
    np.cumsum(pcc, out=pc_off[1:])
    pp = np.zeros(max(1, pp_off[-1]), dtype=np.int32)
    pc = np.zeros(max(1, pc_off[-1]), dtype=np.int32)
    L.or_dual_lists(_p(nodes), len(nodes), _p(data), float(theta), _p(off), int(source),
                    int(target), _p(ppc), _p(pcc), _p(pp_off), _p(pc_off), _p(pp), _p(pc))
    return pp_off, pp[:pp_off[-1]], pc_off, pc[:pc_off[-1]]


def direct(x, y, z, m, eps, targets, nthreads=0):
    x, y, z, m, eps = (_f64(a) for a in (x, y, z, m, eps))
    t = np.ascontiguousarray(targets, dtype=np.int64)
    ax, ay, az = (np.zeros(len(t)) for _ in range(3))
    lib().or_direct(_p(x), _p(y), _p(z), _p(m), _p(eps), len(x), _p(t), len(t),
                    _p(ax), _p(ay), _p(az), int(nthreads))
    return np.stack([ax, ay, az], axis=1)


def direct_vjp(pos, m, eps, w):
    """(da/dx)^T w (n x 3) and (da/dm)^T w (n) of the direct sum."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (_f64(pos[:, k]) for k in range(3))
    w = np.asarray(w, dtype=np.float64)
    wx, wy, wz = (_f64(w[:, k]) for k in range(3))
    n = len(x)
    gx, gy, gz, gm = (np.zeros(n) for _ in range(4))
    lib().or_direct_vjp(_p(x), _p(y), _p(z), _p(_f64(m)), _p(_f64(eps)), n, _p(wx), _p(wy),
                        _p(wz), _p(gx), _p(gy), _p(gz), _p(gm))
    return np.stack([gx, gy, gz], axis=1), gm


def ewald(pos, mass, cm, q, L, n_replicas, a_fac=None, ew_cut=2.6, h_cut=2.8):
    """LeafEwaldFn's quadrupole-form lattice correction for the particles at
    pos (n x 3) from a root (mass, cm, raw second moments q: xx xy xz yy yz
    zz): (acc n x 3, pot n).  a_fac=None uses the reference's constants."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (_f64(pos[:, k]) for k in range(3))
    n = len(x)
    ax, ay, az, pot = (np.zeros(n) for _ in range(4))
    if a_fac is None:
        lib().or_ewald(_p(x), _p(y), _p(z), n, float(mass), _p(_f64(cm)), _p(_f64(q)), float(L),
                       int(n_replicas), _p(ax), _p(ay), _p(az), _p(pot))
    else:
        lib().or_ewald_general(_p(x), _p(y), _p(z), n, float(mass), _p(_f64(cm)), _p(_f64(q)),
                               float(L), int(n_replicas), float(a_fac), float(ew_cut),
                               float(h_cut), _p(ax), _p(ay), _p(az), _p(pot))
    return np.stack([ax, ay, az], axis=1), pot


_ewald_fn = ewald


def ewald_hexa(pos, mass, cm, h, L, n_replicas, a_fac=None, ew_cut=2.6, h_cut=2.8):
    """LeafEwaldFn's HEXADECAPOLE-form lattice correction from a root (mass,
    cm, the 21 reduced moments h in FMOMR order, as hexa_moments()[:, 1:]):
    (acc n x 3, pot n).  a_fac=None uses the reference's constants."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (_f64(pos[:, k]) for k in range(3))
    n = len(x)
    ax, ay, az, pot = (np.zeros(n) for _ in range(4))
    h = _f64(h)
    assert h.shape == (21,)
    if a_fac is None:
        lib().or_ewald_hexa(_p(x), _p(y), _p(z), n, float(mass), _p(_f64(cm)), _p(h), float(L),
                            int(n_replicas), _p(ax), _p(ay), _p(az), _p(pot))
    else:
        lib().or_ewald_hexa_general(_p(x), _p(y), _p(z), n, float(mass), _p(_f64(cm)), _p(h),
                                    float(L), int(n_replicas), float(a_fac), float(ew_cut),
                                    float(h_cut), _p(ax), _p(ay), _p(az), _p(pot))
    return np.stack([ax, ay, az], axis=1), pot


def set_ewald_libm_perturbation(erf_ulps=0.0, exp_ulps=0.0):
    """Scales the real-space erf / erfc and exp values of the Ewald forms by
    1 + ulps * 2^-52 (about that many ulps) until reset with (0, 0): a probe
    of how far a libm that differs by that much moves the restated
    algorithm's result."""
    lib().or_set_ewald_libm_perturbation(float(erf_ulps), float(exp_ulps))


def spline(r2, twoh):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().or_spline(float(r2), float(twoh), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def splineq(invr, r2, twoh):
    v = [ctypes.c_double() for _ in range(4)]
    lib().or_splineq(float(invr), float(r2), float(twoh), *[ctypes.byref(t) for t in v])
    return tuple(t.value for t in v)


def intersect_box_point(lo, hi, p, rsq):
    lo, hi, p = _f64(lo), _f64(hi), _f64(p)
    return bool(lib().or_intersect_box_point(_p(lo), _p(hi), _p(p), float(rsq)))


def kick_drift(pos, vel, acc, dt):
    """Returns new (pos, vel); inputs are (n, 3) arrays."""
    x, y, z = (np.array(pos[:, k], dtype=np.float64) for k in range(3))
    vx, vy, vz = (np.array(vel[:, k], dtype=np.float64) for k in range(3))
    ax, ay, az = (_f64(acc[:, k]) for k in range(3))
    lib().or_kick_drift(_p(x), _p(y), _p(z), _p(vx), _p(vy), _p(vz), _p(ax), _p(ay), _p(az),
                        len(x), float(dt))
    return np.stack([x, y, z], axis=1), np.stack([vx, vy, vz], axis=1)


class Pipeline:
    """Full oracle pipeline on one particle set (original order in, sorted
    order inside).  Attributes mirror the stages of the reference path."""

    def __init__(self, pos, mass, eps, bucket, theta, offset=(0.0, 0.0, 0.0),
                 hexadecapole=False):
        pos = np.asarray(pos, dtype=np.float64)
        self.n = len(pos)
        self.bucket = bucket
        self.theta = theta
        self.offset = offset
        self.box = universe(pos[:, 0], pos[:, 1], pos[:, 2])
        self.keys = keys(pos[:, 0], pos[:, 1], pos[:, 2], self.box)
        self.sorted_keys, self.perm = sort(self.keys)
        self.x = pos[self.perm, 0].copy()
        self.y = pos[self.perm, 1].copy()
        self.z = pos[self.perm, 2].copy()
        self.m = np.asarray(mass, dtype=np.float64)[self.perm].copy()
        self.eps = np.asarray(eps, dtype=np.float64)[self.perm].copy()
        self.nodes = build_tree(self.sorted_keys, bucket)
        self.data = moments(self.nodes, self.x, self.y, self.z, self.m, self.eps)
        self.bucket_nodes = buckets(self.nodes)
        self.hexa = (hexa_moments(self.nodes, self.data, self.x, self.y, self.z, self.m)
                     if hexadecapole else None)

    def gravity(self, bucket_sel=None, nthreads=0):
        acc, pot, cnt = gravity(self.nodes, self.data, self.x, self.y, self.z, self.m,
                                self.eps, self.theta, self.offset, bucket_sel, nthreads,
                                hexa=self.hexa)
        return acc, pot, cnt

    def gravity_unsorted(self, nthreads=0):
        """Accelerations/potentials in the caller's original particle order."""
        acc, pot, cnt = self.gravity(nthreads=nthreads)
        out_a = np.empty_like(acc)
        out_p = np.empty_like(pot)
        out_a[self.perm] = acc
        out_p[self.perm] = pot
        return out_a, out_p, cnt

    def gravity_replicas(self, n_replicas, period, nthreads=0, ewald=False):
        """ExMain::traversalFn with periodic (examples/Gravity.C:35-46): the sum
        of the traversals with offsets (X, Y, Z) * period, X, Y, Z in
        [-N, N] (X outermost), in the caller's original order, with the
        counters summed over the replicas; with ewald, plus LeafEwaldFn's
        lattice correction from the root (L = period[0]): the quadrupole form,
        or in the HEXADECAPOLE build the form with the root's moments to
        fourth order."""
        n = len(self.x)
        acc = np.zeros((n, 3))
        pot = np.zeros(n)
        cnt = np.zeros(4, dtype=np.uint64)
        r = range(-n_replicas, n_replicas + 1)
        for X in r:
            for Y in r:
                for Z in r:
                    off = (X * period[0], Y * period[1], Z * period[2])
                    a, p, c = gravity(self.nodes, self.data, self.x, self.y, self.z, self.m,
                                      self.eps, self.theta, off, None, nthreads, hexa=self.hexa)
                    acc += a
                    pot += p
                    cnt += c
        if ewald:
            root = self.data[0]
            xyz = np.stack([self.x, self.y, self.z], axis=1)
            if self.hexa is None:
                ea, ep = _ewald_fn(xyz, root["mass"], root["cm"], root["q"], period[0],
                                   n_replicas)
            else:
                # the HEXADECAPOLE branch: the root's moments to fourth order
                ea, ep = ewald_hexa(xyz, root["mass"], root["cm"], self.hexa[0, 1:], period[0],
                                    n_replicas)
            acc += ea
            pot += ep
        out_a = np.empty_like(acc)
        out_p = np.empty_like(pot)
        out_a[self.perm] = acc
        out_p[self.perm] = pot
        return out_a, out_p, cnt

    def lists(self):
        return lists(self.nodes, self.data, self.theta, self.offset, hexa=self.hexa)

    def node_id(self, key):
        """Index into self.nodes of the node with this key."""
        i = np.nonzero(self.nodes["key"] == np.uint64(key))[0]
        if len(i) != 1:
            raise KeyError(key)
        return int(i[0])

    def dual(self, source_key=1, target_key=1):
        """Subtree::startDual's traversal from the pair of nodes with these
        keys (root, root by default), in the caller's original order."""
        acc, pot, cnt = dual(self.nodes, self.data, self.x, self.y, self.z, self.m, self.eps,
                             self.theta, self.node_id(source_key), self.node_id(target_key),
                             self.offset)
        out_a = np.empty_like(acc)
        out_p = np.empty_like(pot)
        out_a[self.perm] = acc
        out_p[self.perm] = pot
        return out_a, out_p, cnt

    def dual_lists(self, source_key=1, target_key=1):
        return dual_lists(self.nodes, self.data, self.theta, self.node_id(source_key),
                          self.node_id(target_key), self.offset)

"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Each test fixes a part of the oracle against something independent of it:
the reference's own fixture (tests/gravity/direct.acc), closed forms, brute
force on small inputs, or invariants a wrong index/sign/term would break.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import left_aligned


def _lambs():
    d = np.load("tests/golden/lambs_30k.npz")
    return (d["pos"].astype(np.float64), d["mass"].astype(np.float64),
            d["soft"].astype(np.float64), d["direct_acc"])


# ---------------------------------------------------------------- leaf/SPLINE

def test_direct_sum_reproduces_reference_fixture():
    """or_direct (leaf() kernel + SPLINE, 2h = eps_i + eps_j) against the
    reference's direct.acc for the lambs 30K fixture (tests/gravity/)."""
    pos, m, e, ref = _lambs()
    rng = np.random.default_rng(3)
    t = rng.choice(len(pos), 600, replace=False)
    acc = O.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, e, t)
    rel = np.linalg.norm(acc - ref[t], axis=1) / np.linalg.norm(ref[t], axis=1)
    # the fixture was written from float32 positions with '%.14g'
    assert np.sqrt(np.mean(rel ** 2)) < 3e-8
    assert rel.max() < 3e-7


def test_plummer_softening_does_not_reproduce_fixture():
    """Negative control: a Plummer-softened kernel misses direct.acc by ~10%,
    so the fixture really distinguishes the spline kernel."""
    pos, m, e, ref = _lambs()
    t = np.arange(0, 30000, 500)
    acc = np.zeros((len(t), 3))
    for k, i in enumerate(t):
        d = pos - pos[i]
        r2 = (d * d).sum(1) + (e + e[i]) ** 2 / 4
        w = m / r2 ** 1.5
        w[i] = 0
        acc[k] = (d * w[:, None]).sum(0)
    rel = np.linalg.norm(acc - ref[t], axis=1) / np.linalg.norm(ref[t], axis=1)
    assert np.sqrt(np.mean(rel ** 2)) > 1e-3


@pytest.mark.parametrize("twoh", [0.02, 0.5, 3.0])
def test_spline_force_is_minus_gradient_of_potential(twoh):
    """SPLINE returns potential term a(r) and force term b(r) of the cubic
    spline; b = -(1/r) da/dr at every radius (central differences)."""
    for r in np.concatenate([np.linspace(0.05, 2.5, 60) * twoh]):
        h = 1e-6 * twoh
        ap, _ = O.spline((r + h) ** 2, twoh)
        am, _ = O.spline((r - h) ** 2, twoh)
        _, b = O.spline(r * r, twoh)
        dadr = (ap - am) / (2 * h)
        assert abs(-dadr / r - b) <= 1e-6 * max(abs(b), 1e-12), (r, twoh)


@pytest.mark.parametrize("twoh", [0.01, 1.0])
def test_spline_continuity_and_far_field(twoh):
    for u in (1.0, 2.0):  # branch boundaries r = twoh/2, r = twoh
        r = u * twoh / 2
        a0, b0 = O.spline((r * (1 - 1e-12)) ** 2, twoh)
        a1, b1 = O.spline((r * (1 + 1e-12)) ** 2, twoh)
        assert abs(a0 - a1) <= 1e-9 * abs(a0) and abs(b0 - b1) <= 1e-9 * abs(b0)
    r = 1.7 * twoh
    a, b = O.spline(r * r, twoh)
    assert a == 1.0 / r and b == (1.0 / r) ** 3


def test_splineq_far_field_closed_form():
    """Outside the softening SPLINEQ returns 1/r, 1/r^3, 3/r^5, 15/r^7."""
    for r in (0.3, 1.0, 7.5):
        a, b, c, d = O.splineq(1.0 / r, r * r, 0.01)
        assert np.allclose([a, b, c, d], [1 / r, r ** -3, 3 * r ** -5, 15 * r ** -7], rtol=1e-14)


def _spline_softened_closed_form(r):
    """The cubic-spline softened potential and its derivative chain with
    softening length h = 1 (Monaghan & Lattanzio 1985; Hernquist & Katz 1989,
    eq. A2), u = r: a = -phi, b = -a'/r, c = -b'/r, d = -c'/r."""
    u = r
    if u < 1.0:
        a = 7 / 5 - 2 / 3 * u ** 2 + 3 / 10 * u ** 4 - 1 / 10 * u ** 5
        b = 4 / 3 - 6 / 5 * u ** 2 + 1 / 2 * u ** 3
        c = 12 / 5 - 3 / 2 * u
        d = 3 / 2 / u
    else:
        a = -1 / (15 * u) + 8 / 5 - 4 / 3 * u ** 2 + u ** 3 - 3 / 10 * u ** 4 + 1 / 30 * u ** 5
        b = -1 / (15 * u ** 3) + 8 / 3 - 3 * u + 6 / 5 * u ** 2 - 1 / 6 * u ** 3
        c = -1 / (5 * u ** 5) + 3 / u - 12 / 5 + 1 / 2 * u
        d = -1 / u ** 7 + 3 / u ** 3 - 1 / (2 * u)
    return a, b, c, d


def test_splineq_softened_branches_at_unit_softening_length():
    """SPLINEQ's softened branches (examples/GravityVisitor.h:40-73), restated
    with the reference's grouping, in which the terms in powers of u do not
    carry the powers of dih = 2 / twoh.  At twoh = 2 (dih = 1)
    that grouping coincides with the spline kernel, so there every
    coefficient of both branches is checked against the closed forms, the
    derivative chain b = -a'/r, c = -b'/r, d = -c'/r (central differences),
    and continuity at u = 1 and at r = twoh with the unsoftened forms."""
    twoh = 2.0
    for r in (0.05, 0.3, 0.7, 0.95, 1.05, 1.3, 1.7, 1.95):
        got = O.splineq(1.0 / r, r * r, twoh)
        want = _spline_softened_closed_form(r)
        assert np.allclose(got, want, rtol=1e-13, atol=0), (r, got, want)
        h = 1e-5 * r
        lo = O.splineq(1.0 / (r - h), (r - h) ** 2, twoh)
        hi = O.splineq(1.0 / (r + h), (r + h) ** 2, twoh)
        for k in range(3):
            deriv = (hi[k] - lo[k]) / (2 * h)
            assert abs(-deriv / r - got[k + 1]) <= 1e-7 * abs(got[k + 1]), (r, k)
    for k in range(4):
        below = O.splineq(1.0 / (1 - 1e-9), (1 - 1e-9) ** 2, twoh)[k]
        above = O.splineq(1.0 / (1 + 1e-9), (1 + 1e-9) ** 2, twoh)[k]
        assert abs(below - above) <= 1e-7 * abs(above)
        inside = O.splineq(1.0 / (2 - 1e-9), (2 - 1e-9) ** 2, twoh)[k]
        outside = O.splineq(1.0 / (2 + 1e-9), (2 + 1e-9) ** 2, twoh)[k]
        assert abs(inside - outside) <= 1e-7 * abs(outside)


# ---------------------------------------------------------------- keys

def _keys_numpy(pos, box):
    """Independent Morton keys: per-axis floor quantisation, bitwise loop."""
    q = np.empty((len(pos), 3), dtype=np.uint64)
    for k in range(3):
        u = (pos[:, k] - box[k]) / (box[3 + k] - box[k])
        q[:, k] = np.clip(np.floor(u * 2.0 ** 21), 0, 2 ** 21 - 1).astype(np.uint64)
    key = np.zeros(len(pos), dtype=np.uint64)
    for bit in range(20, -1, -1):
        for k in range(3):
            key = (key << np.uint64(1)) | ((q[:, k] >> np.uint64(bit)) & np.uint64(1))
    return key | (np.uint64(1) << np.uint64(63))


def test_keys_match_independent_interleave():
    rng = np.random.default_rng(5)
    pos = rng.standard_normal((5000, 3)) * [1.0, 3.0, 0.2]
    box = O.universe(pos[:, 0], pos[:, 1], pos[:, 2])
    assert np.array_equal(O.keys(pos[:, 0], pos[:, 1], pos[:, 2], box), _keys_numpy(pos, box))


@pytest.mark.parametrize("level", [1, 2, 3])
def test_keys_brute_force_cells(level):
    """Every cell centre of a 2^level grid gets the key whose top 3*level bits
    interleave the cell indices (x most significant)."""
    n = 2 ** level
    box = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    idx = np.array([(i, j, k) for i in range(n) for j in range(n) for k in range(n)])
    pos = (idx + 0.5) / n
    keys = O.keys(pos[:, 0], pos[:, 1], pos[:, 2], box)
    for (i, j, k), key in zip(idx, keys):
        expect = 1
        for b in range(level - 1, -1, -1):
            expect = (expect << 3) | (((i >> b) & 1) << 2) | (((j >> b) & 1) << 1) | ((k >> b) & 1)
        assert int(key) >> (63 - 3 * level) == expect


def test_universe_is_slopped_cube():
    rng = np.random.default_rng(1)
    pos = rng.uniform(-2, 5, (1000, 3)) * [1, 0.5, 0.25]
    box = O.universe(pos[:, 0], pos[:, 1], pos[:, 2])
    ext = box[3:] - box[:3]
    assert np.allclose(ext, ext[0], rtol=1e-15)
    assert np.isclose(ext[0], (pos.max(0) - pos.min(0)).max() * (1 + 1.91e-6), rtol=1e-15)
    assert np.all(box[:3] < pos.min(0)) and np.all(box[3:] > pos.max(0))


def test_sort_is_stable_lexsort():
    rng = np.random.default_rng(2)
    k = rng.integers(0, 50, 4000).astype(np.uint64) | (np.uint64(1) << np.uint64(63))
    sk, perm = O.sort(k)
    ref = np.lexsort((np.arange(len(k)), k))
    assert np.array_equal(perm, ref) and np.array_equal(sk, k[ref])


# ---------------------------------------------------------------- tree

def _small_pipeline(n=3000, bucket=8, seed=4, theta=0.7):
    pos, _, m = ics.plummer(n, seed=seed)
    return O.Pipeline(pos, m, np.full(n, 0.02), bucket=bucket, theta=theta)


def test_tree_brute_force_prefix_counts():
    """Every node's range is exactly the particles carrying its key prefix."""
    P = _small_pipeline(2000, 6)
    sk = P.sorted_keys
    for nd in P.nodes:
        shift = np.uint64(63 - 3 * nd["depth"])
        prefix = sk >> shift
        members = np.nonzero(prefix == np.uint64(nd["key"]))[0]
        if nd["count"] == 0:
            assert members.size == 0
        else:
            assert members[0] == nd["begin"] and members.size == nd["count"]
            assert members[-1] == nd["begin"] + nd["count"] - 1


def test_tree_invariants():
    P = _small_pipeline(5000, 12)
    nodes = P.nodes
    for i, nd in enumerate(nodes):
        if nd["type"] == O.OR_INTERNAL:
            assert nd["count"] > 12
            ch = nodes[nd["child"]]
            assert len(ch) == 8 and ch["count"].sum() == nd["count"]
            assert np.all(ch["key"] == (nd["key"] << np.uint64(3)) + np.arange(8, dtype=np.uint64))
            starts = np.concatenate([[nd["begin"]], nd["begin"] + np.cumsum(ch["count"])[:-1]])
            assert np.array_equal(ch["begin"][ch["count"] > 0], starts[ch["count"] > 0])
            assert np.all(nd["child"] > i)  # DFS pre-order
        elif nd["type"] == O.OR_LEAF:
            assert 1 <= nd["count"] <= 12
        else:
            assert nd["count"] == 0
    # buckets tile the particles in order
    b = nodes[P.bucket_nodes]
    assert b["begin"][0] == 0 and np.all(b["begin"][1:] == b["begin"][:-1] + b["count"][:-1])
    assert b["count"].sum() == len(P.x)
    # DFS order == ascending left-aligned key among disjoint buckets
    la = left_aligned(b["key"])
    assert np.all(la[1:] > la[:-1])


# ---------------------------------------------------------------- moments

def test_moments_closed_forms():
    P = _small_pipeline(4000, 10)
    x = np.stack([P.x, P.y, P.z], 1)
    for nd, d in zip(P.nodes, P.data):
        if nd["count"] == 0:
            assert d["mass"] == 0 and d["radius"] == 0 and np.all(np.isinf(d["lo"]))
            continue
        sl = slice(nd["begin"], nd["begin"] + nd["count"])
        xm, mm, em = x[sl], P.m[sl], P.eps[sl]
        M = mm.sum()
        cm = (mm[:, None] * xm).sum(0) / M
        r = xm - cm
        Q = np.einsum("i,ij,ik->jk", mm, r, r)
        q = np.array([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]])
        assert np.isclose(d["mass"], M, rtol=1e-13)
        assert np.allclose(d["cm"], cm, rtol=0, atol=1e-13 * (1 + np.abs(cm).max()))
        scale = M * (np.abs(r).max() ** 2 + 1e-300)
        assert np.allclose(d["q"], q, rtol=0, atol=1e-11 * scale)
        assert np.isclose(d["soft"], (mm * em).sum() / M, rtol=1e-13)
        assert np.array_equal(d["lo"], xm.min(0)) and np.array_equal(d["hi"], xm.max(0))
        assert d["count"] == nd["count"]
        if nd["count"] == 1:
            assert d["radius"] == 1.0
        elif nd["type"] == O.OR_LEAF:   # calculateRadiusBox
            assert np.isclose(d["radius"], 0.5 * np.linalg.norm(d["hi"] - d["lo"]), rtol=1e-15)
        else:                            # calculateRadiusFarthestCorner
            far = np.maximum(d["cm"] - d["lo"], d["hi"] - d["cm"])
            assert np.isclose(d["radius"], np.linalg.norm(far), rtol=1e-15)


# ---------------------------------------------------------------- open / lists

def test_intersect_box_point_ties_and_cases():
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([1.0, 2.0, 3.0])
    assert O.intersect_box_point(lo, hi, [0.5, 1.0, 1.0], 0.0)        # inside
    assert O.intersect_box_point(lo, hi, [2.0, 1.0, 1.0], 1.0)        # tie dsq == rsq
    assert not O.intersect_box_point(lo, hi, [2.0, 1.0, 1.0], 0.999999)
    assert O.intersect_box_point(lo, hi, [-1.0, -1.0, 1.0], 2.0)      # corner region
    assert not O.intersect_box_point(lo, hi, [-1.0, -1.0, -1.0], 2.999)
    assert O.intersect_box_point(lo, hi, [-1.0, -1.0, -1.0], 3.0)


def test_lists_partition_sources_and_respect_opening():
    """Per bucket, the P-P leaves plus the P-C nodes cover every particle
    exactly once; each P-C node fails open() and every proper ancestor of a
    listed node passes it (brute force with the predicate)."""
    P = _small_pipeline(3000, 8, theta=0.6)
    pp_off, pp, pc_off, pc = P.lists()
    nodes, data = P.nodes, P.data
    gf = (4.0 / 3.0) / 0.6 ** 2

    def opens(src, tgt):
        if data[src]["count"] <= 6:
            return True
        return O.intersect_box_point(data[tgt]["lo"], data[tgt]["hi"], data[src]["cm"],
                                     data[src]["radius"] * data[src]["radius"] * gf)

    n = len(P.x)
    for bi, bn in enumerate(P.bucket_nodes[::7]):
        b = bi * 7
        cover = np.zeros(n, dtype=np.int64)
        for s in pp[pp_off[b]:pp_off[b + 1]]:
            assert nodes[s]["type"] == O.OR_LEAF
            cover[nodes[s]["begin"]:nodes[s]["begin"] + nodes[s]["count"]] += 1
        for s in pc[pc_off[b]:pc_off[b + 1]]:
            assert nodes[s]["type"] == O.OR_INTERNAL and not opens(s, bn)
            cover[nodes[s]["begin"]:nodes[s]["begin"] + nodes[s]["count"]] += 1
        assert np.all(cover == 1)
        for s in np.concatenate([pp[pp_off[b]:pp_off[b + 1]], pc[pc_off[b]:pc_off[b + 1]]]):
            a = nodes[s]["parent"]
            while a >= 0:
                assert opens(a, bn)
                a = nodes[a]["parent"]


def test_counters_follow_lists():
    P = _small_pipeline(3000, 8)
    pp_off, pp, pc_off, pc = P.lists()
    _, _, cnt = P.gravity()
    nt = P.nodes["count"][P.bucket_nodes]
    n_node = sum(int(nt[b]) * int(pc_off[b + 1] - pc_off[b]) for b in range(len(nt)))
    n_part = sum(int(nt[b]) * int(P.nodes["count"][pp[pp_off[b]:pp_off[b + 1]]].sum())
                 for b in range(len(nt)))
    assert int(cnt[0]) == n_node and int(cnt[1]) == n_part
    assert int(cnt[3]) == len(pc)


# ---------------------------------------------------------------- forces

def test_quadrupole_of_far_cluster():
    """node(): a compact cluster seen from far away: the quadrupole error falls
    as (size/distance)^3 (monopole would fall as ^2)."""
    rng = np.random.default_rng(9)
    clus = rng.standard_normal((64, 3)) * [0.3, 0.1, 0.05]
    errs = []
    for dist in (8.0, 16.0, 32.0):
        tgt = np.array([[dist, 0.0, 0.0], [dist, 0.001, 0.0]])
        pos = np.concatenate([clus, tgt])
        m = np.concatenate([np.full(64, 1.0 / 64), [1e-12, 1e-12]])
        P = O.Pipeline(pos, m, np.full(66, 1e-6), bucket=2, theta=1.0)
        acc, _, cnt = P.gravity_unsorted()
        ref = O.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, np.full(66, 1e-6), [64])
        errs.append(np.linalg.norm(acc[64] - ref[0]) / np.linalg.norm(ref[0]))
        assert cnt[0] > 0  # the target used the node expansion
    assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5


def test_tree_force_error_on_reference_fixture():
    """Full oracle force on lambs 30K (B=12, theta=0.7, quadrupole) against
    direct.acc: the reference states RMS .0008 / max .03 for its (more
    accurate) hexadecapole build; the quadrupole error is a few 1e-3."""
    pos, m, e, ref = _lambs()
    P = O.Pipeline(pos, m, e, bucket=12, theta=0.7)
    acc, _, _ = P.gravity_unsorted()
    rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert 5e-4 < np.sqrt(np.mean(rel ** 2)) < 5e-3
    assert rel.max() < 0.1


def test_kick_drift_closed_form():
    rng = np.random.default_rng(0)
    x, v, a = (rng.standard_normal((50, 3)) for _ in range(3))
    dt = 1e-3
    xn, vn = O.kick_drift(x, v, a, dt)
    assert np.allclose(vn, v + a * dt, rtol=0, atol=1e-15)
    assert np.allclose(xn, x + (v + a * dt) * dt, rtol=0, atol=1e-15)


# ---------------------------------------------------------------- adjoint

def _vjp_set(n=36, seed=5):
    """Small cluster whose pair distances cover the three SPLINE branches
    (u < 1, 1 <= u < 2, beyond 2h) for softening 0.15."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, (n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    e = np.full(n, 0.15)
    r = np.linalg.norm(pos[:, None] - pos[None], axis=2)[np.triu_indices(n, 1)]
    twoh = 0.3
    assert (r < twoh / 2).any() and ((r >= twoh / 2) & (r < twoh)).any() and (r >= twoh).any()
    return pos, m, e, rng.standard_normal((n, 3))


def _dot_acc(pos, m, e, w):
    acc = O.direct(pos[:, 0], pos[:, 1], pos[:, 2], m, e, np.arange(len(pos)))
    return float(np.sum(acc * w))


def test_direct_vjp_matches_finite_differences():
    """or_direct_vjp against central differences of or_direct (h = 1e-6) for
    every position coordinate and every mass."""
    pos, m, e, w = _vjp_set()
    g, gm = O.direct_vjp(pos, m, e, w)
    h = 1e-6
    fd = np.zeros_like(pos)
    for i in range(len(pos)):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            fd[i, k] = (_dot_acc(p1, m, e, w) - _dot_acc(p2, m, e, w)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-6 * np.max(np.abs(g))
    fdm = np.zeros(len(m))
    for j in range(len(m)):
        m1, m2 = m.copy(), m.copy()
        hm = 1e-6 * m[j]
        m1[j] += hm
        m2[j] -= hm
        fdm[j] = (_dot_acc(pos, m1, e, w) - _dot_acc(pos, m2, e, w)) / (2 * hm)
    assert np.max(np.abs(fdm - gm)) < 1e-7 * np.max(np.abs(gm))


def test_direct_vjp_invariants():
    """Translation invariance (sum_j g_j = 0) and degree-1 homogeneity in the
    masses (sum_j m_j gm_j = w . a)."""
    pos, m, e, w = _vjp_set(seed=9)
    g, gm = O.direct_vjp(pos, m, e, w)
    scale = np.abs(g).max()
    assert np.abs(g.sum(0)).max() < 1e-12 * scale * len(pos)
    wa = _dot_acc(pos, m, e, w)
    assert abs(np.dot(m, gm) - wa) < 1e-12 * abs(wa) * len(pos)


# ---------------------------------------------------------------- HEXADECAPOLE

def _stf_numpy(pos, m, cm):
    """Trace-free moments of orders 2-4 by tensor algebra (einsum over the
    full tensors, trace removal with Kronecker deltas), in the 22-value FMOMR
    order -- an independent route from the oracle's component formulas."""
    d = pos - cm
    I = np.eye(3)
    T2 = np.einsum("n,ni,nj->ij", m, d, d)
    T3 = np.einsum("n,ni,nj,nk->ijk", m, d, d, d)
    T4 = np.einsum("n,ni,nj,nk,nl->ijkl", m, d, d, d, d)
    S2 = T2 - np.trace(T2) / 3 * I
    V = np.einsum("ijj->i", T3)
    S3 = T3 - (np.einsum("ij,k->ijk", I, V) + np.einsum("ik,j->ijk", I, V)
               + np.einsum("jk,i->ijk", I, V)) / 5
    T = np.einsum("iikl->kl", T4)
    s = np.trace(T)
    S4 = (T4 - (np.einsum("ij,kl->ijkl", I, T) + np.einsum("ik,jl->ijkl", I, T)
                + np.einsum("il,jk->ijkl", I, T) + np.einsum("jk,il->ijkl", I, T)
                + np.einsum("jl,ik->ijkl", I, T) + np.einsum("kl,ij->ijkl", I, T)) / 7
          + s / 35 * (np.einsum("ij,kl->ijkl", I, I) + np.einsum("ik,jl->ijkl", I, I)
                      + np.einsum("il,jk->ijkl", I, I)))
    assert np.allclose(np.einsum("iikl->kl", S4), 0, atol=1e-12 * np.abs(T4).max())
    x, y, z = 0, 1, 2
    return np.array([m.sum(), S2[x, x], S2[y, y], S2[x, y], S2[x, z], S2[y, z],
                     S3[x, x, x], S3[x, y, y], S3[x, x, y], S3[y, y, y], S3[x, x, z],
                     S3[y, y, z], S3[x, y, z], S4[x, x, x, x], S4[x, y, y, y], S4[x, x, x, y],
                     S4[y, y, y, y], S4[x, x, x, z], S4[y, y, y, z], S4[x, x, y, y],
                     S4[x, x, y, z], S4[x, y, y, z]])


def _cluster(seed=5, n=60):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((n, 3)) * [0.12, 0.08, 0.05] + [0.3, -0.2, 0.1]
    m = rng.uniform(0.5, 1.5, n)
    cm = (m[:, None] * pos).sum(0) / m.sum()
    return pos, m, cm


def _hexa_of(pos, m, cm):
    d = pos - cm
    q = np.array([(m * d[:, a] * d[:, b]).sum() for a, b in
                  ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))])
    return O.hexa_reduce(m.sum(), q, O.hexa_raw_moments(pos, m, cm))


def test_hexa_reduced_moments_match_tensor_algebra():
    pos, m, cm = _cluster()
    h = _hexa_of(pos, m, cm)
    ref = _stf_numpy(pos, m, cm)
    scale = m.sum() * np.abs(pos - cm).max() ** np.array([0] + [2] * 5 + [3] * 7 + [4] * 9)
    assert np.all(np.abs(h - ref) <= 1e-14 * scale)


def test_hexa_potential_matches_legendre_series():
    """pot(r) = -sum_j m_j sum_{l=0,2,3,4} |d_j|^l P_l(cos g) / r^(l+1) (the
    dipole vanishes about the centre of mass): pins the (2l-1)!!/l! weights."""
    from numpy.polynomial import legendre
    pos, m, cm = _cluster()
    h = _hexa_of(pos, m, cm)
    d = pos - cm
    for r in ([1.0, 0.3, -0.2], [0.2, -0.7, 0.9], [-2.0, 1.0, 0.5]):
        r = np.array(r)
        rn = np.linalg.norm(r)
        dn = np.linalg.norm(d, axis=1)
        c = d @ r / (dn * rn)
        ref = 0.0
        for l in (0, 2, 3, 4):
            ref -= (m * dn ** l * legendre.legval(c, [0] * l + [1])).sum() / rn ** (l + 1)
        pot, _ = O.hexa_eval(h, r)
        assert abs(pot - ref) <= 1e-14 * abs(ref)


def test_hexa_acceleration_is_minus_gradient_of_potential():
    pos, m, cm = _cluster()
    h = _hexa_of(pos, m, cm)
    for r in ([1.0, 0.3, -0.2], [-0.4, 0.9, 0.7]):
        r = np.array(r)
        _, acc = O.hexa_eval(h, r)
        g = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-5
            g[k] = -(O.hexa_eval(h, r + e)[0] - O.hexa_eval(h, r - e)[0]) / 2e-5
        assert np.abs(acc - g).max() <= 1e-8 * np.abs(acc).max()


def test_hexa_truncation_error_falls_as_fifth_order():
    """Against the exact point-mass sum, the acceleration error of the order-4
    expansion falls as (size/distance)^5 relative (x32 per doubling); the
    quadrupole truncation would fall as ^3."""
    pos, m, cm = _cluster()
    h = _hexa_of(pos, m, cm)
    u = np.array([0.6, -0.48, 0.64])
    errs = []
    for dist in (1.5, 3.0, 6.0):
        r = u * dist
        _, acc = O.hexa_eval(h, r)
        dd = r - (pos - cm)
        ref = -(m[:, None] * dd / (np.linalg.norm(dd, axis=1) ** 3)[:, None]).sum(0)
        errs.append(np.linalg.norm(acc - ref) / np.linalg.norm(ref))
    assert 20 < errs[0] / errs[1] < 50 and 20 < errs[1] / errs[2] < 50


def test_hexa_tree_moments_match_closed_form_per_node():
    """Leaf sums and the binomial shifts of the upward pass give, for every
    node, the trace-free moments of all its particles about its centre of
    mass (tensor algebra on the particles)."""
    pos, _, m = ics.plummer(2000, seed=21)
    P = O.Pipeline(pos, m, np.full(2000, 0.01), bucket=8, theta=0.7, hexadecapole=True)
    px = np.stack([P.x, P.y, P.z], axis=1)
    for i in range(len(P.nodes)):
        b, c = int(P.nodes[i]["begin"]), int(P.nodes[i]["count"])
        if c == 0:
            continue
        cm = P.data[i]["cm"]
        ref = _stf_numpy(px[b:b + c], P.m[b:b + c], cm)
        ext = np.abs(px[b:b + c] - cm).max()
        scale = P.m[b:b + c].sum() * ext ** np.array([0] + [2] * 5 + [3] * 7 + [4] * 9)
        assert np.all(np.abs(P.hexa[i] - ref) <= 1e-12 * scale + 1e-300), i


def test_hexa_lists_respect_the_hexadecapole_opening():
    """GravityVisitor::open, HEXADECAPOLE branch, by brute force: with large
    softening the softening/monopole branch decides many nodes; every listed
    P-C node fails open() and all its ancestors pass it; the sets differ from
    the quadrupole build's."""
    pos, _, m = ics.plummer(3000, seed=4)
    eps = np.full(3000, 0.15)
    theta = 0.6
    P = O.Pipeline(pos, m, eps, bucket=8, theta=theta, hexadecapole=True)
    Q = O.Pipeline(pos, m, eps, bucket=8, theta=theta)
    pp_off, pp, pc_off, pc = P.lists()
    nodes, data = P.nodes, P.data
    gf = (4.0 / 3.0) / theta ** 2
    gm = np.sqrt(4.0 / 3.0) / theta ** 4

    def soft_meet(a_cm, a_soft, b):
        # openSoftening(source = a, target = b): spheres, then b's box
        r1 = 2.0 * a_soft
        dsq = ((a_cm - data[b]["cm"]) ** 2).sum()
        if dsq <= (r1 + 2.0 * data[b]["soft"]) ** 2:
            return True
        return O.intersect_box_point(data[b]["lo"], data[b]["hi"], a_cm, r1 * r1)

    def opens(src, tgt):
        if data[src]["count"] <= 6:
            return True
        lo, hi, cm = data[tgt]["lo"], data[tgt]["hi"], data[src]["cm"]
        if O.intersect_box_point(lo, hi, cm, data[src]["radius"] ** 2 * gf):
            return True
        if not soft_meet(data[tgt]["cm"], data[tgt]["soft"], src):
            return False
        return O.intersect_box_point(lo, hi, cm, (data[src]["radius"] * gm) ** 2)

    n = len(P.x)
    for bi in range(0, len(P.bucket_nodes), 5):
        bn = P.bucket_nodes[bi]
        cover = np.zeros(n, dtype=np.int64)
        for s in pp[pp_off[bi]:pp_off[bi + 1]]:
            cover[nodes[s]["begin"]:nodes[s]["begin"] + nodes[s]["count"]] += 1
        for s in pc[pc_off[bi]:pc_off[bi + 1]]:
            assert nodes[s]["type"] == O.OR_INTERNAL and not opens(s, bn)
            cover[nodes[s]["begin"]:nodes[s]["begin"] + nodes[s]["count"]] += 1
        assert np.all(cover == 1)
        for s in np.concatenate([pp[pp_off[bi]:pp_off[bi + 1]], pc[pc_off[bi]:pc_off[bi + 1]]]):
            a = nodes[s]["parent"]
            while a >= 0:
                assert opens(a, bn)
                a = nodes[a]["parent"]
    assert len(pc) != len(Q.lists()[3])


def test_hexa_tree_force_meets_reference_accuracy_statement():
    """The reference's accuracy test (tests/gravity/acc_test.sh:5) states, for
    its HEXADECAPOLE build on this fixture with its defaults (bucket 12,
    theta 0.7): RMS relative force error .0008 and maximum .03 against
    direct.acc.  The quadrupole build misses the maximum."""
    pos, m, e, ref = _lambs()
    P = O.Pipeline(pos, m, e, bucket=12, theta=0.7, hexadecapole=True)
    acc, _, _ = P.gravity_unsorted()
    r = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.sqrt(np.mean(r * r)) < 8e-4 and r.max() < 0.03
    Q = O.Pipeline(pos, m, e, bucket=12, theta=0.7)
    acc_q, _, _ = Q.gravity_unsorted()
    rq = np.linalg.norm(acc_q - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rq.max() > 0.03 and np.sqrt(np.mean(r * r)) < np.sqrt(np.mean(rq * rq)) / 3


# ---------------------------------------------------------------- periodic replicas

def test_replica_sum_matches_direct_sum_over_replicated_images():
    """ExMain::traversalFn with periodic (examples/Gravity.C:35-46): the tree
    sum over the 27 replica offsets (nReplicas = 1) against the direct sum of
    the leaf() kernel over an explicitly replicated particle set (targets in
    the central copy): the offsets, their sign and the self-image pairs
    (rsq = period^2, included) agree up to the tree's truncation error."""
    rng = np.random.default_rng(21)
    n = 400
    pos = rng.uniform(0.0, 1.0, (n, 3))
    m = np.full(n, 1.0 / n)
    e = np.full(n, 0.01)
    period = (1.0, 1.0, 1.0)
    P = O.Pipeline(pos, m, e, bucket=8, theta=0.3)
    acc, _, cnt = P.gravity_replicas(1, period)
    reps = [np.array([X, Y, Z], dtype=np.float64) for X in (-1, 0, 1) for Y in (-1, 0, 1)
            for Z in (-1, 0, 1)]
    allpos = np.concatenate([pos + r for r in reps])
    allm = np.tile(m, len(reps))
    alle = np.tile(e, len(reps))
    centre = 13  # (0, 0, 0) in the X-outer order
    t = np.arange(centre * n, (centre + 1) * n)
    ref = O.direct(allpos[:, 0], allpos[:, 1], allpos[:, 2], allm, alle, t)
    rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.sqrt(np.mean(rel ** 2)) < 2e-3
    # the replicas add node() and leaf() calls to the single box's
    _, _, c0 = P.gravity()
    assert cnt[0] > c0[0] and cnt[1] > c0[1] and cnt[3] > c0[3]


def _ewald_setup(seed=3, n=300):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 1.0, (n, 3))
    pos[: n // 3] = 0.4 + 0.05 * rng.standard_normal((n // 3, 3))
    m = rng.uniform(0.5, 1.5, n)
    M = m.sum()
    cm = (m[:, None] * pos).sum(0) / M
    d = pos - cm
    q = np.array([(m * d[:, a] * d[:, b]).sum() for a, b in
                  ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))])
    return pos, M, cm, q


@pytest.mark.parametrize("n_rep", [0, 1])
def test_ewald_is_independent_of_the_split(n_rep):
    """The Ewald split parameter alpha only moves terms between the real-space
    (erfc / -erf) and Fourier sums: alpha = 2/L with the reference's cuts and
    alpha = 3/L with wider cuts give the same correction (a wrong coefficient,
    sign or hole rule in either sum breaks this)."""
    pos, M, cm, q = _ewald_setup()
    t = pos[::30]
    a1, p1 = O.ewald(t, M, cm, q, 1.0, n_rep)
    # (the two-term series below r^2 = 1.2e-3 L^2 is accurate to ~(alpha r)^4,
    # so alpha stays near 2/L; the reference's cuts truncate at ~1e-7)
    a2, p2 = O.ewald(t, M, cm, q, 1.0, n_rep, a_fac=2.0, ew_cut=4.0, h_cut=5.0)
    a3, p3 = O.ewald(t, M, cm, q, 1.0, n_rep, a_fac=1.5, ew_cut=4.0, h_cut=3.0)
    scale = np.abs(a1).max()
    assert np.abs(a1 - a2).max() < 1e-6 * scale and np.abs(a1 - a3).max() < 1e-6 * scale
    assert np.abs(p1 - p2).max() < 1e-6 * np.abs(p1).max()
    assert np.abs(p1 - p3).max() < 1e-6 * np.abs(p1).max()


def test_ewald_small_r_series_is_continuous():
    """Below r^2 = 1.2e-3 L^2 from the root's centre the closed forms give way
    to their two-term Taylor series; both sides of the switch agree to the
    series' truncation, ~(alpha r)^4 of the leading terms (a wrong
    coefficient would open a gap of order 1e-2)."""
    _, M, cm, q = _ewald_setup()
    r0 = np.sqrt(1.2e-3)
    u = np.array([0.6, 0.48, 0.64])
    pts = np.array([cm + u * r0 * (1 - 1e-7), cm + u * r0 * (1 + 1e-7)])
    a, p = O.ewald(pts, M, cm, q, 1.0, 1)
    assert np.abs(a[0] - a[1]).max() < 1e-5 * np.abs(a).max()
    assert abs(p[0] - p[1]) < 1e-6 * abs(p[0])


def test_ewald_of_a_lone_mass_is_symmetric():
    """One mass in the box: the lattice of its images pulls it nowhere (each
    image has a mirror image), and points mirrored through it feel opposite
    forces."""
    cm = np.array([0.3, 0.7, 0.5])
    q = np.zeros(6)
    a, _ = O.ewald(cm[None, :], 2.0, cm, q, 1.0, 1)
    assert np.abs(a).max() < 1e-12
    d = np.array([0.11, -0.07, 0.23])
    a2, p2 = O.ewald(np.array([cm + d, cm - d]), 2.0, cm, q, 1.0, 1)
    assert np.abs(a2[0] + a2[1]).max() < 1e-12 * np.abs(a2).max()
    assert abs(p2[0] - p2[1]) < 1e-12 * abs(p2[0])


def test_periodic_build_approaches_the_exact_periodic_force():
    """Replicas (nReplicas = 1, tree at theta 0.1) + LeafEwaldFn's correction
    from the root's monopole + quadrupole against the exact periodic force:
    the replicated direct sum plus each source's own point-mass Ewald lattice.
    The root expansion leaves ~1% (RMS 6e-3 here); the replicas alone miss by
    a factor of a few, so the correction is what makes the sum periodic."""
    rng = np.random.default_rng(4)
    n, L = 64, 1.0
    pos = rng.uniform(0, 1, (n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    e = np.full(n, 1e-7)
    P = O.Pipeline(pos, m, e, bucket=4, theta=0.1)
    acc, _, _ = P.gravity_replicas(1, (L, L, L), ewald=True)
    acc0, _, _ = P.gravity_replicas(1, (L, L, L), ewald=False)
    reps = [np.array([X, Y, Z], float) for X in (-1, 0, 1) for Y in (-1, 0, 1) for Z in (-1, 0, 1)]
    allpos = np.concatenate([pos + r for r in reps])
    exact = O.direct(allpos[:, 0], allpos[:, 1], allpos[:, 2], np.tile(m, 27), np.tile(e, 27),
                     np.arange(13 * n, 14 * n))
    for j in range(n):
        a, _ = O.ewald(pos, m[j], pos[j], np.zeros(6), L, 1)
        exact += a
    rel = np.linalg.norm(acc - exact, axis=1) / np.linalg.norm(exact, axis=1)
    rel0 = np.linalg.norm(acc0 - exact, axis=1) / np.linalg.norm(exact, axis=1)
    assert np.sqrt(np.mean(rel ** 2)) < 2e-2 and rel.max() < 5e-2
    assert np.sqrt(np.mean(rel0 ** 2)) > 0.5
    # the exact periodic forces conserve momentum
    assert np.abs((m[:, None] * exact).sum(0)).max() < 1e-12


def _ewald_hexa_setup(seed=3, n=300):
    """A lopsided cluster whose root has sizeable octupole and hexadecapole
    moments: the reduced moments from the oracle's own per-node moments."""
    pos, M, cm, q = _ewald_setup(seed, n)
    m = np.random.default_rng(seed).uniform(0.5, 1.5, n)
    m *= M / m.sum()
    raw = O.hexa_raw_moments(pos, m, cm)
    h = O.hexa_reduce(M, q, raw)[1:]
    return pos, M, cm, q, h


def test_ewald_hexa_reduces_to_the_quadrupole_form():
    """With the octupole and hexadecapole set to zero the fourth-order form is
    the quadrupole form fed the trace-free second moments (whose traces, the
    Q2 terms, then vanish): the two restatements agree to rounding."""
    pos, M, cm, q, h = _ewald_hexa_setup()
    h2 = np.zeros(21)
    h2[:5] = h[:5]
    # the trace-free second moments as the quadrupole form's six components
    xx, yy, xy, xz, yz = h[:5]
    qtf = np.array([xx, xy, xz, yy, yz, -(xx + yy)])
    t = pos[::25]
    a1, p1 = O.ewald_hexa(t, M, cm, h2, 1.0, 1)
    a2, p2 = O.ewald(t, M, cm, qtf, 1.0, 1)
    assert np.abs(a1 - a2).max() < 1e-13 * np.abs(a2).max()
    assert np.abs(p1 - p2).max() < 1e-13 * np.abs(p2).max()


@pytest.mark.parametrize("n_rep", [0, 1])
def test_ewald_hexa_is_independent_of_the_split(n_rep):
    """The third- and fourth-order terms of the real-space sum (g4, g5) and of
    the Fourier sum (the sine and h^4 coefficients) must balance for every
    alpha: alpha = 2/L with the reference's cuts, with wider cuts and at
    1.5/L agree (a wrong coefficient in either sum breaks this), and the
    higher moments change the correction (they are not dead terms)."""
    pos, M, cm, q, h = _ewald_hexa_setup()
    t = pos[::30]
    a1, p1 = O.ewald_hexa(t, M, cm, h, 1.0, n_rep)
    a2, p2 = O.ewald_hexa(t, M, cm, h, 1.0, n_rep, a_fac=2.0, ew_cut=4.0, h_cut=5.0)
    a3, p3 = O.ewald_hexa(t, M, cm, h, 1.0, n_rep, a_fac=1.5, ew_cut=4.0, h_cut=3.0)
    scale = np.abs(a1).max()
    assert np.abs(a1 - a2).max() < 1e-6 * scale and np.abs(a1 - a3).max() < 1e-6 * scale
    assert np.abs(p1 - p2).max() < 1e-6 * np.abs(p1).max()
    assert np.abs(p1 - p3).max() < 1e-6 * np.abs(p1).max()
    hq = h.copy()
    hq[5:] = 0.0
    aq, _ = O.ewald_hexa(t, M, cm, hq, 1.0, n_rep)
    assert np.abs(a1 - aq).max() > 1e-4 * scale


def test_ewald_libm_perturbation_knob():
    """The knob the GPU parity test uses to measure the libm sensitivity: at
    (0, 0) the results are bit for bit the unperturbed ones (also after a
    perturbed call), and a one-ulp change moves the fourth-order form's hole
    terms just outside the series radius by far more than an ulp, and far
    more than it moves the quadrupole form's."""
    pos, M, cm, q, h = _ewald_hexa_setup()
    r0 = np.sqrt(1.2e-3)
    pts = np.array([cm + np.array([0.6, 0.48, 0.64]) * r0 * 1.5, pos[0]])
    a0, p0 = O.ewald_hexa(pts, M, cm, h, 1.0, 1)
    q0, _ = O.ewald(pts, M, cm, q, 1.0, 1)
    O.set_ewald_libm_perturbation(1, 1)
    try:
        a1, _ = O.ewald_hexa(pts, M, cm, h, 1.0, 1)
        q1, _ = O.ewald(pts, M, cm, q, 1.0, 1)
    finally:
        O.set_ewald_libm_perturbation(0, 0)
    a2, p2 = O.ewald_hexa(pts, M, cm, h, 1.0, 1)
    assert np.array_equal(a0, a2) and np.array_equal(p0, p2)
    d_hexa = np.abs(a1[0] - a0[0]).max() / np.abs(a0[0]).max()
    d_quad = np.abs(q1[0] - q0[0]).max() / np.abs(q0[0]).max()
    assert d_hexa > 1e-11 and d_hexa > 30.0 * d_quad


def test_ewald_hexa_small_r_series_is_continuous():
    _, M, cm, q, h = _ewald_hexa_setup()
    r0 = np.sqrt(1.2e-3)
    u = np.array([0.6, 0.48, 0.64])
    pts = np.array([cm + u * r0 * (1 - 1e-7), cm + u * r0 * (1 + 1e-7)])
    a, p = O.ewald_hexa(pts, M, cm, h, 1.0, 1)
    assert np.abs(a[0] - a[1]).max() < 1e-5 * np.abs(a).max()
    assert abs(p[0] - p[1]) < 1e-6 * abs(p[0])


def test_periodic_hexadecapole_build_approaches_the_exact_periodic_force():
    """The HEXADECAPOLE build's replicas + its fourth-order lattice correction
    against the exact periodic force (as for the quadrupole build): closer
    than the quadrupole-form correction of the same root."""
    rng = np.random.default_rng(4)
    n, L = 64, 1.0
    pos = rng.uniform(0, 1, (n, 3))
    pos[:24] = 0.3 + 0.05 * rng.standard_normal((24, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    e = np.full(n, 1e-7)
    P = O.Pipeline(pos, m, e, bucket=4, theta=0.1, hexadecapole=True)
    acc, _, _ = P.gravity_replicas(1, (L, L, L), ewald=True)
    acc0, _, _ = P.gravity_replicas(1, (L, L, L), ewald=False)
    root = P.data[0]
    xyz = np.stack([P.x, P.y, P.z], axis=1)
    eq, _ = O.ewald(xyz, root["mass"], root["cm"], root["q"], L, 1)
    accq = acc0.copy()
    accq[P.perm] += eq
    reps = [np.array([X, Y, Z], float) for X in (-1, 0, 1) for Y in (-1, 0, 1) for Z in (-1, 0, 1)]
    allpos = np.concatenate([pos + r for r in reps])
    exact = O.direct(allpos[:, 0], allpos[:, 1], allpos[:, 2], np.tile(m, 27), np.tile(e, 27),
                     np.arange(13 * n, 14 * n))
    for j in range(n):
        a, _ = O.ewald(pos, m[j], pos[j], np.zeros(6), L, 1)
        exact += a
    rms = lambda a: float(np.sqrt(np.mean(  # noqa: E731
        (np.linalg.norm(a - exact, axis=1) / np.linalg.norm(exact, axis=1)) ** 2)))
    print("periodic hexadecapole vs exact: rms %.3g (quadrupole-form correction %.3g, "
          "replicas alone %.3g)" % (rms(acc), rms(accq), rms(acc0)))
    assert rms(acc) < rms(accq) and rms(acc) < 1e-2


def test_walk_only_counters_equal_the_gravity_counters():
    """or_traversal_counters runs or_gravity's walk without the forces: the
    same four counters on a Plummer and a clumped input."""
    for pos, bucket, theta in ((ics.plummer(20000, seed=3)[0], 12, 0.7),
                               (ics.halo_mixture(20000, n_halos=8, seed=5)[0], 32, 0.5)):
        n = len(pos)
        P = O.Pipeline(pos, np.full(n, 1.0 / n), np.full(n, 0.01), bucket=bucket, theta=theta)
        _, _, c = P.gravity()
        assert np.array_equal(O.traversal_counters(P.nodes, P.data, theta), c)

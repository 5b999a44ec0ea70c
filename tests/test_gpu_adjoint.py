"""GPU tests of the hand-written adjoint (through the C ABI).

* the bucket groups and member-masked lists the tree adjoint walks against
  the per-bucket lists;
* direct-sum VJP against the oracle's (itself pinned by finite differences);
* tree VJP against central finite differences of the GPU tree force, the
  translation-invariance and mass-homogeneity identities, and the direct-sum
  VJP (within the tree approximation);
* leapfrog adjoint (loss sum m |x_T - x_0|^2) against central finite
  differences (h = 1e-6, 64 seeded coordinates) in direct and tree mode,
  checkpoint schedules agreeing, and the tree gradient against the FP64
  direct-sum gradient;
* the torch.autograd.Function over the C ABI.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics

pytestmark = pytest.mark.gpu

H = 1e-6


def _driver(bucket=12):
    from paratreet_amd.gravity import Configuration, Driver
    return Driver(Configuration(max_particles_per_leaf=bucket))


def _plummer(n, seed, eps):
    pos, vel, m = ics.plummer(n, seed=seed)
    return pos, vel, m, np.full(n, eps)


# ---------------------------------------------------------------- single VJP

def _tree_acc(d, pos, m, e, theta):
    d.load(pos, m, e)
    d.compute_forces(theta=theta)
    return d.accelerations()[0]


@pytest.mark.parametrize("case", ["plummer20k_b16", "halo12k_b32", "plummer300_b12"])
def test_group_lists_match_bucket_lists(case):
    """The adjoint walks groups of consecutive buckets: the groups partition
    the buckets in order (at most 8 buckets and 16 targets unless one bucket
    is larger), no node appears twice in a group's list, and for every member
    bucket the entries carrying its bit are exactly its pt_traverse lists."""
    from paratreet_amd.gravity import GravityVisitor
    if case == "plummer20k_b16":
        pos, _, m, e = _plummer(20000, 11, 0.005)
        bucket, theta = 16, 0.7
    elif case == "halo12k_b32":
        pos, _, m = ics.halo_mixture(12000, n_halos=6, seed=8)
        e = np.full(len(pos), 0.05)
        bucket, theta = 32, 0.9
    else:
        pos, _, m, e = _plummer(300, 3, 0.01)
        bucket, theta = 12, 0.5
    d = _driver(bucket)
    d.load(pos, m, e)
    d.remake_universe()
    d.assign_keys()
    d.sort()
    d._check(d.L.pt_build_tree(d.ctx))
    d._check(d.L.pt_upward(d.ctx))
    d.traverse(GravityVisitor((0.0, 0.0, 0.0), theta))
    pp_off, pp, pc_off, pc = d.lists()
    gfirst, gnb, gpp_off, gpp, gppm, gpc_off, gpc, gpcm = d.group_lists(theta)
    counts = np.asarray(d.nodes()["count"])[d.buckets()]
    nb = len(counts)
    assert len(gnb) == -(-nb // 8) * 8
    b_next = 0
    for s in range(len(gnb)):
        k = int(gnb[s])
        if k == 0:
            assert gpp_off[s + 1] == gpp_off[s] and gpc_off[s + 1] == gpc_off[s]
            continue
        assert gfirst[s] == b_next and 1 <= k <= 8
        assert k == 1 or counts[b_next:b_next + k].sum() <= 16
        for off, ids, msk, boff, bids in ((gpc_off, gpc, gpcm, pc_off, pc),
                                          (gpp_off, gpp, gppm, pp_off, pp)):
            a, z = off[s], off[s + 1]
            gi, gm = ids[a:z], msk[a:z].astype(np.int64)
            assert len(np.unique(gi)) == len(gi)
            assert np.all(gm > 0) and np.all(gm < (1 << k))
            for j in range(k):
                b = b_next + j
                mine = np.sort(gi[(gm >> j) & 1 == 1])
                assert np.array_equal(mine, np.sort(bids[boff[b]:boff[b + 1]])), (s, j)
        b_next += k
    assert b_next == nb
    d.close()


def test_direct_vjp_matches_oracle():
    from paratreet_amd.gravity import FORCE_DIRECT
    pos, _, m, e = _plummer(3000, 21, 0.02)
    w = np.random.default_rng(1).standard_normal(pos.shape)
    d = _driver()
    d.load(pos, m, e)
    g, gm = d.force_vjp(w, mode=FORCE_DIRECT)
    go, gmo = O.direct_vjp(pos, m, e, w)
    assert np.linalg.norm(g - go) <= 1e-12 * np.linalg.norm(go)
    assert np.linalg.norm(gm - gmo) <= 1e-12 * np.linalg.norm(gmo)


@pytest.mark.parametrize("case", ["plummer5k", "halo8k"])
def test_tree_vjp_finite_differences(case):
    """(da/dx)^T w and (da/dm)^T w on the traversal lists against central
    differences of the GPU FP64 tree force (the lists are the same on both
    sides of each difference for all but rare threshold cases)."""
    from paratreet_amd.gravity import FORCE_TREE
    if case == "plummer5k":
        pos, _, m, e = _plummer(5000, 22, 0.01)
        bucket, theta = 12, 0.7
    else:
        pos, _, m = ics.halo_mixture(8000, n_halos=4, seed=5)
        e = np.full(len(pos), 0.05)
        bucket, theta = 16, 0.9
    rng = np.random.default_rng(2)
    w = rng.standard_normal(pos.shape)
    d = _driver(bucket)
    d.load(pos, m, e)
    g, gm = d.force_vjp(w, theta=theta, mode=FORCE_TREE)
    coords = rng.choice(pos.size, 48, replace=False)
    err = []
    for c in coords:
        i, k = divmod(int(c), 3)
        p1, p2 = pos.copy(), pos.copy()
        p1[i, k] += H
        p2[i, k] -= H
        da = _tree_acc(d, p1, m, e, theta) - _tree_acc(d, p2, m, e, theta)
        fd = np.sum(da * w) / (2 * H)
        err.append(abs(fd - g[i, k]) / np.abs(g).max())
    err = np.array(err)
    print("tree vjp FD rel err: median %.2e max %.2e" % (np.median(err), err.max()))
    assert np.median(err) < 1e-7
    assert (err < 1e-5).mean() >= 0.95
    merr = []
    for j in rng.choice(len(m), 12, replace=False):
        hm = 1e-6 * m[j]
        m1, m2 = m.copy(), m.copy()
        m1[j] += hm
        m2[j] -= hm
        da = _tree_acc(d, pos, m1, e, theta) - _tree_acc(d, pos, m2, e, theta)
        fd = np.sum(da * w) / (2 * hm)
        merr.append(abs(fd - gm[j]) / np.abs(gm).max())
    merr = np.array(merr)
    print("tree vjp mass FD rel err: median %.2e max %.2e" % (np.median(merr), merr.max()))
    assert np.median(merr) < 1e-7
    assert (merr < 1e-5).mean() >= 0.9


def test_tree_vjp_identities_and_direct():
    """Translation invariance: sum_j g_j = 0; homogeneity of degree 1 in the
    masses: sum_j m_j gm_j = w . a; and closeness to the direct-sum VJP."""
    from paratreet_amd.gravity import FORCE_DIRECT, FORCE_TREE
    pos, _, m, e = _plummer(30000, 23, 0.005)
    w = np.random.default_rng(3).standard_normal(pos.shape)
    d = _driver(16)
    a = _tree_acc(d, pos, m, e, 0.7)
    g, gm = d.force_vjp(w, theta=0.7, mode=FORCE_TREE)
    assert np.abs(g.sum(0)).max() < 1e-10 * np.abs(g).sum()
    wa = np.sum(w * a)
    assert abs(np.dot(m, gm) - wa) < 1e-11 * np.sum(np.abs(w * a))
    gd, gmd = d.force_vjp(w, mode=FORCE_DIRECT)
    rx = np.linalg.norm(g - gd) / np.linalg.norm(gd)
    rm = np.linalg.norm(gm - gmd) / np.linalg.norm(gmd)
    print("tree vs direct VJP: positions %.2e masses %.2e" % (rx, rm))
    assert rx < 2e-2 and rm < 2e-2


def test_tree_vjp_fp32c_against_fp64():
    """The FP32 adjoint (FP32 pairs on bucket-centre-shifted coordinates, FP64
    sums) against the FP64 adjoint on the same lists, which finite
    differences pin above; the identities hold to FP32 rounding."""
    from paratreet_amd import _lib
    from paratreet_amd.gravity import FORCE_TREE
    pos, _, m, e = _plummer(30000, 23, 0.005)
    w = np.random.default_rng(5).standard_normal(pos.shape)
    d = _driver(16)
    a = _tree_acc(d, pos, m, e, 0.7)
    g64, gm64 = d.force_vjp(w, theta=0.7, mode=FORCE_TREE)
    d.set_adjoint_precision(_lib.PT_FP32C)
    g32, gm32 = d.force_vjp(w, theta=0.7, mode=FORCE_TREE)
    d.set_adjoint_precision(_lib.PT_FP64)
    rx = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    rm = np.linalg.norm(gm32 - gm64) / np.linalg.norm(gm64)
    print("fp32c vs fp64 tree VJP: positions %.2e masses %.2e" % (rx, rm))
    assert rx < 1e-5 and rm < 1e-5
    assert np.abs(g32.sum(0)).max() < 1e-6 * np.abs(g32).sum()
    assert abs(np.dot(m, gm32) - np.sum(w * a)) < 1e-6 * np.sum(np.abs(w * a))
    # the setting is back to FP64 (equal up to the order of the atomic sums)
    g2, gm2 = d.force_vjp(w, theta=0.7, mode=FORCE_TREE)
    assert np.linalg.norm(g2 - g64) <= 1e-12 * np.linalg.norm(g64)
    assert np.linalg.norm(gm2 - gm64) <= 1e-12 * np.linalg.norm(gm64)
    with pytest.raises(_lib.PtError):
        d.set_adjoint_precision(7)


# ---------------------------------------------------------------- trajectory

STEPS, DT = 10, 1e-3


def _final(d, pos, vel, m, e, mode, theta=0.7):
    d.load(pos, m, e, vel=vel)
    d.leapfrog_forward(STEPS, DT, theta=theta, mode=mode)
    return d.state()[0]


def _loss_terms(xT, x0, m):
    return m * np.sum((xT - x0) ** 2, axis=1)


def _gradients(d, pos, vel, m, e, mode, theta=0.7, every=0):
    """dL/dx0, dL/dv0, dL/dm of L = sum m |x_T - x_0|^2 via the adjoint."""
    d.load(pos, m, e, vel=vel)
    d.leapfrog_forward(STEPS, DT, theta=theta, mode=mode, checkpoint_every=every)
    xT = d.state()[0]
    lam = 2.0 * m[:, None] * (xT - pos)
    gx, gv, gm = d.leapfrog_adjoint(lam_x=lam)
    return gx - lam, gv, gm + np.sum((xT - pos) ** 2, axis=1)


def _fd_check(d, pos, vel, m, e, mode, gx, gv, gm, ncoord, nmass, theta=0.7, seed=4):
    rng = np.random.default_rng(seed)
    errs = {"x": [], "v": [], "m": []}
    for c in rng.choice(2 * pos.size, ncoord, replace=False):
        which, c = divmod(int(c), pos.size)
        i, k = divmod(c, 3)
        if which == 0:
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += H
            p2[i, k] -= H
            l1 = _loss_terms(_final(d, p1, vel, m, e, mode, theta), p1, m)
            l2 = _loss_terms(_final(d, p2, vel, m, e, mode, theta), p2, m)
            ref, key = gx[i, k], "x"
        else:
            v1, v2 = vel.copy(), vel.copy()
            v1[i, k] += H
            v2[i, k] -= H
            l1 = _loss_terms(_final(d, pos, v1, m, e, mode, theta), pos, m)
            l2 = _loss_terms(_final(d, pos, v2, m, e, mode, theta), pos, m)
            ref, key = gv[i, k], "v"
        fd = np.sum(l1 - l2) / (2 * H)
        scale = np.abs(gx).max() if key == "x" else np.abs(gv).max()
        errs[key].append(abs(fd - ref) / scale)
    for j in rng.choice(len(m), nmass, replace=False):
        hm = 1e-6 * m[j]
        m1, m2 = m.copy(), m.copy()
        m1[j] += hm
        m2[j] -= hm
        l1 = _loss_terms(_final(d, pos, vel, m1, e, mode, theta), pos, m1)
        l2 = _loss_terms(_final(d, pos, vel, m2, e, mode, theta), pos, m2)
        fd = np.sum(l1 - l2) / (2 * hm)
        errs["m"].append(abs(fd - gm[j]) / np.abs(gm).max())
    return {k: np.array(v) for k, v in errs.items()}


def test_leapfrog_adjoint_direct_finite_differences():
    from paratreet_amd.gravity import FORCE_DIRECT
    pos, vel, m, e = _plummer(512, 31, 0.01)
    d = _driver()
    gx, gv, gm = _gradients(d, pos, vel, m, e, FORCE_DIRECT)
    errs = _fd_check(d, pos, vel, m, e, FORCE_DIRECT, gx, gv, gm, 64, 16)
    for k, v in errs.items():
        print("direct trajectory FD %s: median %.2e max %.2e" % (k, np.median(v), v.max()))
    # FD noise: positions carry ~1e-16 absolute rounding per step while the
    # loss sees x0 only through dt^2 (da/dx) h ~ 1e-11, so the position
    # differences resolve ~1e-6 of the gradient scale; v0 and m are better
    # conditioned
    assert errs["x"].max() < 1e-5
    assert errs["v"].max() < 1e-6
    assert errs["m"].max() < 1e-6


def test_leapfrog_adjoint_tree_finite_differences():
    """Config-4 shape at 20k particles: 10 steps dt = 1e-3, theta = 0.7."""
    from paratreet_amd.gravity import FORCE_TREE
    pos, vel, m, e = _plummer(20000, 32, 0.005)
    d = _driver(16)
    gx, gv, gm = _gradients(d, pos, vel, m, e, FORCE_TREE)
    errs = _fd_check(d, pos, vel, m, e, FORCE_TREE, gx, gv, gm, 64, 8)
    for k, v in errs.items():
        print("tree trajectory FD %s: median %.2e max %.2e" % (k, np.median(v), v.max()))
    for k, v in errs.items():
        assert np.median(v) < 1e-5, k
        assert (v < 1e-4).mean() >= 0.9, k


def test_checkpoint_schedules_agree():
    from paratreet_amd.gravity import FORCE_TREE
    pos, vel, m, e = _plummer(8000, 33, 0.005)
    d = _driver(16)
    ref = _gradients(d, pos, vel, m, e, FORCE_TREE, every=0)
    final = d.state()
    for every in (1, 3, 10):
        got = _gradients(d, pos, vel, m, e, FORCE_TREE, every=every)
        for a, b in zip(got, ref):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), every
    # the adjoint restores the final state
    assert np.array_equal(d.state()[0], final[0])


def test_tree_gradient_close_to_direct_gradient():
    from paratreet_amd.gravity import FORCE_DIRECT, FORCE_TREE
    pos, vel, m, e = _plummer(16384, 34, 0.005)
    d = _driver(16)
    t = _gradients(d, pos, vel, m, e, FORCE_TREE)
    r = _gradients(d, pos, vel, m, e, FORCE_DIRECT)
    errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(t, r)]
    print("tree vs direct gradient: x %.2e v %.2e m %.2e" % tuple(errs))
    assert max(errs) < 1e-2


def test_leapfrog_fp32c_forward_and_adjoint():
    """FP32C forward and FP32C adjoint against the FP64 forward and adjoint
    (config-5 form at 20k particles, checkpoints every 3 steps)."""
    from paratreet_amd import _lib
    from paratreet_amd.gravity import FORCE_TREE
    pos, vel, m, e = _plummer(20000, 36, 0.005)
    d = _driver(16)
    ref = _gradients(d, pos, vel, m, e, FORCE_TREE, every=3)
    d.load(pos, m, e, vel=vel)
    d.set_adjoint_precision(_lib.PT_FP32C)
    d.leapfrog_forward(STEPS, DT, theta=0.7, precision=_lib.PT_FP32C, mode=FORCE_TREE,
                       checkpoint_every=3)
    xT = d.state()[0]
    lam = 2.0 * m[:, None] * (xT - pos)
    gx, gv, gm = d.leapfrog_adjoint(lam_x=lam)
    d.set_adjoint_precision(_lib.PT_FP64)
    got = (gx - lam, gv, gm + np.sum((xT - pos) ** 2, axis=1))
    errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(got, ref)]
    print("fp32c vs fp64 trajectory gradient: x %.2e v %.2e m %.2e" % tuple(errs))
    assert max(errs) < 1e-5


def test_autograd_function():
    import torch
    from paratreet_amd.autograd import trajectory
    pos, vel, m, e = _plummer(4096, 35, 0.01)
    dev = torch.device("cuda:0")
    x0 = torch.tensor(pos, device=dev, requires_grad=True)
    v0 = torch.tensor(vel, device=dev, requires_grad=True)
    mm = torch.tensor(m, device=dev, requires_grad=True)
    ee = torch.tensor(e, device=dev)
    d = _driver(12)
    xT, vT = trajectory(x0, v0, mm, ee, steps=STEPS, dt=DT, driver=d)
    loss = (mm[:, None] * (xT - x0) ** 2).sum()
    loss.backward()
    gx, gv, gm = _gradients(d, pos, vel, m, e, 0)
    assert np.abs(x0.grad.cpu().numpy() - gx).max() <= 1e-12 * np.abs(gx).max()
    assert np.abs(v0.grad.cpu().numpy() - gv).max() <= 1e-12 * np.abs(gv).max()
    assert np.abs(mm.grad.cpu().numpy() - gm).max() <= 1e-12 * np.abs(gm).max()
    # forward values are the library's trajectory
    assert np.array_equal(xT.detach().cpu().numpy(), d.state()[0])
    # a second trajectory on the same Driver invalidates the first's backward
    xa, _ = trajectory(x0, v0, mm, ee, steps=2, dt=DT, driver=d)
    xb, _ = trajectory(x0, v0, mm, ee, steps=2, dt=DT, driver=d)
    with pytest.raises(RuntimeError):
        xa.sum().backward()
    xb.sum().backward()

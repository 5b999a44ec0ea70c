"""The five benchmark workloads at their stated sizes on the GPU.

C1  100k Plummer (seed 12345), bucket 12, theta 0.7, eps 0.01: universe, keys,
    sort, tree and CentroidData bit-exact, lists and counters exact, FP64
    forces within 1e-12 of the oracle (list and fused paths), one kick-drift
    at dt = 1e-3 bit-exact.
C2  1M Plummer: tests/test_gpu_fullsize.py.
C3  16M, 64 Plummer halos + 20% background, bucket 32, theta 0.5 / 0.7 / 0.9,
    eps 0.01: keys, sort and the node table bit-exact against the oracle;
    sorted keys ascending, the permutation a permutation, the buckets a
    partition; per theta the counters of the fused walk equal the oracle's
    (walk-only restatement), FP64 forces within 1e-12 of the oracle on 2048
    seeded buckets, the compensated-FP32 form with the same counters and
    within its RMS / maximum bounds of FP64 on all 16M particles.
C4  256k Plummer, 10 leapfrog steps dt = 1e-3, FP64, L = sum m |x_T - x_0|^2:
    the adjoint's gradients against central finite differences (h = 1e-6)
    on 64 seeded coordinates of x0 / v0 and 8 masses, and against the FP64
    direct-sum trajectory's gradient.
C5  4M halo mixture (16 halos), bucket 32, 50 steps dt = 5e-4, checkpoints
    every 10: compensated-FP32 forward and adjoint against the FP64 forward
    and adjoint, gradient errors and the device memory the context holds.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import bits, check_fp32c, oracle_node_data, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ C1

def test_c1_plummer100k_full_parity_and_step():
    from paratreet_amd.gravity import Configuration, Driver, GravityVisitor
    n, bucket, theta, eps, dt = 100_000, 12, 0.7, 0.01, 1e-3
    pos, vel, m = ics.plummer(n, seed=12345)
    e = np.full(n, eps)
    _say("C1: oracle pipeline")
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=theta)
    d = Driver(Configuration(max_particles_per_leaf=bucket))
    d.load(pos, m, e, vel)
    box = d.remake_universe()
    d.assign_keys()
    d.build_tree()
    assert np.array_equal(bits(box), bits(P.box))
    assert np.array_equal(d.keys(), P.keys)
    sk, perm = d.sorted()
    assert np.array_equal(sk, P.sorted_keys)
    assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64))
    g = d.nodes()
    gi, oi = np.argsort(g["key"]), np.argsort(P.nodes["key"])
    assert np.array_equal(g["key"][gi], P.nodes["key"][oi])
    for f in ("type", "depth", "begin", "count"):
        assert np.array_equal(np.asarray(g[f])[gi], np.asarray(P.nodes[f])[oi]), f
    assert np.count_nonzero(bits(d.node_data()[gi]) != bits(oracle_node_data(P)[oi])) == 0
    _say("C1: lists")
    cnt = d.traverse(GravityVisitor((0.0, 0.0, 0.0), theta))
    gpp_off, gpp, gpc_off, gpc = d.lists()
    opp_off, opp, opc_off, opc = P.lists()
    assert np.array_equal(gpp_off, opp_off) and np.array_equal(gpc_off, opc_off)
    assert np.array_equal(g["key"][gpp], P.nodes["key"][opp])
    assert np.array_equal(g["key"][gpc], P.nodes["key"][opc])
    oacc, opot, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    d.interact(0)
    acc, pot = d.accelerations()
    assert rel_err(acc, oacc).max() <= FP64_TOL
    assert (np.abs(pot - opot) / np.abs(opot)).max() <= FP64_TOL
    cntf = d.compute_forces(theta=theta, precision=0)
    accf, _ = d.accelerations()
    assert tuple(cntf) == tuple(cnt)
    e64 = rel_err(accf, oacc)
    _say("C1: fused fp64 vs oracle max rel %.3g" % e64.max())
    assert e64.max() <= FP64_TOL
    d.kick_drift(dt)
    gpos, gvel = d.state()
    opos, ovel = O.kick_drift(pos, vel, accf, dt)
    assert np.array_equal(bits(gpos), bits(opos)) and np.array_equal(bits(gvel), bits(ovel))
    # the same step from the oracle's own accelerations differs by the 1e-12
    # force agreement times dt only
    opos2, ovel2 = O.kick_drift(pos, vel, oacc, dt)
    assert np.abs(gvel - ovel2).max() <= 1e-12 * dt * np.abs(oacc).max() * 2
    d.compute_forces(theta=theta, precision=1)
    d.close()


# ------------------------------------------------------------------ C3

C3_N, C3_BUCKET, C3_EPS = 16_000_000, 32, 0.01


@pytest.fixture(scope="module")
def c3():
    from paratreet_amd.gravity import Configuration, Driver
    _say("C3: generating 16M halo mixture")
    pos, _, m = ics.halo_mixture(C3_N, n_halos=64, seed=4242)
    e = np.full(C3_N, C3_EPS)
    _say("C3: oracle keys / sort / tree / moments")
    P = O.Pipeline(pos, m, e, bucket=C3_BUCKET, theta=0.7)
    _say("C3: GPU build")
    d = Driver(Configuration(max_particles_per_leaf=C3_BUCKET))
    d.load(pos, m, e)
    d.remake_universe()
    d.assign_keys()
    d.build_tree()
    yield P, d
    d.close()


def test_c3_keys_sort_tree(c3):
    P, d = c3
    assert np.array_equal(d.keys(), P.keys)
    sk, perm = d.sorted()
    assert np.all(sk[1:] >= sk[:-1])
    assert np.array_equal(np.sort(perm.astype(np.int64)), np.arange(C3_N))
    assert np.array_equal(sk, P.sorted_keys)
    assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64))
    g = d.nodes()
    gi, oi = np.argsort(g["key"]), np.argsort(P.nodes["key"])
    assert np.array_equal(g["key"][gi], P.nodes["key"][oi])
    for f in ("type", "depth", "begin", "count"):
        assert np.array_equal(np.asarray(g[f])[gi], np.asarray(P.nodes[f])[oi]), f
    gb = d.buckets()
    counts = np.asarray(g["count"])[gb]
    begins = np.asarray(g["begin"])[gb]
    assert counts.sum() == C3_N and counts.max() <= C3_BUCKET
    assert np.array_equal(begins, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    _say("C3: %d nodes, %d buckets, depth %d" % (len(g["key"]), len(gb), d.tree_size()[2]))


@pytest.mark.parametrize("theta", [0.5, 0.7, 0.9])
def test_c3_theta(c3, theta):
    P, d = c3
    _say("C3 theta %.1f: oracle counters" % theta)
    ocnt = O.traversal_counters(P.nodes, P.data, theta)
    _say("C3 theta %.1f: GPU fp64" % theta)
    cnt = d.compute_forces(theta=theta, precision=0)
    a64, _ = d.accelerations()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    rng = np.random.default_rng(int(theta * 10))
    sel = np.sort(rng.choice(len(P.bucket_nodes), 2048, replace=False)).astype(np.int32)
    oacc, _, scnt = O.gravity(P.nodes, P.data, P.x, P.y, P.z, P.m, P.eps, theta, bucket_sel=sel)
    idx = np.concatenate([np.arange(P.nodes["begin"][b], P.nodes["begin"][b] + P.nodes["count"][b])
                          for b in P.bucket_nodes[sel]])
    e = rel_err(a64[P.perm[idx]], oacc[idx])
    _say("C3 theta %.1f: %d interactions, fp64 vs oracle on %d particles max rel %.3g" % (
        theta, int(cnt[0]) + int(cnt[1]), len(idx), e.max()))
    assert e.max() <= FP64_TOL
    cnt32 = d.compute_forces(theta=theta, precision=1)
    a32, _ = d.accelerations()
    assert tuple(cnt32) == tuple(cnt)
    check_fp32c(a32, a64, "C3 theta %.1f" % theta)


# ------------------------------------------------------------------ C4 / C5

def _gradients(d, pos, vel, m, e, steps, dt, theta, precision, mode, every):
    d.load(pos, m, e, vel=vel)
    d.leapfrog_forward(steps, dt, theta=theta, precision=precision, mode=mode,
                       checkpoint_every=every)
    xT = d.state()[0]
    lam = 2.0 * m[:, None] * (xT - pos)
    gx, gv, gm = d.leapfrog_adjoint(lam_x=lam)
    return gx - lam, gv, gm + np.sum((xT - pos) ** 2, axis=1)


def test_c4_plummer256k_adjoint_fd_and_direct():
    from paratreet_amd.gravity import FORCE_DIRECT, FORCE_TREE, PT_FP64, Configuration, Driver
    n, steps, dt, theta, H = 262_144, 10, 1e-3, 0.7, 1e-6
    pos, vel, m = ics.plummer(n, seed=2025)
    e = np.full(n, 0.005)
    d = Driver(Configuration(max_particles_per_leaf=16))
    d.set_adjoint_precision(PT_FP64)
    _say("C4: tree adjoint")
    gx, gv, gm = _gradients(d, pos, vel, m, e, steps, dt, theta, PT_FP64, FORCE_TREE, 0)

    def loss(p, v, mm):
        # per-particle terms: the differences are summed, not the two losses
        # (each ~1e-5, whose difference would lose the gradient to rounding)
        d.load(p, mm, e, vel=v)
        d.leapfrog_forward(steps, dt, theta=theta, precision=PT_FP64, mode=FORCE_TREE)
        return mm * np.sum((d.state()[0] - p) ** 2, axis=1)

    rng = np.random.default_rng(7)
    errs = {"x": [], "v": [], "m": []}
    _say("C4: finite differences")
    for c in rng.choice(2 * pos.size, 64, replace=False):
        which, c = divmod(int(c), pos.size)
        i, k = divmod(c, 3)
        if which == 0:
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += H
            p2[i, k] -= H
            fd = np.sum(loss(p1, vel, m) - loss(p2, vel, m)) / (2 * H)
            errs["x"].append(abs(fd - gx[i, k]) / np.abs(gx).max())
        else:
            v1, v2 = vel.copy(), vel.copy()
            v1[i, k] += H
            v2[i, k] -= H
            fd = np.sum(loss(pos, v1, m) - loss(pos, v2, m)) / (2 * H)
            errs["v"].append(abs(fd - gv[i, k]) / np.abs(gv).max())
    for j in rng.choice(n, 8, replace=False):
        hm = 1e-6 * m[j]
        m1, m2 = m.copy(), m.copy()
        m1[j] += hm
        m2[j] -= hm
        fd = np.sum(loss(pos, vel, m1) - loss(pos, vel, m2)) / (2 * hm)
        errs["m"].append(abs(fd - gm[j]) / np.abs(gm).max())
    errs = {k: np.array(v) for k, v in errs.items()}
    for k, v in errs.items():
        _say("C4 FD %s (%d): median %.2e max %.2e of the gradient scale" % (
            k, len(v), np.median(v), v.max()))
    for k, v in errs.items():
        assert np.median(v) < 1e-5, k
        assert (v < 1e-4).mean() >= 0.9, k
    _say("C4: FP64 direct-sum trajectory gradient")
    r = _gradients(d, pos, vel, m, e, steps, dt, theta, PT_FP64, FORCE_DIRECT, 0)
    d.close()
    rel = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip((gx, gv, gm), r)]
    _say("C4 tree vs direct gradient rel L2: x0 %.2e v0 %.2e m %.2e" % tuple(rel))
    assert rel[0] < 1e-2 and rel[1] < 1e-4 and rel[2] < 1e-4


def test_c5_halo4m_fp32c_adjoint():
    import torch
    from paratreet_amd.gravity import FORCE_TREE, PT_FP32C, PT_FP64, Configuration, Driver
    n, steps, dt, theta, every = 4_194_304, 50, 5e-4, 0.7, 10
    pos, vel, m = ics.halo_mixture(n, n_halos=16, seed=4243)
    e = np.full(n, 0.01)
    free0 = torch.cuda.mem_get_info()[0]
    d = Driver(Configuration(max_particles_per_leaf=32))
    _say("C5: FP32C forward + FP32C adjoint")
    d.set_adjoint_precision(PT_FP32C)
    g32 = _gradients(d, pos, vel, m, e, steps, dt, theta, PT_FP32C, FORCE_TREE, every)
    held = free0 - torch.cuda.mem_get_info()[0]
    _say("C5: FP64 forward + FP64 adjoint")
    d.set_adjoint_precision(PT_FP64)
    g64 = _gradients(d, pos, vel, m, e, steps, dt, theta, PT_FP64, FORCE_TREE, every)
    d.close()
    rel = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(g32, g64)]
    _say("C5 fp32c vs fp64 gradient rel L2: x0 %.2e v0 %.2e m %.2e; device memory held %.2f GB" % (
        *rel, held / 1e9))
    assert max(rel) < 1e-5
    assert held < 32e9

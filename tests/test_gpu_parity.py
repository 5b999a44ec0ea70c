"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bit-exact: universe box, Morton keys, sorted order + permutation, node table
(key, type, depth, range, parent), bucket order, CentroidData of every node,
per-bucket interaction lists and the interaction counters.  FP64 accelerations
and potentials within 1e-12 relative per particle.  FP32-compensated
accelerations within 1e-5 RMS relative of the FP64 GPU path on the same lists.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import (bits, check_fp32c, left_aligned, oracle_node_data, rel_err,
                           two_cluster)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _lambs():
    d = np.load("tests/golden/lambs_30k.npz")
    return (d["pos"].astype(np.float64), d["mass"].astype(np.float64),
            d["soft"].astype(np.float64), d["direct_acc"])


def _case(name):
    if name == "plummer2k":
        pos, _, m = ics.plummer(2000, seed=11)
        return pos, m, np.full(2000, 0.01), 12, 0.7
    if name == "plummer30k":
        pos, _, m = ics.plummer(30000, seed=12)
        return pos, m, np.full(30000, 0.005), 16, 0.7
    if name == "twocluster5k":
        pos, m = two_cluster(5000)
        return pos, m, np.full(5000, 0.02), 8, 0.5
    if name == "halo20k":
        pos, _, m = ics.halo_mixture(20000, n_halos=8, seed=3)
        return pos, m, np.full(20000, 0.05), 32, 0.9
    if name == "lambs30k":
        pos, m, e, _ = _lambs()
        return pos, m, e, 12, 0.7
    raise KeyError(name)


CASES = ["plummer2k", "plummer30k", "twocluster5k", "halo20k", "lambs30k"]


def _gpu(pos, m, e, bucket, theta, precision=0):
    from paratreet_amd.gravity import Configuration, Driver
    cfg = Configuration(max_particles_per_leaf=bucket)
    d = Driver(cfg)
    d.load(pos, m, e)
    return d


def _run_stages(d, theta):
    box = d.remake_universe()
    d.assign_keys()
    d.sort()
    from paratreet_amd import _lib
    d._check(d.L.pt_build_tree(d.ctx))
    d._check(d.L.pt_upward(d.ctx))
    from paratreet_amd.gravity import GravityVisitor
    cnt = d.traverse(GravityVisitor((0.0, 0.0, 0.0), theta))
    return box, cnt


@pytest.fixture(scope="module", params=CASES)
def pair(request):
    pos, m, e, bucket, theta = _case(request.param)
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=theta)
    d = _gpu(pos, m, e, bucket, theta)
    box, cnt = _run_stages(d, theta)
    yield request.param, P, d, box, cnt
    d.close()


def test_universe_and_keys_bitexact(pair):
    name, P, d, box, _ = pair
    assert np.array_equal(bits(box), bits(P.box)), name
    assert np.array_equal(d.keys(), P.keys), name


def test_sort_bitexact(pair):
    name, P, d, _, _ = pair
    sk, perm = d.sorted()
    assert np.array_equal(sk, P.sorted_keys), name
    assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64)), name


def test_tree_bitexact(pair):
    name, P, d, _, _ = pair
    g = d.nodes()
    o = P.nodes
    assert len(g["key"]) == len(o), name
    gi = {int(k): i for i, k in enumerate(g["key"])}
    assert set(gi) == set(int(k) for k in o["key"]), name
    for j in range(len(o)):
        i = gi[int(o["key"][j])]
        assert g["type"][i] == o["type"][j]
        assert g["depth"][i] == o["depth"][j]
        assert g["begin"][i] == o["begin"][j]
        assert g["count"][i] == o["count"][j]
        if o["parent"][j] >= 0:
            assert int(g["key"][g["parent"][i]]) == int(o["key"][o["parent"][j]])
        else:
            assert g["parent"][i] == -1
    gb = d.buckets()
    assert np.array_equal(g["key"][gb], o["key"][P.bucket_nodes]), name


def test_moments_bitexact(pair):
    name, P, d, _, _ = pair
    g = d.node_data()
    keys = d.nodes()["key"]
    gi = {int(k): i for i, k in enumerate(keys)}
    order = np.array([gi[int(k)] for k in P.nodes["key"]])
    od = oracle_node_data(P)
    gd = g[order]
    mism = np.nonzero(bits(gd) != bits(od))
    assert mism[0].size == 0, "%s: %d mismatching fields, first node %d field %d: %r vs %r" % (
        name, mism[0].size, mism[0][0], mism[1][0], gd[mism[0][0], mism[1][0]],
        od[mism[0][0], mism[1][0]])


def test_lists_and_counters_bitexact(pair):
    name, P, d, _, cnt = pair
    gpp_off, gpp, gpc_off, gpc = d.lists()
    opp_off, opp, opc_off, opc = P.lists()
    gkeys = d.nodes()["key"]
    okeys = P.nodes["key"]
    assert np.array_equal(gpp_off, opp_off), name
    assert np.array_equal(gpc_off, opc_off), name
    # canonical order = DFS order = ascending left-aligned key within a bucket
    assert np.array_equal(gkeys[gpp], okeys[opp]), name
    assert np.array_equal(gkeys[gpc], okeys[opc]), name
    _, _, ocnt = P.gravity()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt), name


def test_forces_fp64_within_1e12(pair):
    name, P, d, _, _ = pair
    d.interact(0)
    acc, pot = d.accelerations()
    oacc, opot, _ = P.gravity_unsorted()
    e = rel_err(acc, oacc)
    ep = np.abs(pot - opot) / np.abs(opot)
    print("%s fp64 vs oracle: max rel acc %.3g pot %.3g" % (name, e.max(), ep.max()))
    assert e.max() <= FP64_TOL, name
    assert ep.max() <= FP64_TOL, name


def test_forces_fp32c_precision(pair):
    name, P, d, _, _ = pair
    d.interact(0)
    a64, p64 = d.accelerations()
    d.interact(1)
    a32, p32 = d.accelerations()
    check_fp32c(a32, a64, name + " (same lists)")


@pytest.mark.parametrize("grouping", [64, 24, 0], ids=["grouped64", "grouped24", "per_bucket"])
def test_fused_path_matches_oracle(pair, grouping):
    """pt_compute_forces (fused walk + interactions, no lists) against the
    oracle, with the grouped walk (bucket runs of at most 64 or 24 particles
    sharing a walk, then sub-runs of at most 16) and with one walk per
    bucket: counters exact, FP64 within 1e-12, FP32C within its bounds."""
    name, P, d, _, _ = pair
    d.set_fused_grouping(grouping)
    try:
        cnt = d.compute_forces(theta=P.theta, precision=0)
        acc, pot = d.accelerations()
        cnt32 = d.compute_forces(theta=P.theta, precision=1)
        a32, _ = d.accelerations()
    finally:
        d.set_fused_grouping(1)
    oacc, opot, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt), name
    e = rel_err(acc, oacc)
    ep = np.abs(pot - opot) / np.abs(opot)
    print("%s fused fp64 vs oracle (%s): max rel acc %.3g pot %.3g" % (
        name, "grouped" if grouping else "per bucket", e.max(), ep.max()))
    assert e.max() <= FP64_TOL and ep.max() <= FP64_TOL, name
    assert tuple(cnt32) == tuple(cnt), name
    check_fp32c(a32, acc, name + " fused")


@pytest.mark.parametrize("grouping", [64, 0], ids=["grouped64", "per_bucket"])
def test_walk_stack_overflow_rerun(grouping):
    """A deep walk (30k particles in 8 halos, bucket 4, theta 0.1).  One walk
    per bucket: a CPU replay of the walk order reaches 226 pending entries,
    which overflows the 192-entry LDS stack, so the evaluation reruns with the
    big stack.  Grouped: the same replay gives residual lists of up to 533
    nodes, so 25 groups (FP32C, 512-entry stack) and 92 (FP64, 384) leave the
    shared walk for one walk per member.  Both match the oracle exactly
    (counters) and to 1e-12 (FP64)."""
    pos, _, m = ics.halo_mixture(30000, n_halos=8, seed=3)
    e = np.full(len(pos), 0.05)
    P = O.Pipeline(pos, m, e, bucket=4, theta=0.1)
    d = _gpu(pos, m, e, 4, 0.1)
    d.set_fused_grouping(grouping)
    cnt = d.compute_forces(theta=0.1, precision=0)
    acc, pot = d.accelerations()
    oacc, opot, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    assert rel_err(acc, oacc).max() <= FP64_TOL
    cnt32 = d.compute_forces(theta=0.1, precision=1)
    a32, _ = d.accelerations()
    d.close()
    assert tuple(cnt32) == tuple(cnt)
    check_fp32c(a32, acc, "deep walk")


@pytest.mark.parametrize("grouping", [64, 24], ids=["grouped64", "grouped24"])
def test_grouped_walk_with_large_buckets(grouping):
    """Buckets of up to 100 particles (bucket size 100, a clumpy halo
    mixture): runs mix groups of several small buckets with buckets larger
    than the cap, which form groups of one, and with buckets of more than 64
    particles, which take the per-bucket walk's several target passes inside
    the grouped kernel.  Counters exact, FP64 within 1e-12, FP32C within its
    bounds."""
    pos, _, m = ics.halo_mixture(20000, n_halos=8, seed=21)
    e = np.full(len(pos), 0.02)
    P = O.Pipeline(pos, m, e, bucket=100, theta=0.7)
    counts = P.nodes["count"][P.bucket_nodes]
    assert counts.max() > 64 and (counts <= 8).any()
    d = _gpu(pos, m, e, 100, 0.7)
    d.set_fused_grouping(grouping)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    cnt32 = d.compute_forces(theta=0.7, precision=1)
    a32, _ = d.accelerations()
    d.close()
    oacc, _, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    assert rel_err(acc, oacc).max() <= FP64_TOL
    assert tuple(cnt32) == tuple(cnt)
    check_fp32c(a32, acc, "large buckets")


def test_lambs_tree_error_vs_reference_direct():
    """Tree force error on the reference's accuracy fixture against its
    direct.acc: same order as the oracle's (quadrupole, theta 0.7)."""
    pos, m, e, ref = _lambs()
    d = _gpu(pos, m, e, 12, 0.7)
    d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    d.close()
    r = rel_err(acc, ref)
    rms = float(np.sqrt(np.mean(r * r)))
    print("lambs fp64 tree vs direct.acc: rms %.3g max %.3g" % (rms, r.max()))
    assert rms < 5e-3 and r.max() < 0.1


def test_direct_sum_matches_reference_fixture():
    pos, m, e, ref = _lambs()
    d = _gpu(pos, m, e, 12, 0.7)
    t = np.arange(0, 30000, 97)
    acc = d.direct_sum(t)
    d.close()
    r = rel_err(acc, ref[t])
    print("gpu direct vs direct.acc: rms %.3g max %.3g" % (np.sqrt(np.mean(r * r)), r.max()))
    assert r.max() < 1e-6


@pytest.mark.parametrize("n,bucket", [(1, 12), (12, 12), (13, 12), (100, 1), (257, 4)])
def test_small_and_degenerate(n, bucket):
    rng = np.random.default_rng(n * 31 + bucket)
    pos = rng.standard_normal((n, 3))
    if n >= 100:
        pos[: n // 3] = pos[0]  # coincident particles: identical keys, rsq == 0 pairs
    m = np.full(n, 1.0 / n)
    e = np.full(n, 0.01)
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=0.7)
    d = _gpu(pos, m, e, bucket, 0.7)
    _run_stages(d, 0.7)
    assert np.array_equal(d.keys(), P.keys)
    sk, perm = d.sorted()
    assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64))
    nn, nb, _ = d.tree_size()
    assert nn == len(P.nodes) and nb == len(P.bucket_nodes)
    d.interact(0)
    acc, _ = d.accelerations()
    oacc, _, _ = P.gravity_unsorted()
    nz = np.linalg.norm(oacc, axis=1) > 0
    assert np.all(acc[~nz] == 0.0)
    if nz.any():
        assert rel_err(acc[nz], oacc[nz]).max() <= FP64_TOL
    d.close()


@pytest.mark.parametrize("n,bucket", [(1, 12), (12, 12), (13, 12), (100, 1), (257, 4)])
def test_small_and_degenerate_fused_grouped(n, bucket):
    """The same small and degenerate sets through pt_compute_forces with the
    grouped walk (runs of at most 64 particles): counters exact, FP64 within
    1e-12, zero force where the oracle's is zero."""
    rng = np.random.default_rng(n * 31 + bucket)
    pos = rng.standard_normal((n, 3))
    if n >= 100:
        pos[: n // 3] = pos[0]  # coincident particles: identical keys, rsq == 0 pairs
    m = np.full(n, 1.0 / n)
    e = np.full(n, 0.01)
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=0.7)
    d = _gpu(pos, m, e, bucket, 0.7)
    d.set_fused_grouping(64)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    d.close()
    oacc, _, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    nz = np.linalg.norm(oacc, axis=1) > 0
    assert np.all(acc[~nz] == 0.0)
    if nz.any():
        assert rel_err(acc[nz], oacc[nz]).max() <= FP64_TOL


def test_leapfrog_matches_oracle():
    pos, vel, m = ics.plummer(3000, seed=5)
    e = np.full(3000, 0.01)
    from paratreet_amd.gravity import Configuration, Driver
    d = Driver(Configuration(max_particles_per_leaf=12))
    d.load(pos, m, e, vel)
    d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    d.kick_drift(1e-3)
    gpos, gvel = d.state()
    d.close()
    opos, ovel = O.kick_drift(pos, vel, acc, 1e-3)
    assert np.array_equal(bits(gpos), bits(opos))
    assert np.array_equal(bits(gvel), bits(ovel))


def test_run_matches_oracle_iterations():
    """pt_run (Driver::run, src/Driver.h:183-288): every step remakes the
    universe, regenerates keys, rebuilds the tree, evaluates the forces and
    kicks/drifts; against the oracle doing the same per step (the FP64
    forces agree to ~1e-14, so the trajectories agree far below 1e-12)."""
    pos, vel, m = ics.plummer(4000, seed=8)
    e = np.full(4000, 0.01)
    steps, dt, theta = 4, 2e-3, 0.7
    from paratreet_amd.gravity import Configuration, Driver
    d = Driver(Configuration(max_particles_per_leaf=12))
    d.load(pos, m, e, vel)
    cnt = d.run(steps, dt, theta=theta, precision=0)
    gpos, gvel = d.state()
    d.close()
    p, v = pos.copy(), vel.copy()
    ocnt = np.zeros(4, dtype=np.int64)
    for _ in range(steps):
        P = O.Pipeline(p, m, e, bucket=12, theta=theta)
        acc, _, c = P.gravity_unsorted()
        ocnt += np.asarray(c, dtype=np.int64)
        p, v = O.kick_drift(p, v, acc, dt)
    dp = np.abs(gpos - p).max() / np.abs(p).max()
    dv = np.abs(gvel - v).max() / np.abs(v).max()
    print("run vs oracle after %d steps: positions %.2e velocities %.2e" % (steps, dp, dv))
    assert dp <= 1e-12 and dv <= 1e-12
    assert np.abs(np.asarray(cnt, dtype=np.int64) - ocnt).max() <= 1e-6 * ocnt.max()


def test_top_bits_sort_and_fallback():
    """pt_compute_forces sorts only the key bits the previous tree's splits
    read, then orders each leaf by the full key: the sorted keys, the
    permutation and the forces equal the oracle's on the second evaluation;
    a much deeper tree on the next input fails the check and is rebuilt
    from the full sort with the same result."""
    from paratreet_amd.gravity import Configuration, Driver
    pos, _, m = ics.plummer(20000, seed=41)
    e = np.full(20000, 0.01)
    d = Driver(Configuration(max_particles_per_leaf=12))

    def check(pos, m, e):
        P = O.Pipeline(pos, m, e, bucket=12, theta=0.7)
        sk, perm = d.sorted()
        assert np.array_equal(sk, P.sorted_keys)
        assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64))
        acc, _ = d.accelerations()
        oacc, _, ocnt = P.gravity_unsorted()
        assert rel_err(acc, oacc).max() <= FP64_TOL
        return ocnt

    d.load(pos, m, e)
    d.compute_forces(theta=0.7, precision=0)   # full sort
    cnt = d.compute_forces(theta=0.7, precision=0)  # top bits + leaf order
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in check(pos, m, e))
    assert d.tree_size()[2] <= 10
    # a tight clump with distant outliers: a far deeper tree
    rng = np.random.default_rng(3)
    deep = rng.standard_normal((20000, 3)) * 1e-3
    deep[:8] = rng.uniform(-10.0, 10.0, (8, 3))
    d.load(deep, m, e)
    cnt = d.compute_forces(theta=0.7, precision=0)
    assert d.tree_size()[2] >= 12
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in check(deep, m, e))
    d.close()


def test_deeper_tree_in_same_context():
    """The build launches the previous tree's depth + 1 levels of splits; a
    particle set whose tree is deeper is detected and rebuilt with every
    level (counters and FP64 forces still match the oracle)."""
    rng = np.random.default_rng(8)
    pos0 = rng.uniform(-1, 1, (2000, 3))
    d = _gpu(pos0, np.full(2000, 1.0 / 2000), np.full(2000, 0.01), 12, 0.7)
    d.compute_forces(theta=0.7, precision=0)
    pos, _, m = ics.halo_mixture(30000, n_halos=8, seed=3)
    e = np.full(len(pos), 0.01)
    d.load(pos, m, e)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    P = O.Pipeline(pos, m, e, bucket=12, theta=0.7)
    oacc, _, ocnt = P.gravity_unsorted()
    nn, nb, depth = d.tree_size()
    d.close()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    assert nn == len(P.nodes) and nb == len(P.bucket_nodes)
    assert rel_err(acc, oacc).max() <= FP64_TOL

"""Unequal softenings: a gas + dark mixture (softenings 0.5 and 0.002, ratio
250) in which P-C pairs reach SPLINEQ's softened branch
(examples/GravityVisitor.h:40-73), both its u < 1 and u >= 1 parts, after
openSoftening (:122-128) was false.  With one softening for every particle
openSoftening sends every pair with r^2 < twoh^2 to addGravity, so the
other parity cases never evaluate that branch.

GPU against the oracle: lists and counters exact, FP64 accelerations and
potentials within 1e-12 relative, the compensated-FP32 form within the RMS
and maximum bounds of the FP64 path on the same interactions, for the list
path and the fused walk.  A second case puts every neighbour pair of a
lattice just inside the softening reach, where the FP32 kernels' near/far
classification has the least room.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import check_fp32c, gas_dark_mixture, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _lattice_at_reach(k=20, seed=3):
    """k^3 points on a unit lattice, jittered by 1e-7, softening 0.5 (1 + 1e-6):
    every nearest-neighbour pair sits just inside twoh = 1 + 2e-6."""
    rng = np.random.default_rng(seed)
    g = np.arange(k, dtype=np.float64)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = pos + rng.uniform(-1e-7, 1e-7, pos.shape)
    n = len(pos)
    return pos, np.full(n, 1.0 / n), np.full(n, 0.5 * (1.0 + 1e-6))


def _gasdark():
    pos, _, m, e = gas_dark_mixture(30000)
    return pos, m, e


CASES = {
    "gasdark30k": lambda: _gasdark() + (16, 0.7),
    "lattice8k_at_reach": lambda: _lattice_at_reach() + (12, 0.7),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    from paratreet_amd.gravity import Configuration, Driver
    pos, m, e, bucket, theta = CASES[request.param]()
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=theta)
    d = Driver(Configuration(max_particles_per_leaf=bucket))
    d.load(pos, m, e)
    yield request.param, P, d, theta
    d.close()


def test_softened_branch_is_reached():
    pos, _, m, e = gas_dark_mixture(30000)
    P = O.Pipeline(pos, m, e, bucket=16, theta=0.7)
    _, _, _, sp = O.gravity_soft_pairs(P.nodes, P.data, P.x, P.y, P.z, P.m, P.eps, 0.7)
    print("gasdark30k: P-C pairs in SPLINEQ's softened branch: u<1 %d, u>=1 %d" % sp)
    assert sp[0] > 0 and sp[1] > 0


def test_lists_counters_and_fp64(case):
    from paratreet_amd.gravity import GravityVisitor
    name, P, d, theta = case
    d.remake_universe()
    d.assign_keys()
    d.build_tree()
    cnt = d.traverse(GravityVisitor((0.0, 0.0, 0.0), theta))
    gpp_off, gpp, gpc_off, gpc = d.lists()
    opp_off, opp, opc_off, opc = P.lists()
    gkeys, okeys = d.nodes()["key"], P.nodes["key"]
    assert np.array_equal(gpp_off, opp_off) and np.array_equal(gpc_off, opc_off), name
    assert np.array_equal(gkeys[gpp], okeys[opp]) and np.array_equal(gkeys[gpc], okeys[opc])
    oacc, opot, ocnt = P.gravity_unsorted()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt), name
    d.interact(0)
    acc, pot = d.accelerations()
    e = rel_err(acc, oacc)
    ep = np.abs(pot - opot) / np.abs(opot)
    print("%s list fp64 vs oracle: max rel acc %.3g pot %.3g" % (name, e.max(), ep.max()))
    assert e.max() <= FP64_TOL and ep.max() <= FP64_TOL, name
    d.interact(1)
    a32, _ = d.accelerations()
    check_fp32c(a32, acc, name + " list")


@pytest.mark.parametrize("grouping", [64, 0], ids=["grouped64", "per_bucket"])
def test_fused_fp64_and_fp32c(case, grouping):
    name, P, d, theta = case
    oacc, opot, ocnt = P.gravity_unsorted()
    d.set_fused_grouping(grouping)
    try:
        cnt = d.compute_forces(theta=theta, precision=0)
        acc, pot = d.accelerations()
        cnt32 = d.compute_forces(theta=theta, precision=1)
        a32, _ = d.accelerations()
    finally:
        d.set_fused_grouping(1)
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt), name
    e = rel_err(acc, oacc)
    ep = np.abs(pot - opot) / np.abs(opot)
    print("%s fused fp64 vs oracle (grouping %d): max rel acc %.3g pot %.3g" % (
        name, grouping, e.max(), ep.max()))
    assert e.max() <= FP64_TOL and ep.max() <= FP64_TOL, name
    assert tuple(int(v) for v in cnt32) == tuple(int(v) for v in cnt), name
    check_fp32c(a32, acc, name + " fused")


def test_fp32c_refuses_softenings_below_its_range():
    """The compensated-FP32 pair terms grow as powers of 1/softening (to the
    eighth in the adjoint) and leave the FP32 range below ~2e-5 length units:
    the FP32C forms return PT_ERR_UNSUPPORTED there (forward and adjoint)
    while FP64 evaluates the same input; at the threshold FP32C runs."""
    from paratreet_amd import _lib
    from paratreet_amd.gravity import Configuration, Driver
    rng = np.random.default_rng(11)
    n = 3000
    pos = rng.normal(size=(n, 3)) * 1e-3
    m = np.full(n, 1.0 / n)
    e = np.full(n, 2.5e-5)
    e[17] = 1e-6
    d = Driver(Configuration(max_particles_per_leaf=12))
    d.load(pos, m, e)
    with pytest.raises(_lib.PtError) as err:
        d.compute_forces(theta=0.7, precision=1)
    assert err.value.status == _lib.PT_ERR_UNSUPPORTED and "softening" in str(err.value)
    d.compute_forces(theta=0.7, precision=0)
    a64, _ = d.accelerations()
    assert np.all(np.isfinite(a64))
    d.set_adjoint_precision(1)
    with pytest.raises(_lib.PtError) as err:
        d.force_vjp(np.ones((n, 3)), theta=0.7)
    assert err.value.status == _lib.PT_ERR_UNSUPPORTED
    d.set_adjoint_precision(0)
    d.force_vjp(np.ones((n, 3)), theta=0.7)
    e[17] = 2.5e-5
    d.load(pos, m, e)
    d.compute_forces(theta=0.7, precision=1)
    a32, _ = d.accelerations()
    d.compute_forces(theta=0.7, precision=0)
    a64, _ = d.accelerations()
    assert np.all(np.isfinite(a32))
    check_fp32c(a32, a64, "softening 2.5e-5")
    d.close()

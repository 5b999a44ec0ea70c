"""GPU parity of the periodic build (ExMain::traversalFn with bPeriodic,
examples/Gravity.C:35-50): pt_compute_forces sums the (2N+1)^3 replica
traversals and adds LeafEwaldFn's lattice correction (the quadrupole form,
and in the hexadecapole build the form with the root's moments to fourth
order); counters exact and FP64 forces within 1e-12 of the oracle, FP32C
within the RMS and maximum bounds of FP64, for both builds."""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _driver(bucket, period, n_rep, hexa=False):
    from paratreet_amd.gravity import Configuration, Driver
    return Driver(Configuration(max_particles_per_leaf=bucket, periodic=1, fPeriod=period,
                                nReplicas=n_rep, hexadecapole=hexa))


@pytest.mark.parametrize("hexa,grouping", [(False, 64), (False, 0), (True, 64), (True, 0)],
                         ids=["quad-grouped64", "quad-per_bucket", "hexa-grouped64",
                              "hexa-per_bucket"])
def test_periodic_replicas_match_oracle(hexa, grouping):
    """The replica sum (each replica's walk adding to the last) and the Ewald
    correction against the oracle, in both builds, with the grouped walk and
    with one walk per bucket."""
    rng = np.random.default_rng(5)
    n = 4000
    pos = rng.uniform(0.0, 2.0, (n, 3))
    pos[: n // 2] = 1.0 + 0.15 * rng.standard_normal((n // 2, 3))  # a clump
    m = np.full(n, 1.0 / n)
    e = np.full(n, 0.02)
    period = (2.0, 2.0, 2.0)
    P = O.Pipeline(pos, m, e, bucket=8, theta=0.7, hexadecapole=hexa)
    oacc, opot, ocnt = P.gravity_replicas(1, period, ewald=True)
    d = _driver(8, period, 1, hexa)
    d.load(pos, m, e)
    d.set_fused_grouping(grouping)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, pot = d.accelerations()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    # relative to the particle's force, or to the median force where the
    # replicas nearly cancel
    on = np.linalg.norm(oacc, axis=1)
    scale = np.maximum(on, np.median(on))
    err = np.linalg.norm(acc - oacc, axis=1) / scale
    # The real-space hole terms g1..g5 of the Ewald recursion cancel
    # catastrophically just outside the series radius (r^2 = 1.2e-3 L^2), so
    # the erf / exp values of two correct libms (the device's and the host's,
    # each within a few ulps) move the result by far more than their ulps:
    # the oracle measures how far one ulp of each moves it, per particle, and
    # the bound allows two.  In the quadrupole form (g up to g3) that reach
    # stays below 1e-13; in the fourth-order form it reaches 6e-11.
    sens_a, sens_p = np.zeros(n), np.zeros(n)
    for eu, xu in ((1, 1), (-1, -1), (1, -1), (-1, 1)):
        O.set_ewald_libm_perturbation(eu, xu)
        try:
            pa, pp, _ = P.gravity_replicas(1, period, ewald=True)
        finally:
            O.set_ewald_libm_perturbation(0, 0)
        sens_a = np.maximum(sens_a, np.linalg.norm(pa - oacc, axis=1))
        sens_p = np.maximum(sens_p, np.abs(pp - opot))
    tol = FP64_TOL + 2.0 * sens_a / scale
    print("periodic %s fp64 vs oracle: max rel %.3g (largest libm reach %.3g)" % (
        "hexa" if hexa else "quad", err.max(), (sens_a / scale).max()))
    worst = np.argsort(err / tol)[-5:]
    xs = np.stack([P.x, P.y, P.z], axis=1)[np.argsort(P.perm)]
    cm = P.data[0]["cm"]
    for i in worst:
        r2 = float(np.sum((xs[i] - cm) ** 2))
        print("  particle %d: err %.3g tol %.3g libm reach %.3g |a| %.3g r^2/L^2 %.4g" % (
            i, err[i], tol[i], sens_a[i] / scale[i], on[i], r2 / period[0] ** 2))
    if not hexa:
        assert (sens_a / scale).max() <= 1e-13
    assert np.all(err <= tol)
    assert np.all(np.abs(pot - opot) <= FP64_TOL * np.abs(opot).max() + 2.0 * sens_p)
    cnt32 = d.compute_forces(theta=0.7, precision=1)
    a32, _ = d.accelerations()
    assert tuple(cnt32) == tuple(cnt)
    e32 = np.linalg.norm(a32 - acc, axis=1) / np.maximum(on, np.median(on))
    print("periodic %s fp32c vs fp64: rms %.3g max %.3g" % (
        "hexa" if hexa else "quad", float(np.sqrt(np.mean(e32 * e32))), e32.max()))
    assert float(np.sqrt(np.mean(e32 * e32))) <= 1e-5
    assert e32.max() <= 1e-4
    d.close()


def test_periodic_zero_replicas_is_the_single_box_plus_ewald():
    pos, _, m = ics.plummer(3000, seed=9)
    e = np.full(3000, 0.01)
    d = _driver(12, (25.0, 25.0, 25.0), 0)
    d.load(pos, m, e)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, _ = d.accelerations()
    d.close()
    P = O.Pipeline(pos, m, e, bucket=12, theta=0.7)
    oacc, _, ocnt = P.gravity_replicas(0, (25.0, 25.0, 25.0), ewald=True)
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    assert rel_err(acc, oacc).max() <= FP64_TOL

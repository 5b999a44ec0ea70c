"""Shared helpers for the parity tests (oracle = checker, GPU = product)."""
import numpy as np

from oracle import oracle as O


def left_aligned(keys):
    keys = np.asarray(keys, dtype=np.uint64)
    depth = np.array([(int(k).bit_length() - 1) // 3 for k in keys], dtype=np.int64)
    shift = (63 - 3 * depth).astype(np.uint64)
    return keys << shift


def node_table_by_key(keys, types, depths, begins, counts):
    return {int(k): (int(t), int(d), int(b), int(c))
            for k, t, d, b, c in zip(keys, types, depths, begins, counts)}


def bits(a):
    """float64 array -> uint64 bit patterns (for bit-exact comparisons)."""
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def oracle_node_data(P):
    """Oracle CentroidData flattened like pt_get_node_data (18 per node)."""
    d = P.data
    return np.concatenate([d["radius"][:, None], d["soft"][:, None], d["mass"][:, None],
                           d["cm"], d["q"], d["lo"], d["hi"]], axis=1)


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def two_cluster(n, seed=7):
    """Two offset Plummer clusters (the shape of the reference's inputgen
    two-cluster model, built from our own generator)."""
    from paratreet_amd import ics
    h = n // 2
    p1, v1, _ = ics.plummer(h, seed=seed)
    p2, v2, _ = ics.plummer(n - h, seed=seed + 1)
    p2 = p2 + 4.0
    pos = np.concatenate([p1, p2])
    return pos, np.full(n, 1.0 / n)


def oracle_pipeline(pos, mass, eps, bucket, theta):
    return O.Pipeline(pos, mass, eps, bucket=bucket, theta=theta)


# FP32-compensated forces against the FP64 path on the same interactions:
# RMS and per-particle maximum of the relative error.  The maximum catches a
# near/far misclassification on a few targets, which the RMS alone would not.
FP32C_RMS_TOL = 1e-5
FP32C_MAX_TOL = 1e-4


def check_fp32c(a32, a64, name=""):
    e = rel_err(a32, a64)
    rms = float(np.sqrt(np.mean(e * e)))
    mx = float(e.max())
    print("%s fp32c vs fp64: rms %.3g p99 %.3g max %.3g" % (
        name, rms, float(np.percentile(e, 99)), mx))
    assert rms <= FP32C_RMS_TOL, "%s: fp32c rms %.3g" % (name, rms)
    assert mx <= FP32C_MAX_TOL, "%s: fp32c max %.3g at particle %d" % (name, mx, int(e.argmax()))
    return rms, mx


def gas_dark_mixture(n, seed=12, gas_fraction=0.2, eps_gas=0.5, eps_dark=0.002,
                     gas_mass_factor=0.05):
    """A Plummer sphere whose particles are two species, as a Tipsy file of
    gas (large hsmooth softening, small mass) and dark matter (small eps):
    unequal softenings let a P-C pair reach SPLINEQ's softened branch, which
    one softening for all particles never does."""
    from paratreet_amd import ics
    pos, vel, m0 = ics.plummer(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    gas = rng.random(n) < gas_fraction
    eps = np.where(gas, eps_gas, eps_dark)
    m = np.where(gas, gas_mass_factor, 1.0) * m0
    return pos, vel, m, eps

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

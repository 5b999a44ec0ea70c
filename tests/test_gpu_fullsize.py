"""GPU parity at the benchmark's full size (1M Plummer, bucket 16, theta 0.7,
eps 0.005), against the oracle's C restatement on the same input:

* universe box, Morton keys, sorted keys and permutation bit-exact;
* node table (keys, types, depths, ranges) and bucket order bit-exact;
* CentroidData of every node bit-exact;
* interaction counters of the fused walk exact;
* FP64 accelerations and potentials within 1e-12 relative per particle, and
  the FP32-compensated form within 1e-5 RMS and 1e-4 maximum of FP64;
* size-independent invariants: sorted keys ascending, the permutation a
  permutation, buckets partition the particles, the root holds the total mass.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import bits, check_fp32c, oracle_node_data, rel_err

pytestmark = pytest.mark.gpu

N, BUCKET, THETA, EPS = 1_000_000, 16, 0.7, 0.005


@pytest.fixture(scope="module")
def full():
    from paratreet_amd.gravity import Configuration, Driver, GravityVisitor
    pos, _, m = ics.plummer(N, seed=2024)
    e = np.full(N, EPS)
    print("fullsize: oracle pipeline, n=%d" % N, flush=True)
    P = O.Pipeline(pos, m, e, bucket=BUCKET, theta=THETA)
    print("fullsize: GPU stages", flush=True)
    d = Driver(Configuration(max_particles_per_leaf=BUCKET))
    d.load(pos, m, e)
    box = d.remake_universe()
    d.assign_keys()
    d.sort()
    d._check(d.L.pt_build_tree(d.ctx))
    d._check(d.L.pt_upward(d.ctx))
    yield P, d, box, pos, m
    d.close()


def test_keys_sort_tree_bitexact(full):
    P, d, box, _, _ = full
    assert np.array_equal(bits(box), bits(P.box))
    assert np.array_equal(d.keys(), P.keys)
    sk, perm = d.sorted()
    assert np.array_equal(sk, P.sorted_keys)
    assert np.array_equal(perm.astype(np.int64), P.perm.astype(np.int64))
    assert np.all(sk[1:] >= sk[:-1])
    assert np.array_equal(np.sort(perm.astype(np.int64)), np.arange(N))
    g = d.nodes()
    o = P.nodes
    assert len(g["key"]) == len(o)
    gi = np.argsort(g["key"])
    oi = np.argsort(o["key"])
    assert np.array_equal(g["key"][gi], o["key"][oi])
    for f in ("type", "depth", "begin", "count"):
        assert np.array_equal(np.asarray(g[f])[gi], np.asarray(o[f])[oi]), f
    gb = d.buckets()
    assert np.array_equal(g["key"][gb], o["key"][P.bucket_nodes])
    # buckets partition the sorted particles in order
    counts = np.asarray(g["count"])[gb]
    begins = np.asarray(g["begin"])[gb]
    assert counts.sum() == N and counts.max() <= BUCKET
    assert np.array_equal(begins, np.concatenate([[0], np.cumsum(counts)[:-1]]))


def test_moments_bitexact(full):
    P, d, _, _, m = full
    g = d.node_data()
    keys = d.nodes()["key"]
    gi = np.argsort(keys)
    oi = np.argsort(P.nodes["key"])
    od = oracle_node_data(P)
    mism = np.count_nonzero(bits(g[gi]) != bits(od[oi]))
    assert mism == 0, "%d mismatching CentroidData fields" % mism
    root = int(np.nonzero(np.asarray(d.nodes()["depth"]) == 0)[0][0])
    assert abs(g[root, 2] - m.sum()) <= 1e-12 * m.sum()


def test_forces_and_counters(full):
    P, d, _, _, _ = full
    print("fullsize: oracle forces", flush=True)
    oacc, opot, ocnt = P.gravity_unsorted()
    print("fullsize: GPU forces", flush=True)
    cnt = d.compute_forces(theta=THETA, precision=0)
    acc, pot = d.accelerations()
    assert tuple(int(v) for v in cnt) == tuple(int(v) for v in ocnt)
    e = rel_err(acc, oacc)
    ep = np.abs(pot - opot) / np.abs(opot)
    print("fullsize fp64 vs oracle: max rel acc %.3g pot %.3g" % (e.max(), ep.max()))
    assert e.max() <= 1e-12 and ep.max() <= 1e-12
    cnt32 = d.compute_forces(theta=THETA, precision=1)
    a32, _ = d.accelerations()
    assert tuple(int(v) for v in cnt32) == tuple(int(v) for v in ocnt)
    check_fp32c(a32, acc, "fullsize")

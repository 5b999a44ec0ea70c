"""Pins of the oracle's dual-tree traversal (oracle/pt_oracle.c or_dual,
restating DualTraverser::doTrav / runInvertedTraversal, src/Traverser.h:
485-610, with GravityVisitor::open/node/leaf/cell).

What the definitions fix, checked independently of the restatement:

* startDual's pair (root, root): the root's box encloses itself, so cell()
  is false and runInvertedTraversal(root, root) runs with an internal
  source, whose n_particles is -1: no particle moves, every internal node of
  the tree is opened (the root's opening sphere holds its whole box), every
  leaf adds -n to n_part_ints.  Counters = (0, 2^64 - N, #internal, 1).
* A pair of disjoint nodes of one depth: cell() is true for every pair of
  internal nodes below it (disjoint octree cells never enclose one another),
  so the walk splits both until a leaf is reached; with buckets of at most
  six particles open() is true for every leaf source (nMinParticleNode), so
  nothing is lost: every target bucket's leaf() and node() sources partition
  the source node's particles exactly once, every node() source fails open()
  while its ancestors from the bucket's depth down pass it, the counters follow from the
  lists, and the forces approximate the direct sum of the source's particles
  on the target's at the quadrupole tree's accuracy.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics


def _pipe(n=3000, bucket=6, seed=4, eps=1e-5):
    pos, _, m = ics.plummer(n, seed=seed)
    return O.Pipeline(pos, m, np.full(n, eps), bucket=bucket, theta=0.7)


def test_root_pair_closed_form():
    P = _pipe(bucket=8)
    acc, pot, cnt = P.dual()
    n_internal = int(np.count_nonzero(P.nodes["type"] == O.OR_INTERNAL))
    assert tuple(int(v) for v in cnt) == (0, 2 ** 64 - P.n, n_internal, 1)
    assert np.all(acc == 0.0) and np.all(pot == 0.0)
    pp_off, pp, pc_off, pc = P.dual_lists()
    assert len(pp) == 0 and len(pc) == 0


def _subtree_range(P, i):
    return int(P.nodes["begin"][i]), int(P.nodes["begin"][i] + P.nodes["count"][i])


def _ancestors(P, i, stop):
    out = []
    p = int(P.nodes["parent"][i])
    while p >= 0 and p != stop:
        out.append(p)
        p = int(P.nodes["parent"][p])
    return out + [stop]


def _open(P, s, t):
    d = P.data
    if d["count"][s] <= 6:
        return True
    gf = 4.0 / 3.0 / (P.theta * P.theta)
    return O.intersect_box_point(d["lo"][t], d["hi"][t], d["cm"][s], d["radius"][s] ** 2 * gf)


@pytest.mark.parametrize("pair", [(8, 15), (9, 10), (12, 11)])
def test_disjoint_pair_partitions_sources(pair):
    P = _pipe()
    s, t = (P.node_id(k) for k in pair)
    pp_off, pp, pc_off, pc = P.dual_lists(*pair)
    s0, s1 = _subtree_range(P, s)
    t0, t1 = _subtree_range(P, t)
    bn = P.bucket_nodes
    n_node = n_part = 0
    touched = 0
    for b, node in enumerate(bn):
        b0, b1 = _subtree_range(P, node)
        inside = t0 <= b0 and b1 <= t1
        ents = list(pp[pp_off[b]:pp_off[b + 1]]) + list(pc[pc_off[b]:pc_off[b + 1]])
        if not inside:
            assert not ents
            continue
        touched += 1
        cover = np.zeros(s1 - s0, dtype=np.int64)
        for e in ents:
            e0, e1 = _subtree_range(P, e)
            assert s0 <= e0 and e1 <= s1
            cover[e0 - s0:e1 - s0] += 1
        assert np.all(cover == 1), "bucket %d: sources do not partition the source node" % b
        # the pair walk keeps source and target at one depth, so open() is
        # tested against this bucket from the bucket's depth down
        dt = int(P.nodes["depth"][node])
        for e in pc[pc_off[b]:pc_off[b + 1]]:
            assert not _open(P, e, node)
            assert all(_open(P, a, node) for a in _ancestors(P, e, s)
                       if a != e and P.nodes["depth"][a] >= dt)
        n_part += sum(int(P.nodes["count"][e]) for e in pp[pp_off[b]:pp_off[b + 1]]) * (b1 - b0)
        n_node += (pc_off[b + 1] - pc_off[b]) * (b1 - b0)
    assert touched > 0
    acc, _, cnt = P.dual(*pair)
    assert int(cnt[0]) == n_node and int(cnt[1]) == n_part
    # forces against the direct sum of the source's particles on the target's
    x = np.stack([P.x, P.y, P.z], axis=1)
    d = x[s0:s1][None, :, :] - x[t0:t1][:, None, :]
    r2 = np.sum(d * d, axis=2)
    ref = np.sum(d * (P.m[s0:s1][None, :] / (r2 * np.sqrt(r2)))[:, :, None], axis=1)
    a_sorted = acc[P.perm]
    err = np.linalg.norm(a_sorted[t0:t1] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("pair %s: %d target buckets, rms %.2g max %.2g vs direct" % (
        pair, touched, np.sqrt(np.mean(err * err)), err.max()))
    assert np.sqrt(np.mean(err * err)) < 5e-3 and err.max() < 5e-2
    # no force anywhere outside the target
    mask = np.ones(P.n, dtype=bool)
    mask[t0:t1] = False
    assert np.all(a_sorted[mask] == 0.0)

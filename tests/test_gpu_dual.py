"""GPU parity of the dual-tree traversal (pt_start_dual: Subtree::startDual ->
DualTraverser::doTrav / runInvertedTraversal, src/Traverser.h:485-610, with
GravityVisitor::cell, examples/GravityVisitor.h:226-230) against the oracle's
restatement: counters exact (modulo 2^64, as the reference's), per-bucket
lists equal as multisets, FP64 accelerations and potentials within 1e-12
relative (exactly zero where the oracle's are), the compensated-FP32 form
within its bounds of FP64.  Start pairs cover every branch of doTrav: the
root pair (cell() false at once, an inverted walk with an internal source),
sibling pairs (cell() true down to the leaves), a leaf against an internal
node (inverted walk with a leaf source), an internal node against a leaf,
a node against its own descendant, and a target deeper than its source.
pt_compute_forces with dual_tree runs the root-pair traversal before the
down traversal, both adding to the accelerations and the counters
(examples/Gravity.C:30-33)."""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics
from tests.helpers import check_fp32c, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _pick_pairs(P):
    """Key pairs for every doTrav branch on this tree."""
    nodes = P.nodes
    keys = nodes["key"]
    typ = nodes["type"]
    depth = nodes["depth"]
    ch = nodes["child"][0]
    internal_kids = [int(c) for c in ch if typ[c] == O.OR_INTERNAL]
    a, b = internal_kids[0], internal_kids[-1]
    pairs = [(1, 1), (int(keys[a]), int(keys[b])), (int(keys[b]), int(keys[a]))]
    if len(internal_kids) > 2:
        pairs.append((int(keys[internal_kids[1]]), int(keys[internal_kids[2]])))

    def first_desc(root, want):
        stack = [root]
        while stack:
            i = stack.pop()
            if typ[i] == want and i != root:
                return i
            if typ[i] == O.OR_INTERNAL:
                stack.extend(int(c) for c in nodes["child"][i][::-1])
        return None

    leaf_a = first_desc(a, O.OR_LEAF)
    deep_b = first_desc(int(nodes["child"][b][0]), O.OR_INTERNAL) or b
    pairs.append((int(keys[leaf_a]), int(keys[b])))         # leaf source, internal target
    pairs.append((int(keys[a]), int(keys[first_desc(b, O.OR_LEAF)])))  # internal, leaf
    pairs.append((int(keys[a]), int(keys[first_desc(a, O.OR_INTERNAL) or a])))  # own descendant
    if depth[deep_b] > depth[a]:
        pairs.append((int(keys[a]), int(keys[deep_b])))      # target deeper than source
    return pairs


CASES = {
    "plummer4k_b8": lambda: ics.plummer(4000, seed=31)[0::2] + (8, 0.01),
    "twohalo3k_b12": lambda: ics.halo_mixture(3000, n_halos=2, seed=6)[0::2] + (12, 0.05),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    from paratreet_amd.gravity import Configuration, Driver
    pos, m, bucket, eps = CASES[request.param]()
    e = np.full(len(pos), eps)
    P = O.Pipeline(pos, m, e, bucket=bucket, theta=0.7)
    d = Driver(Configuration(max_particles_per_leaf=bucket))
    d.load(pos, m, e)
    d.remake_universe()
    d.assign_keys()
    d.build_tree()
    yield request.param, P, d, (pos, m, e, bucket)
    d.close()


def _bucket_multisets(off, ids, keys):
    return [sorted(int(k) for k in keys[ids[off[b]:off[b + 1]]]) for b in range(len(off) - 1)]


def test_dual_pairs_match_oracle(case):
    from paratreet_amd.gravity import GravityVisitor
    name, P, d, _ = case
    gkeys = d.nodes()["key"]
    v = GravityVisitor((0.0, 0.0, 0.0), 0.7)
    for sk, tk in _pick_pairs(P):
        label = "%s pair (%d, %d)" % (name, sk, tk)
        cnt = d.start_dual(v, sk, tk, precision=0)
        acc, pot = d.accelerations()
        oacc, opot, ocnt = P.dual(sk, tk)
        assert tuple(int(c) for c in cnt) == tuple(int(c) for c in ocnt), label
        gpp_off, gpp, gpc_off, gpc = d.lists()
        opp_off, opp, opc_off, opc = P.dual_lists(sk, tk)
        okeys = P.nodes["key"]
        assert _bucket_multisets(gpp_off, gpp, gkeys) == _bucket_multisets(opp_off, opp, okeys)
        assert _bucket_multisets(gpc_off, gpc, gkeys) == _bucket_multisets(opc_off, opc, okeys)
        nz = np.linalg.norm(oacc, axis=1) > 0
        assert np.all(acc[~nz] == 0.0) and np.all(pot[np.abs(opot) == 0] == 0.0), label
        if nz.any():
            e = rel_err(acc[nz], oacc[nz])
            pz = np.abs(opot) > 0
            ep = np.abs(pot[pz] - opot[pz]) / np.abs(opot[pz]) if pz.any() else np.zeros(1)
            print("%s: %d particles moved, fp64 max rel %.3g pot %.3g" % (
                label, int(nz.sum()), e.max(), ep.max()))
            assert e.max() <= FP64_TOL and ep.max() <= FP64_TOL, label
            cnt32 = d.start_dual(v, sk, tk, precision=1)
            a32, _ = d.accelerations()
            assert tuple(cnt32) == tuple(cnt), label
            check_fp32c(a32[nz], acc[nz], label)


@pytest.mark.parametrize("grouping", [64, 0], ids=["grouped64", "per_bucket"])
def test_dual_tree_config_adds_to_down_traversal(case, grouping):
    from paratreet_amd.gravity import Configuration, Driver
    name, P, _, (pos, m, e, bucket) = case
    d = Driver(Configuration(max_particles_per_leaf=bucket, dual_tree=True))
    d.load(pos, m, e)
    d.set_fused_grouping(grouping)
    cnt = d.compute_forces(theta=0.7, precision=0)
    acc, pot = d.accelerations()
    cnt32 = d.compute_forces(theta=0.7, precision=1)
    a32, _ = d.accelerations()
    d.close()
    dacc, dpot, dcnt = P.dual()
    oacc, opot, ocnt = P.gravity_unsorted()
    want = tuple((int(a) + int(b)) % 2 ** 64 for a, b in zip(dcnt, ocnt))
    assert tuple(int(c) for c in cnt) == want, name
    assert tuple(cnt32) == tuple(cnt)
    ref = oacc + dacc
    assert rel_err(acc, ref).max() <= FP64_TOL, name
    assert (np.abs(pot - (opot + dpot)) / np.abs(opot + dpot)).max() <= FP64_TOL
    check_fp32c(a32, acc, name + " dual_tree")


def test_dual_rejections():
    from paratreet_amd import _lib
    from paratreet_amd.gravity import Configuration, Driver, GravityVisitor
    pos, _, m = ics.plummer(500, seed=2)
    e = np.full(500, 0.01)
    d = Driver(Configuration(max_particles_per_leaf=8))
    d.load(pos, m, e)
    v = GravityVisitor((0.0, 0.0, 0.0), 0.7)
    with pytest.raises(_lib.PtError) as ex:
        d.start_dual(v)
    assert ex.value.status == _lib.PT_ERR_STATE
    d.remake_universe()
    d.assign_keys()
    d.build_tree()
    with pytest.raises(_lib.PtError) as ex:
        d.start_dual(v, 1, 0x7ffffff)  # no node has this key
    assert ex.value.status == _lib.PT_ERR_ARG
    d.close()
    with pytest.raises(_lib.PtError):
        Driver(Configuration(max_particles_per_leaf=8, dual_tree=True, periodic=1,
                             fPeriod=(4.0, 4.0, 4.0), nReplicas=1))

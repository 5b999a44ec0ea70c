"""Morton key rounding: the oracle's definition against the float-mantissa one.

SFC::generateKey lives in the unvendored N-BodyShop/utility submodule, so its
exact bit derivation cannot be read here.  The oracle (and the GPU path) use
q = floor(u * 2^21) per axis with u = (x - lesser) / (greater - lesser) in
FP64.  The utility of the reference's era takes the bits from the FP32
mantissa of u + 1 (in [1, 2) the 23 mantissa bits are u * 2^23 rounded to
nearest; the top 21 of them are the coordinate).  This test measures how far
the two definitions lead apart on the 100k (bucket 12) and 1M (bucket 16)
Plummer inputs: keys, tree nodes, buckets, per-bucket interaction lists and
FP64 forces.

What the two definitions fix, and what is asserted: with v = u * 2^23,
floor(round(v) / 4) differs from floor(v / 4) exactly when frac(v / 4) >= 7/8
(the rounding carries into the 21st bit), and then by +1; so about
1 - (7/8)^3 = 33% of keys differ, every difference is a +1 step on the axes
where the carry happened, and only a carry into a bit that a split reads
moves a particle between tree nodes.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paratreet_amd import ics


def _deinterleave(keys):
    q = np.zeros((len(keys), 3), dtype=np.int64)
    k = keys.astype(np.uint64)
    for b in range(21):
        for axis in range(3):
            bit = (k >> np.uint64(3 * b + 2 - axis)) & np.uint64(1)
            q[:, axis] |= bit.astype(np.int64) << b
    return q


def _interleave(q):
    k = np.zeros(len(q), dtype=np.uint64)
    for b in range(21):
        for axis in range(3):
            bit = ((q[:, axis] >> b) & 1).astype(np.uint64)
            k |= bit << np.uint64(3 * b + 2 - axis)
    return k | (np.uint64(1) << np.uint64(63))


def float_mantissa_keys(pos, box):
    """Keys from the FP32 mantissa of u + 1, u in FP64 as the oracle has it."""
    lo, hi = np.asarray(box[:3]), np.asarray(box[3:])
    u = (pos - lo) / (hi - lo)
    f = (1.0 + u).astype(np.float32)
    mant = f.view(np.uint32).astype(np.int64) & 0x7FFFFF
    # u + 1 == 2.0f (u within half an ulp of 1) has mantissa 0: clamp as the
    # oracle clamps u * 2^21 >= 2^21 - 1
    q = np.where(f >= np.float32(2.0), 0x1FFFFF, mant >> 2)
    return _interleave(q)


def _tree(pos, m, eps, keys, bucket, theta):
    sk, perm = O.sort(keys)
    x, y, z = (pos[perm, k].copy() for k in range(3))
    nodes = O.build_tree(sk, bucket)
    data = O.moments(nodes, x, y, z, m[perm], eps[perm])
    return perm, nodes, data, (x, y, z, m[perm].copy(), eps[perm].copy())


@pytest.mark.parametrize("n,bucket,eps,lists", [(100_000, 12, 0.01, True),
                                                 (1_000_000, 16, 0.005, False)])
def test_float_mantissa_keys_against_floor(n, bucket, eps, lists):
    theta = 0.7
    pos, _, m = ics.plummer(n, seed=2024 if n == 1_000_000 else 12345)
    e = np.full(n, eps)
    box = O.universe(pos[:, 0], pos[:, 1], pos[:, 2])
    k64 = O.keys(pos[:, 0], pos[:, 1], pos[:, 2], box)
    k32 = float_mantissa_keys(pos, box)

    q64, q32 = _deinterleave(k64), _deinterleave(k32)
    dq = q32 - q64
    assert set(np.unique(dq)) <= {0, 1}, "a difference other than one rounding carry"
    v4 = (pos - box[:3]) / (box[3:] - box[:3]) * 2.0 ** 21
    carry = (v4 - np.floor(v4)) >= 0.875
    # away from the exact 7/8 boundary the carry rule decides every axis
    clear = np.abs((v4 - np.floor(v4)) - 0.875) > 1e-6
    assert np.array_equal(dq[clear] == 1, carry[clear])
    frac_keys = float(np.mean(np.any(dq != 0, axis=1)))
    assert abs(frac_keys - (1 - (7 / 8) ** 3)) < 0.01, frac_keys

    p64, n64, d64, s64 = _tree(pos, m, e, k64, bucket, theta)
    p32, n32, d32, s32 = _tree(pos, m, e, k32, bucket, theta)
    keys64, keys32 = set(n64["key"].tolist()), set(n32["key"].tolist())
    node_diff = len(keys64 ^ keys32)
    b64 = O.buckets(n64)
    b32 = O.buckets(n32)
    # a bucket is unchanged when its key and its particles are the same
    def bucket_sets(nodes, bn, perm):
        return {int(nodes["key"][b]): frozenset(perm[nodes["begin"][b]:nodes["begin"][b]
                                                     + nodes["count"][b]].tolist())
                for b in bn}
    bs64, bs32 = bucket_sets(n64, b64, p64), bucket_sets(n32, b32, p32)
    same = [k for k in bs64 if bs32.get(k) == bs64[k]]
    msg = ("n=%d: keys differing %d (%.1f%%), nodes in one tree only %d of %d, buckets "
           "changed %d of %d" % (n, int(np.count_nonzero(k64 != k32)), 100 * frac_keys,
                                 node_diff, len(n64), len(bs64) - len(same), len(bs64)))
    assert node_diff <= 0.01 * len(n64)
    if lists:
        l64 = O.lists(n64, d64, theta)
        l32 = O.lists(n32, d32, theta)
        i64 = {int(n64["key"][b]): j for j, b in enumerate(b64)}
        i32 = {int(n32["key"][b]): j for j, b in enumerate(b32)}
        diff_entries = 0
        total = 0
        for kb in same:
            a, b = i64[kb], i32[kb]
            for off, ids, nodes, off2, ids2, nodes2 in (
                    (l64[0], l64[1], n64, l32[0], l32[1], n32),
                    (l64[2], l64[3], n64, l32[2], l32[3], n32)):
                s1 = nodes["key"][ids[off[a]:off[a + 1]]].tolist()
                s2 = nodes2["key"][ids2[off2[b]:off2[b + 1]]].tolist()
                total += len(s1)
                diff_entries += len(set(s1) ^ set(s2))
        msg += ", list entries differing %d of %d in the unchanged buckets" % (diff_entries,
                                                                             total)
        a64, _, _ = O.gravity(n64, d64, *s64, theta)
        a32, _, _ = O.gravity(n32, d32, *s32, theta)
        g64 = np.empty_like(a64)
        g32 = np.empty_like(a32)
        g64[p64] = a64
        g32[p32] = a32
        r = np.linalg.norm(g32 - g64, axis=1) / np.linalg.norm(g64, axis=1)
        msg += ", FP64 forces differ by rms %.2g max %.2g" % (np.sqrt(np.mean(r * r)), r.max())
        assert diff_entries <= 0.01 * total
    print(msg)

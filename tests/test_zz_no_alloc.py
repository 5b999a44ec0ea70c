"""pt_sort_noalloc (the sort of pt_sort with its temporaries in a fixed
per-GPU device array) against pt_sort on the same keys: two calls in a row,
the second smaller, both bit-exact with pt_sort's sorted keys and
permutation; plus the size limit and the one-key case."""
import ctypes

import numpy as np
import pytest

from paratreet_amd import _lib, ics

pytestmark = pytest.mark.gpu

MAX_N = 1 << 21  # PT_SORT_NOALLOC_MAX_N


def _noalloc_fn():
    L = _lib.load()
    fn = L.pt_sort_noalloc
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p]
    return L, fn


def _context_sort(pos, n):
    """pt_sort through a context: its keys stay on the device (the device
    view), its sorted keys and permutation come back to the host."""
    from paratreet_amd.gravity import Configuration, Driver
    d = Driver(Configuration(max_particles_per_leaf=16))
    d.load(pos[:n], np.full(n, 1.0 / n), np.full(n, 0.01))
    d.remake_universe()
    d.assign_keys()
    d.sort()
    sk, perm = d.sorted()
    return d, sk, perm


def _call(fn, view, n):
    import torch
    out_k = torch.empty(n, dtype=torch.int64, device="cuda")
    out_p = torch.empty(n, dtype=torch.int32, device="cuda")
    # poison the outputs so a call that writes nothing cannot pass
    out_k.fill_(-1)
    out_p.fill_(-1)
    torch.cuda.synchronize()
    st = fn(view.keys, n, out_k.data_ptr(), out_p.data_ptr(), view.stream)
    torch.cuda.synchronize()
    return st, out_k.cpu().numpy().view(np.uint64), out_p.cpu().numpy().view(np.uint32)


def test_sort_noalloc_matches_pt_sort_twice():
    L, fn = _noalloc_fn()
    # first call: the largest size the array holds (a Plummer sphere)
    pos, _, _ = ics.plummer(MAX_N, seed=31)
    d1, sk1, p1 = _context_sort(pos, MAX_N)
    print("noalloc call 1: n=%d" % MAX_N, flush=True)
    st, k1, q1 = _call(fn, d1.device_view(), MAX_N)
    assert st == 0, L.pt_last_error(None)
    assert np.array_equal(k1, sk1)
    assert np.array_equal(q1, p1)
    d1.close()
    # second call, smaller, with many equal keys (positions on a coarse grid):
    # ties must come out in input order, as pt_sort orders them
    n2 = 100_003
    pos2, _, _ = ics.plummer(n2, seed=32)
    pos2 = np.round(pos2 * 4.0) / 4.0
    d2, sk2, p2 = _context_sort(pos2, n2)
    assert len(np.unique(sk2)) < n2 // 2
    print("noalloc call 2: n=%d" % n2, flush=True)
    st, k2, q2 = _call(fn, d2.device_view(), n2)
    assert st == 0, L.pt_last_error(None)
    assert np.array_equal(k2, sk2)
    assert np.array_equal(q2, p2)
    d2.close()


def test_sort_noalloc_limits():
    import torch
    L, fn = _noalloc_fn()
    buf = torch.zeros(4, dtype=torch.int64, device="cuda")
    p = buf.data_ptr()
    st = fn(p, MAX_N + 1, p, p, None)
    assert st == _lib.PT_ERR_ARG
    assert b"exceeds the maximum" in L.pt_last_error(None)
    assert fn(p, -1, p, p, None) == _lib.PT_ERR_ARG
    assert fn(p, 0, p, p, None) == 0
    # one key: the copy path
    key = torch.tensor([12345], dtype=torch.int64, device="cuda")
    ok = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    op = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    assert fn(key.data_ptr(), 1, ok.data_ptr(), op.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(ok.item()) == 12345 and int(op.item()) == 0

"""The allocation-free entry point bs_conn_bernoulli_count_noalloc against
bs_conn_bernoulli_count: identical row pointers, edge counts and largest
in-degree on two calls in a row (the second smaller, so stale scratch from
the first call would show), and the size limit rejected before any device
work."""
import ctypes as C

import numpy as np
import pytest

from brainpy.state_amd import _lib as L


def _count(fn, seed, n_pre, n_post, p, excl):
    import torch
    row = torch.full((n_pre + 1,), -7, dtype=torch.int64, device='cuda')
    ne, mx = C.c_int64(-1), C.c_int32(-1)
    torch.cuda.synchronize()
    L.check(fn(seed, n_pre, n_post, p, excl, C.c_void_p(row.data_ptr()), C.byref(ne), C.byref(mx)))
    torch.cuda.synchronize()
    return row.cpu().numpy(), ne.value, mx.value


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = L.lib()
    ne, mx = C.c_int64(), C.c_int32()
    rc = lib.bs_conn_bernoulli_count_noalloc(1, (1 << 20) + 1, 10, 0.1, 0, C.c_void_p(0), C.byref(ne), C.byref(mx))
    assert rc == -1 and b'at most' in lib.bs_last_error()
    rc = lib.bs_conn_bernoulli_count_noalloc(1, 10, (1 << 20) + 1, 0.1, 0, C.c_void_p(0), C.byref(ne), C.byref(mx))
    assert rc == -1


@pytest.mark.gpu
def test_noalloc_count_equals_the_allocating_entry_point():
    lib = L.lib()
    L.require_device()
    cases = [(11, 3000, 5000, 0.07, 1),      # large first: the second call finds stale scratch
             (12, 1500, 900, 0.2, 0),
             (13, 1, 1, 1.0, 0),
             (14, 2049, 0, 0.5, 0)]
    for seed, n_pre, n_post, p, excl in cases:
        a = _count(lib.bs_conn_bernoulli_count, seed, n_pre, n_post, p, excl)
        b = _count(lib.bs_conn_bernoulli_count_noalloc, seed, n_pre, n_post, p, excl)
        assert np.array_equal(a[0], b[0]), (seed, n_pre, n_post)
        assert a[1] == b[1] and a[2] == b[2], (seed, a[1:], b[1:])
        assert a[0][0] == 0 and a[0][-1] == a[1]
        if n_post and p > 0:
            assert a[1] > 0

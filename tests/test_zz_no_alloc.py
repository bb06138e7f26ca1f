"""The allocation-free exact kNN (psq_knn_exact_device_noalloc) against the
allocating entry it mirrors (psq_knn_exact_device), through the C-ABI: bit for
bit on two calls in a row, the second smaller (the scratch array keeps the
first call's contents), with duplicate points that send queries down the
fallback, and the size limit rejected before any work."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from polariseq_amd import _native as N
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def _run(entry, x, q, k):
    import torch
    n, d = x.shape
    xd = torch.from_numpy(x).cuda()
    qd = None if q is None else torch.from_numpy(q.view(np.int32)).cuda()
    nq = n if q is None else len(q)
    idx = torch.full((nq, k), -7, dtype=torch.int32, device="cuda")
    dist = torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(N.lib(), entry)(xd.data_ptr(), n, d, None if qd is None else qd.data_ptr(), nq, k,
                                 idx.data_ptr(), dist.data_ptr(), stream)
    assert rc == N.PSQ_OK, N.lib().psq_last_error().decode()
    torch.cuda.synchronize()
    return idx.cpu().numpy().view(np.uint32), dist.cpu().numpy()


def _same(x, q, k):
    print(f"noalloc exact kNN n={x.shape[0]} d={x.shape[1]} "
          f"n_q={x.shape[0] if q is None else len(q)} k={k}", flush=True)
    a = _run("psq_knn_exact_device", x, q, k)
    b = _run("psq_knn_exact_device_noalloc", x, q, k)
    assert np.array_equal(a[0], b[0]), "indices differ from the allocating entry"
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), "distances differ"


def test_two_calls_in_a_row_second_smaller():
    x = ds.mixture(60000, 50, 20, seed=1)
    q = np.arange(0, 60000, 3, dtype=np.uint32)
    _same(x, q, 15)
    y = ds.mixture(5000, 50, 8, seed=2)
    _same(y, None, 10)


def test_duplicates_take_the_fallback():
    rng = np.random.default_rng(5)
    base = ds.mixture(100, 30, 4, seed=5)
    x = np.ascontiguousarray(base[rng.integers(0, 100, 20000)])  # ~200 copies of each point
    _same(x, None, 32)
    _same(x[:7000], np.arange(0, 7000, 2, dtype=np.uint32), 64)


def test_generic_width_and_segments_shrunk_to_fit():
    x = ds.mixture(200000, 20, 10, seed=3)  # the standard lists would not fit: fewer segments
    _same(x, np.arange(0, 200000, 5, dtype=np.uint32), 16)


def test_limits_rejected_before_any_work():
    import torch
    x = torch.zeros((300000, 50), dtype=torch.float32, device="cuda")
    out_i = torch.zeros((300000, 15), dtype=torch.int32, device="cuda")
    out_d = torch.zeros((300000, 15), dtype=torch.float32, device="cuda")
    L = N.lib()
    rc = L.psq_knn_exact_device_noalloc(x.data_ptr(), 300000, 50, None, 300000, 15,
                                        out_i.data_ptr(), out_d.data_ptr(), None)
    assert rc == N.PSQ_E_INVALID and "scratch" in L.psq_last_error().decode()
    rc = L.psq_knn_exact_device_noalloc(x.data_ptr(), 1000, 50, None, 1000, 65,
                                        out_i.data_ptr(), out_d.data_ptr(), None)
    assert rc == N.PSQ_E_UNSUPPORTED
    torch.cuda.synchronize()
    assert int(out_i.abs().sum()) == 0 and float(out_d.abs().sum()) == 0.0

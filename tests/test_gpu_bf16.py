"""GPU checks of the precision form (BF16 MFMA join tiles + FP32 re-rank).

The candidate ranking uses BF16 MFMA dot products whose internal accumulation
order is the hardware's, so this form is not compared bit for bit with a CPU
model. What is checked: every output distance is the reference's exact f32
sq_dist (simd.rs:45-62) square-rooted, bit for bit; the graph invariants of
nndescent.rs:1081-1107; determinism for a seed; recall close to the plain form's.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def _valid(g, n, k):
    ip, ix, dx = g["indptr"], g["indices"], g["data"]
    assert (np.diff(ip) == k).all()
    rows = ix.reshape(n, k)
    assert not (rows == np.arange(n)[:, None]).any()
    s = np.sort(rows, axis=1)
    assert not (s[:, 1:] == s[:, :-1]).any()
    assert (np.diff(dx.reshape(n, k), axis=1) >= 0).all()


@pytest.mark.parametrize("k,heap_k", [(15, 0), (30, 0), (15, 15)])
def test_bf16_form_exact_distances_invariants_and_recall(orc, k, heap_k):
    n = 20000
    x = ds.mixture(n, 50, 20, seed=0)
    g = P.nndescent_dense(x, k, seed=0, precision="bf16", heap_k=heap_k)
    assert P.last_stats()["heap_k"] == (heap_k or min(2 * k, 64))
    _valid(g, n, k)
    ix, dx = g["indices"].reshape(n, k), g["data"].reshape(n, k)
    for i in range(0, n, 211):
        for s in range(k):
            want = np.float32(np.sqrt(orc.sq_dist(x[i], x[ix[i, s]])))
            assert dx[i, s].view(np.uint32) == want.view(np.uint32)
    again = P.nndescent_dense(x, k, seed=0, precision="bf16", heap_k=heap_k)
    assert np.array_equal(again["indices"], g["indices"])
    q = np.random.default_rng(1).choice(n, 3000, replace=False).astype(np.uint32)
    ei, _ = orc.knn_exact64(x, k, q)
    plain = P.nndescent_dense(x, k, seed=0)
    rb = ds.recall(ix[q], ei).mean()
    rp = ds.recall(plain["indices"].reshape(n, k)[q], ei).mean()
    assert rb >= 0.95 and rb >= rp - 0.01, (rb, rp)


def test_renumbering_returns_the_callers_numbering(orc):
    # The BF16 form renumbers the points in a locality order while iterating;
    # the graph must come back in the caller's numbering, with exact distances.
    n, k = 12000, 10
    x = ds.mixture(n, 50, 15, seed=5)
    on = P.nndescent_dense(x, k, seed=1, precision="bf16", n_trees=6, reorder="on")
    assert P.last_stats()["reordered"] == 1
    off = P.nndescent_dense(x, k, seed=1, precision="bf16", n_trees=6, reorder="off")
    assert P.last_stats()["reordered"] == 0
    for g in (on, off):
        _valid(g, n, k)
        ix, dx = g["indices"].reshape(n, k), g["data"].reshape(n, k)
        for i in range(0, n, 97):
            for s in range(k):
                want = np.float32(np.sqrt(orc.sq_dist(x[i], x[ix[i, s]])))
                assert dx[i, s].view(np.uint32) == want.view(np.uint32)
    q = np.arange(0, n, 6, dtype=np.uint32)
    ei, _ = orc.knn_exact64(x, k, q)
    r_on = ds.recall(on["indices"].reshape(n, k)[q], ei).mean()
    r_off = ds.recall(off["indices"].reshape(n, k)[q], ei).mean()
    assert abs(r_on - r_off) < 0.01 and r_on > 0.95, (r_on, r_off)

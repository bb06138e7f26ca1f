"""CPU checks of pp.neighbors' conventions around the graph step
(python/polariseq/__init__.py:2514-2600) and of their oracle (oracle/conventions.py).

* The summation order k_unit_rows uses (csrc/rows_kernels.hip), restated here in
  numpy float32, reproduces np.linalg.norm bit for bit for every row width the
  kernel accepts; a plain left-to-right sum does not.
* 1 - cos from the Euclidean distance of unit rows: d^2 / 2 equals 1 - a.b.
* The control flow: k = n_neighbors - 1, exact up to 10,000 cells, the metric
  check and its ValueError text, raised before any device work.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import conventions as C

f32 = np.float32


def _block(a):
    n = len(a)
    if n < 8:
        r = f32(0)
        for v in a:
            r = f32(r + v)
        return r
    r = [f32(a[j]) for j in range(8)]
    i = 8
    while i < n - (n % 8):
        for j in range(8):
            r[j] = f32(r[j] + a[i + j])
        i += 8
    res = f32(f32(f32(r[0] + r[1]) + f32(r[2] + r[3])) + f32(f32(r[4] + r[5]) + f32(r[6] + r[7])))
    while i < n:
        res = f32(res + a[i])
        i += 1
    return res


def _split(n):
    return n // 2 - (n // 2) % 8


def _kernel_sumsq(x):
    """k_unit_rows' order: 8-way blocks of <= 128, split at n/2 rounded down to 8."""
    a = (x * x).astype(f32)
    n = len(a)
    if n <= 128:
        return _block(a)
    n2 = _split(n)
    m = n - n2
    left = _block(a[:n2])
    if m <= 128:
        return f32(left + _block(a[n2:]))
    m2 = _split(m)
    return f32(left + f32(_block(a[n2:n2 + m2]) + _block(a[n2 + m2:])))


def test_kernel_summation_order_reproduces_numpy_norm():
    rng = np.random.default_rng(3)
    seq_differs = 0
    for d in range(1, 257):
        x = (rng.normal(size=(12, d)) * rng.uniform(0.01, 100, size=(12, 1))).astype(f32)
        want = np.linalg.norm(x, axis=1)
        got = np.array([np.sqrt(_kernel_sumsq(r)) for r in x], f32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), d
        seq = np.array([np.sqrt(np.cumsum((r * r).astype(f32), dtype=f32)[-1]) for r in x], f32)
        seq_differs += int((seq.view(np.uint32) != want.view(np.uint32)).sum())
    assert seq_differs > 100  # the order matters: a sequential sum is caught


def test_unit_rows_and_cosine_conversion():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(300, 10)).astype(f32) * rng.uniform(0.2, 5.0, size=(300, 1)).astype(f32)
    x[7] = 0.0
    u = C.unit_rows(x)
    assert u.dtype == np.float32 and np.all(u[7] == 0)
    nrm = np.linalg.norm(np.delete(u, 7, axis=0).astype(np.float64), axis=1)
    assert np.abs(nrm - 1).max() < 1e-6
    a, b = u[:150].astype(np.float64), u[150:].astype(np.float64)
    d = np.sqrt(((a - b) ** 2).sum(1)).astype(f32)
    c = C.cosine_from_unit_euclidean(d).astype(np.float64)
    cos = 1.0 - (a * b).sum(1)
    keep = np.arange(150) != 7
    assert np.abs(c - cos)[keep].max() < 1e-6
    assert np.abs(d.astype(np.float64) ** 2 - cos)[keep].max() > 0.1  # the halving is there


def test_control_flow_of_the_oracle_pipeline():
    calls = []

    def nnd(search, k, seed):
        calls.append(("nndescent", k, search.shape))
        n = search.shape[0]
        return (np.arange(n + 1, dtype=np.uint32), np.zeros(n, np.uint32), np.full(n, 2.0, f32))

    def exact(emb, n_neighbors, metric):
        calls.append(("exact", n_neighbors, metric))
        n = emb.shape[0]
        return (np.arange(n + 1, dtype=np.uint32), np.zeros(n, np.uint32), np.ones(n, f32))

    def fuzzy(ip, ix, dx, n):
        return {"n": n}

    emb = np.ones((10_000, 3), f32)
    g = C.neighbors_graph(emb, 15, "euclidean", None, 0, nnd, exact, fuzzy)
    assert g["method"] == "exact" and calls[-1] == ("exact", 15, "euclidean")
    big = np.ones((10_001, 3), f32)
    g = C.neighbors_graph(big, 15, "cosine", None, 0, nnd, exact, fuzzy)
    assert g["method"] == "nndescent" and calls[-1][:2] == ("nndescent", 14)
    assert g["distances"]["data"][0] == f32(2.0)  # 2^2 / 2
    g = C.neighbors_graph(emb, 1, "euclidean", True, 0, nnd, exact, fuzzy)
    assert calls[-1][:2] == ("nndescent", 1)  # k = max(1, n_neighbors - 1)
    with pytest.raises(ValueError, match="metric must be 'euclidean' or 'cosine', not 'manhattan'"):
        C.neighbors_graph(emb, 15, "manhattan", None, 0, nnd, exact, fuzzy)


def test_library_rejects_an_unknown_metric_before_device_work():
    import polariseq_amd as P
    with pytest.raises(ValueError, match="metric must be 'euclidean' or 'cosine', not 'manhattan'"):
        P.neighbors_graph(np.ones((20, 4), f32), 5, metric="manhattan")

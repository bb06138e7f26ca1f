"""GPU parity of the exact neighbour graph (neighbors() / neighbors_dense,
core/neighbors.rs:101-241; the path pp.neighbors takes up to 10,000 cells)
against the oracle's restatement of the reference, bit for bit: row pointers,
indices and the f32 distances. Covers both metrics, duplicate points (the
reference can keep the query in its own row), zero-norm rows under cosine
(distance 1.0), n_neighbors at and above n_obs, heaps in LDS and in global
memory, the d = 50 / d = 30 register paths and the generic width, and the
device entry."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())


def _same(x, nn, metric, orc):
    g = P.neighbors_dense(x, nn, metric)
    w = orc.neighbors_exact(x, nn, metric)
    print(f"neighbors_exact n={x.shape[0]} d={x.shape[1]} n_neighbors={nn} {metric}", flush=True)
    assert np.array_equal(g["indptr"], w["indptr"]), "indptr differs from the oracle"
    assert np.array_equal(g["indices"], w["indices"]), "indices differ from the oracle"
    assert np.array_equal(g["data"].view(np.uint32), w["data"].view(np.uint32)), "distances differ"
    return g


def _dupes(n, d, distinct, seed):
    rng = np.random.default_rng(seed)
    base = ds.mixture(distinct, d, 5, seed=seed)
    return np.ascontiguousarray(base[rng.integers(0, distinct, n)])


def test_kats(orc):
    g = _same(np.array(KATS["four_cells"]["points"], np.float32), 3, "euclidean", orc)
    assert set(g["indices"][:2].tolist()) == {1, 3} and np.allclose(g["data"][:2], 1.0)
    g = _same(np.array(KATS["cosine"]["points"], np.float32), 2, "cosine", orc)
    assert g["indices"][0] == 1 and abs(g["data"][0]) < 1e-5
    # the heap's boundary rule: the later of two equal distances survives here
    x = np.array([[0.0], [-1.0], [1.0], [0.5]], np.float32)
    g = _same(x, 3, "euclidean", orc)
    assert g["indices"][:2].tolist() == [3, 2]
    # a duplicate ahead of the query takes entry 0, and the query stays in its row
    x = np.array([[0.0, 0.0], [0.0, 0.0], [1.0, 0.0], [2.0, 0.0]], np.float32)
    g = _same(x, 2, "euclidean", orc)
    assert g["indices"].tolist() == [1, 1, 1, 2]


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("n,d,nn", [(3000, 50, 15), (2500, 30, 31), (1000, 7, 10), (777, 50, 65),
                                    (513, 50, 129), (300, 16, 300)])
def test_mixture(n, d, nn, metric, orc):
    x = ds.mixture(n, d, 12, seed=n + d)
    _same(x, nn, metric, orc)


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("nn", [2, 15, 100, 200])
def test_duplicate_points(nn, metric, orc):
    x = _dupes(900, 50, 120, seed=nn)
    _same(x, nn, metric, orc)


def test_cosine_zero_rows(orc):
    x = ds.mixture(700, 50, 6, seed=4)
    x[[0, 3, 64, 699]] = 0.0
    g = _same(x, 15, "cosine", orc)
    for i in (0, 3, 64, 699):
        s, e = g["indptr"][i], g["indptr"][i + 1]
        assert (g["data"][s:e] == 1.0).all()


def test_small_and_degenerate(orc):
    x = ds.mixture(50, 50, 3, seed=1)
    _same(x, 80, "euclidean", orc)  # n_neighbors > n_obs: every other point
    g = _same(x, 1, "euclidean", orc)  # nothing but self
    assert len(g["indices"]) == 0 and (g["indptr"] == 0).all()
    g = _same(x[:1], 15, "cosine", orc)  # one cell
    assert g["indptr"].tolist() == [0, 0]


def test_device_entry_equals_host_entry(orc):
    import torch
    x = ds.mixture(2000, 50, 10, seed=7)
    g = P.neighbors_dense(x, 15)
    ip, ix, dx = P.neighbors_dense_device(torch.from_numpy(x).cuda(), 15)
    torch.cuda.synchronize()
    assert np.array_equal(ip.cpu().numpy().view(np.uint32), g["indptr"])
    assert np.array_equal(ix.cpu().numpy().view(np.uint32), g["indices"])
    assert np.array_equal(dx.cpu().numpy().view(np.uint32), g["data"].view(np.uint32))


def test_errors():
    x = ds.mixture(10, 4, 2, seed=0)
    with pytest.raises(RuntimeError, match="n_neighbors must be"):
        P.neighbors_dense(x, 0)
    with pytest.raises(RuntimeError, match="unknown metric"):
        P.neighbors_dense(x, 3, "manhattan")

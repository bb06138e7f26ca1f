"""GPU checks of the pp.neighbors graph computation (polariseq_amd/graph.py,
csrc/rows_kernels.hip) against python/polariseq/__init__.py:2514-2600.

* The unit-row scaling and the cosine conversion, bit for bit with the numpy
  expressions of the reference (oracle/conventions.py), at every width class
  (under 8, multiples of 8, one and two splits of the pairwise sum), with zero
  rows and rows below the 1e-12 floor.
* The whole approximate pipeline against the oracle's pipeline (the GPU
  algorithm's model for NN-Descent, the FP64 exact search, the fuzzy oracle):
  distances bit-exact, connectivities as the fuzzy tests accept them.
* The reference's own pp.neighbors tests (pyapi/tests/test_neighbors.py:53-66,
  82-104): the exact/approximate switch with n_neighbors - 1 stored neighbours,
  and cosine honoured on the approximate path.
* Results written into memory-mapped arrays equal the in-memory ones.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from oracle import conventions as C

pytestmark = pytest.mark.gpu

WEIGHT_ULPS = 2


def _mixture(n: int, d: int = 30, seed: int = 0) -> np.ndarray:
    """A 25-component Gaussian mixture (the reference's pp.neighbors tests use
    one of this size, test_neighbors.py:19-23; this is polariseq_amd.datasets')."""
    from polariseq_amd import datasets as ds
    return ds.mixture(n, d, 25, seed=seed)


@pytest.mark.parametrize("d", [1, 5, 8, 30, 50, 128, 129, 200, 250, 256])
def test_unit_rows_bit_exact_with_numpy(d):
    import torch
    rng = np.random.default_rng(d)
    x = (rng.normal(size=(777, d)) * rng.uniform(1e-3, 1e3, size=(777, 1))).astype(np.float32)
    x[3] = 0.0
    x[5] = 1e-20  # norm below the 1e-12 floor
    print(f"unit rows d={d}", flush=True)
    got = P.unit_rows_device(torch.from_numpy(x).cuda()).cpu().numpy()
    want = C.unit_rows(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_unit_rows_rejects_wider_rows():
    import torch
    with pytest.raises(RuntimeError, match="256"):
        P.unit_rows_device(torch.ones((4, 257), device="cuda"))


def test_cosine_conversion_bit_exact_with_numpy():
    import torch
    rng = np.random.default_rng(2)
    d = np.concatenate([rng.uniform(0, 2, 100_000), [0.0, 2.0, 1e-30, 1.4142135]]).astype(np.float32)
    t = torch.from_numpy(d.copy()).cuda()
    P.cosine_distances_device(t, 100_002)
    got = t.cpu().numpy()
    want = d.copy()
    want[:100_002] = C.cosine_from_unit_euclidean(d[:100_002])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("metric,K", [("euclidean", 15), ("cosine", 15), ("euclidean", 100),
                                      ("cosine", 100)])
def test_approximate_pipeline_matches_the_oracle_pipeline(orc, metric, K):
    x = _mixture(3000, d=20, seed=7)
    x *= np.random.default_rng(8).uniform(0.5, 2.0, size=(3000, 1)).astype(np.float32)

    def nnd(search, k, seed):
        g, _ = orc.gpu_model_nndescent(search, k, seed=seed)
        return g["indptr"], g["indices"], g["data"]

    want = C.neighbors_graph(x, K, metric, True, 3, nnd, orc.neighbors_exact,
                             orc.fuzzy_connectivities)
    wd = want["distances"]
    print(f"pipeline {metric} n_neighbors={K}", flush=True)
    got = P.neighbors_graph(x, K, metric=metric, approximate=True, seed=3)
    gd, gc = got["distances"], got["connectivities"]
    assert got["uns"]["method"] == "nndescent"
    assert np.array_equal(gd["indptr"], wd["indptr"])
    assert np.array_equal(gd["indices"], wd["indices"])
    assert np.array_equal(gd["data"].view(np.uint32), wd["data"].view(np.uint32))
    wc = want["connectivities"]
    assert np.array_equal(gc["indptr"], wc["indptr"]) and np.array_equal(gc["indices"], wc["indices"])
    ulps = np.abs(gc["data"].view(np.int32).astype(np.int64) - wc["data"].view(np.int32).astype(np.int64))
    assert ulps.max() <= WEIGHT_ULPS


def test_auto_switch_stores_n_neighbors_minus_one():
    """test_neighbors.py:53-66."""
    K = 15
    big = P.neighbors_graph(_mixture(12_000, seed=5), n_neighbors=K)
    assert big["uns"]["method"] == "nndescent"
    assert len(big["distances"]["indptr"]) == 12_001
    assert len(big["distances"]["indices"]) == 12_000 * (K - 1)
    small = P.neighbors_graph(_mixture(2_000, seed=5), n_neighbors=K)
    assert small["uns"]["method"] == "exact"
    assert len(small["distances"]["indices"]) == 2_000 * (K - 1)


def test_cosine_is_honoured_on_the_approximate_path():
    """test_neighbors.py:82-104."""
    rng = np.random.default_rng(0)
    base = rng.normal(size=(12_000, 10)).astype(np.float32)
    x = base * rng.uniform(0.2, 5.0, size=(12_000, 1)).astype(np.float32)
    e = P.neighbors_graph(x, n_neighbors=10, metric="cosine", approximate=False)["distances"]
    a = P.neighbors_graph(x, n_neighbors=10, metric="cosine", approximate=True)["distances"]
    overlap = np.mean([len(set(e["indices"][e["indptr"][i]:e["indptr"][i + 1]])
                           & set(a["indices"][a["indptr"][i]:a["indptr"][i + 1]])) / 9
                       for i in range(0, 12_000, 97)])
    assert overlap > 0.9, overlap
    assert float(np.max(a["data"])) <= 2.0 + 1e-5


@pytest.mark.parametrize("approximate", [True, False])
def test_results_into_memory_mapped_arrays(tmp_path, approximate):
    x = _mixture(4_000, seed=9)
    made = {}

    def alloc(name, shape, dtype):
        made[name] = np.memmap(tmp_path / f"{name}.bin", dtype=dtype, mode="w+", shape=shape)
        return made[name]

    ref = P.neighbors_graph(x, 15, approximate=approximate, seed=1)
    mapped = P.neighbors_graph(x, 15, approximate=approximate, seed=1, alloc=alloc)
    assert set(made) == {f"{g}_{p}" for g in ("distances", "connectivities")
                         for p in ("indptr", "indices", "data")}
    for g in ("distances", "connectivities"):
        for p in ("indptr", "indices", "data"):
            assert isinstance(mapped[g][p], np.memmap)
            assert np.array_equal(np.asarray(mapped[g][p]), ref[g][p]), (g, p)

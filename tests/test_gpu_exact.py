"""GPU parity of the exact tiled kNN: indices bit-exact against the FP64 oracle
(order = distance then index, self excluded; core/doublets.rs:544-651), distances
within 1e-5 relative. Covers the reference KATs, duplicate points (exact ties),
rows with fewer than k others, the d = 50 register path and the generic path, and
the large-k path (64 < k <= 2048) that Scrublet's k_adj needs
(core/doublets.rs:1083-1090: round(0.5 sqrt(n)) * (1 + n_sim / n))."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
RTOL = 1e-5  # distance tolerance against the FP64 oracle


def _check(x, k, queries, orc):
    idx, dist = P.knn_exact(x, k, queries)
    ei, ed = orc.knn_exact64(x, k, queries)
    assert np.array_equal(idx, ei), "indices differ from the FP64 oracle"
    fin = np.isfinite(ed)
    assert np.array_equal(np.isfinite(dist), fin)
    assert np.allclose(dist[fin], np.sqrt(ed[fin]), rtol=RTOL, atol=0)
    return idx, dist


def test_line_points_kat(orc):
    kat = KATS["line_points"]
    x = np.array(kat["points"], np.float32)
    idx, _ = _check(x, kat["k"], None, orc)
    for r, want in kat["expect_rows"].items():
        assert list(idx[int(r)]) == want


def test_four_cells_kat(orc):
    kat = KATS["four_cells"]
    x = np.array(kat["points"], np.float32)
    idx, dist = _check(x, 2, None, orc)
    assert set(idx[0]) == {1, 3} and np.allclose(dist[0], 1.0, atol=1e-5)


@pytest.mark.parametrize("n,d,k", [(2000, 50, 15), (2000, 50, 30), (1500, 50, 64), (3000, 8, 10),
                                   (1200, 20, 1), (700, 30, 33), (1800, 30, 12), (300, 3, 16)])
def test_random_with_ties_bit_exact(orc, n, d, k):
    rng = np.random.default_rng(n * 31 + d)
    x = ds.mixture(n, d, 6, seed=n + d) if d >= 8 else rng.random((n, d), dtype=np.float32)
    for i in range(0, n, 97):  # exact duplicates (doublets.rs:1668-1673)
        x[(i + 1) % n] = x[i]
    x[::13, 0] = 0.5  # shared coordinate values
    q = np.arange(0, n, 3, dtype=np.uint32)
    _check(x, k, q, orc)


def test_all_rows_and_integer_grid_ties(orc):
    # Points on an integer grid: many exactly equal distances, broken by index.
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(6), indexing="ij"), -1)
    x = g.reshape(-1, 3).astype(np.float32)
    _check(x, 26, None, orc)


def test_fewer_points_than_k(orc):
    x = np.random.default_rng(0).random((5, 4), dtype=np.float32)
    idx, dist = _check(x, 10, None, orc)
    assert (idx[:, 4:] == 0xFFFFFFFF).all() and np.isinf(dist[:, 4:]).all()


def test_many_duplicates_take_the_overflow_path(orc):
    # 300 copies of one point: more candidates tie within the margin than the
    # per-query buffer holds, so those queries are finished by the fallback.
    rng = np.random.default_rng(4)
    x = rng.random((800, 50), dtype=np.float32)
    x[100:400] = x[100]
    q = np.array([100, 101, 250, 399, 0, 500], np.uint32)
    _check(x, 15, q, orc)


def test_device_entry_matches_host_entry(orc):
    import torch
    x = ds.mixture(4000, 50, 10, seed=7)
    q = np.arange(0, 4000, 5, dtype=np.uint32)
    hi, hd = P.knn_exact(x, 15, q)
    di, dd = P.knn_exact_device(torch.from_numpy(x).cuda(), 15,
                                torch.from_numpy(q.astype(np.int32)).cuda())
    assert np.array_equal(di.cpu().numpy().view(np.uint32), hi)
    assert np.array_equal(dd.cpu().numpy(), hd)


def test_100k_subset_of_the_1p3m_workload(orc):
    """1.3M x 50 (60 components): exact kNN of a 100k random subset as its own
    dataset, on the GPU for all 100k rows. 1,000 rows are checked bit-exact
    against the oracle; every row is checked for the invariants."""
    x = ds.mixture(1_300_000, 50, 60, seed=0)
    sub = np.sort(np.random.default_rng(0).choice(1_300_000, 100_000, replace=False))
    xs = np.ascontiguousarray(x[sub])
    del x
    idx, dist = P.knn_exact(xs, 15)
    n = xs.shape[0]
    assert not (idx == np.arange(n)[:, None]).any()
    assert (np.diff(dist, axis=1) >= 0).all()
    assert all(len(set(r)) == 15 for r in idx[::997])
    rows = np.random.default_rng(1).choice(n, 1000, replace=False).astype(np.uint32)
    ei, ed = orc.knn_exact64(xs, 15, rows)
    assert np.array_equal(idx[rows], ei)
    assert np.allclose(dist[rows], np.sqrt(ed), rtol=RTOL, atol=0)


# ── large k (64 < k <= 2048) ──

@pytest.mark.parametrize("n,d,k", [(3000, 50, 65), (3000, 50, 474), (2500, 20, 1161),
                                   (2300, 50, 2048), (900, 7, 300),
                                   # n >= 64 k / 32 k: pass 1 keeps 16 / 32 per segment
                                   (8000, 30, 100), (20000, 30, 474)])
def test_large_k_bit_exact(orc, n, d, k):
    rng = np.random.default_rng(n + k)
    x = ds.mixture(n, d, 6, seed=k) if d >= 8 else rng.random((n, d), dtype=np.float32)
    for i in range(0, n, 89):  # exact duplicates
        x[(i + 1) % n] = x[i]
    q = np.arange(0, n, 7, dtype=np.uint32)
    print(f"exact large k n={n} d={d} k={k}", flush=True)
    _check(x, k, q, orc)


def test_large_k_scrublet_shape_all_rows(orc):
    """Scrublet's exact search: cells plus twice as many simulated doublets
    (means of two random cells), k_adj = round(0.5 sqrt(n)) * 3, every row."""
    rng = np.random.default_rng(5)
    cells = ds.mixture(3000, 30, 8, seed=5)
    a, b = rng.integers(0, 3000, (2, 6000))
    dbl = ((cells[a] + cells[b]) / 2).astype(np.float32)
    x = np.concatenate([cells, dbl])
    k = int(np.round(0.5 * np.sqrt(3000))) * 3
    assert k > 64
    print(f"exact scrublet shape k={k}", flush=True)
    _check(x, k, None, orc)


@pytest.mark.parametrize("n", [7000, 16000])
def test_large_k_short_lists_with_duplicates(orc, n):
    """n >= 32 k (32 smallest per segment in pass 1) and n >= 64 k (16) with 600
    copies of one point: ties beyond the candidate buffer go to the fallback."""
    y = np.random.default_rng(3).random((n, 30), dtype=np.float32)
    y[1000:1600] = y[1000]
    print(f"exact large k, short lists, duplicates, n={n}", flush=True)
    _check(y, 120, np.array([1000, 1001, 1599, 0, n - 1, 3500], np.uint32), orc)


def test_large_k_fewer_points_and_overflow(orc):
    x = np.random.default_rng(1).random((90, 6), dtype=np.float32)
    idx, dist = _check(x, 120, None, orc)
    assert (idx[:, 89:] == 0xFFFFFFFF).all() and np.isinf(dist[:, 89:]).all()
    # 700 copies of one point: ties beyond the candidate buffer go to the fallback
    y = np.random.default_rng(2).random((1500, 50), dtype=np.float32)
    y[200:900] = y[200]
    _check(y, 100, np.array([200, 201, 650, 899, 0, 1400], np.uint32), orc)

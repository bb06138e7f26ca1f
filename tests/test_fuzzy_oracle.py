"""CPU tests of the fuzzy-connectivities oracle (oracle/fuzzy_oracle.cpp) and of
the host-side checks of its GPU boundary.

The oracle is pinned by
* the umap-learn values the reference tests hold (crates/polariseq-core/src/umap.rs:2261-2276);
* the equation the bandwidth search solves, sum_j exp(-max(0, d_j - rho)/sigma) =
  log2(k + 1), solved here independently with scipy's brentq, and the
  1e-3 * mean floor in a case where only the floor keeps a weight above zero;
* the reference's HashMap formulation of the symmetric union
  (umap.rs:1891-1918), restated below over undirected pairs in numpy float32;
* the invariants of umap.rs:1954-2000 (symmetric with bitwise-equal weights,
  rows sorted, no self-loops).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from polariseq_amd import _native as N
from polariseq_amd import datasets as ds

# umap.rs:2261-2264: umap-learn 0.5.12 on the row [0, 0.5, 1, 2, 4] (self first).
UMAP_LEARN_REFERENCE = np.array([1.0, 0.754181, 0.42896995, 0.13878077], np.float32)


def _strengths_row(orc, d):
    d = np.asarray(d, np.float32)
    return orc.fuzzy_strengths(np.array([0, len(d)], np.uint32), d)


def test_smooth_knn_dist_matches_umap_learn(orc):
    w = _strengths_row(orc, [0.5, 1.0, 2.0, 4.0])
    assert np.all(np.abs(w - UMAP_LEARN_REFERENCE) < 1e-5), w
    assert abs(float(w.sum(dtype=np.float32)) - np.log2(5.0)) < 1e-4


def test_weights_are_valid_and_non_increasing(orc):
    w = _strengths_row(orc, [0.5, 1.0, 2.0, 4.0])
    assert len(w) == 4 and np.all((w >= 0) & (w <= 1))
    assert w[0] >= w[1] >= w[2]


def _brentq_weights(d: np.ndarray, target: float) -> np.ndarray:
    from scipy.optimize import brentq
    d = d.astype(np.float64)
    rho = d[d > 0][0]
    x = np.maximum(d - rho, 0.0)
    f = lambda s: np.exp(-x / s).sum() - target  # noqa: E731
    sigma = brentq(f, 1e-12, 1e6, xtol=1e-14, rtol=1e-14)
    return np.exp(-x / sigma)


def test_bandwidth_solves_the_log2_target(orc):
    """Every row's weights equal exp(-(d - rho)/sigma*) with sigma* the root of
    the log2(k+1) equation, to the bisection's tolerance. A search in the wrong
    direction (all weights 1) or an ln(k+1) target fails this."""
    rng = np.random.default_rng(3)
    data = ds.mixture(600, 12, seed=4)
    g = orc.neighbors_exact(data, 16)
    ip, dx = g["indptr"], g["data"]
    w = orc.fuzzy_strengths(ip, dx)
    rows = rng.choice(len(ip) - 1, 60, replace=False)
    worst = 0.0
    for i in rows:
        d = dx[ip[i]:ip[i + 1]]
        want = _brentq_weights(d, np.log2(len(d) + 1))
        worst = max(worst, float(np.abs(w[ip[i]:ip[i + 1]] - want).max()))
        # the mutants this must reject
        wrong_target = _brentq_weights(d, np.log(len(d) + 1))
        assert np.abs(wrong_target - want).max() > 1e-3
        assert np.abs(np.ones_like(want) - want).max() > 1e-2
    assert worst < 2e-5, worst


def test_zero_distances_and_the_mean_floor(orc):
    # all zero: rho = 0 and every x <= 0, weight 1 (umap.rs:600-606)
    assert np.array_equal(_strengths_row(orc, [0.0, 0.0, 0.0]), np.ones(3, np.float32))
    # rho is the nearest NON-zero distance: zeros and rho itself weigh 1
    w = _strengths_row(orc, [0.0, 1.0, 2.0, 3.0, 4.0])
    assert w[0] == w[1] == 1.0 and 0 < w[4] < w[3] < w[2] < 1
    # psum >= 3 > log2(5) for every sigma: the search shrinks sigma to 2^-64 and
    # only the 1e-3 * mean floor (the mean counts the self zero) keeps w[3] > 0.
    d = np.array([1.0, 1.0, 1.0, 1.0001], np.float32)
    d64 = d.astype(np.float64)
    sigma = 1e-3 * (d64.sum() / 5.0)
    want = np.float32(np.exp(-((d64[3] - 1.0) / sigma)))
    w = _strengths_row(orc, d)
    assert w[:3].tolist() == [1.0, 1.0, 1.0]
    assert abs(float(w[3]) - float(want)) <= 2e-7 and w[3] > 0.5
    # an empty row is no row
    assert len(orc.fuzzy_strengths(np.array([0, 0], np.uint32), np.zeros(0, np.float32))) == 0


def _hashmap_union(ip, ix, w):
    """umap.rs:1891-1918: fold directed edges into undirected pairs in row order."""
    edges = {}
    n = len(ip) - 1
    for i in range(n):
        for e in range(ip[i], ip[i + 1]):
            j = int(ix[e])
            key = (min(i, j), max(i, j))
            old = edges.get(key, np.float32(0.0))
            wij = np.float32(w[e])
            edges[key] = np.float32(np.float32(old + wij) - np.float32(old * wij))
    return sorted((s, t, v) for (s, t), v in edges.items() if v > np.float32(1e-5))


def _upper_edges(g):
    ip, ix, w = g["indptr"], g["indices"], g["data"]
    out = []
    for i in range(len(ip) - 1):
        for e in range(ip[i], ip[i + 1]):
            if ix[e] > i:
                out.append((i, int(ix[e]), w[e]))
    return out


@pytest.mark.parametrize("n,k,seed", [(80, 5, 7), (300, 10, 1)])
def test_union_matches_the_hashmap_formulation(orc, n, k, seed):
    data = ds.mixture(n, 8, seed=seed)
    g = orc.neighbors_exact(data, k + 1)
    fz = orc.fuzzy_connectivities(g["indptr"], g["indices"], g["data"], n)
    w = orc.fuzzy_strengths(g["indptr"], g["data"])
    want = _hashmap_union(g["indptr"], g["indices"], w)
    got = _upper_edges(fz)
    assert len(got) == len(want)
    for (gs, gt, gw), (ws, wt, ww) in zip(got, want):
        assert (gs, gt) == (ws, wt)
        assert np.float32(gw).view(np.uint32) == np.float32(ww).view(np.uint32)


def test_graph_is_symmetric_sorted_and_canonical(orc):
    n = 120
    data = ds.mixture(n, 6, seed=3)
    g = orc.neighbors_exact(data, 9)
    fz = orc.fuzzy_connectivities(g["indptr"], g["indices"], g["data"], n)
    ip, ix, w = fz["indptr"], fz["indices"], fz["data"]
    assert ip[0] == 0 and len(ip) == n + 1 and np.all(np.diff(ip.astype(np.int64)) >= 0)
    lookup = {}
    for i in range(n):
        row = ix[ip[i]:ip[i + 1]]
        assert np.all(row != i)
        assert np.all(np.diff(row.astype(np.int64)) > 0)
        for e in range(ip[i], ip[i + 1]):
            lookup[(i, int(ix[e]))] = w[e].view(np.uint32)
    for (i, j), bits in lookup.items():
        assert lookup[(j, i)] == bits


def test_repeated_columns_self_entries_and_empty_rows(orc):
    """Inputs the kNN step never makes but the function accepts: a neighbour
    named twice (folded in ascending weight order), a self entry (own and
    transposed fold into one self-loop) and empty rows."""
    ip = np.array([0, 3, 3, 5, 6, 6], np.uint32)
    ix = np.array([1, 1, 2, 0, 2, 2], np.uint32)
    dx = np.array([0.5, 0.9, 1.3, 0.2, 0.7, 0.4], np.float32)
    n = 5
    fz = orc.fuzzy_connectivities(ip, ix, dx, n)
    w = orc.fuzzy_strengths(ip, dx)

    def fold(vals):
        acc = np.float32(0.0)
        for v in vals:
            acc = np.float32(np.float32(acc + v) - np.float32(acc * v))
        return acc

    f32 = np.float32
    # row 0: own (1, w0), (1, w1), (2, w2); transposed (0 <- 2: w3)
    want = {
        0: {1: fold(sorted([f32(w[0]), f32(w[1])], key=lambda v: v.view(np.uint32))),
            2: fold([w[2], w[3]])},
        1: {0: fold(sorted([f32(w[0]), f32(w[1])], key=lambda v: v.view(np.uint32)))},
        2: {0: fold([w[3], w[2]]), 2: fold([w[4], w[4]])},
        3: {2: fold([w[5]])},
        4: {},
    }
    want[2][3] = fold([w[5]])
    for i in range(n):
        row = dict(zip(fz["indices"][fz["indptr"][i]:fz["indptr"][i + 1]].tolist(),
                       fz["data"][fz["indptr"][i]:fz["indptr"][i + 1]]))
        assert sorted(row) == sorted(want[i]), (i, row, want[i])
        for c, v in want[i].items():
            assert np.float32(row[c]).view(np.uint32) == np.float32(v).view(np.uint32), (i, c)


# ── host-side checks of the boundary (no device call is made) ──

def test_binding_errors_mirror_the_reference():
    import polariseq_amd as psq
    ip = np.array([0, 1, 2], np.uint32)
    ix = np.array([1, 0], np.uint32)
    dx = np.array([1.0, 1.0], np.float32)
    with pytest.raises(ValueError, match="indptr must have n_obs \\+ 1 entries"):
        psq.fuzzy_connectivities_into(ip, ix, dx, 3, lambda r, e: None)
    with pytest.raises(TypeError):
        psq.fuzzy_connectivities(ip.astype(np.int64), ix, dx, 2)
    with pytest.raises(TypeError):
        psq.fuzzy_connectivities(ip, ix, dx.astype(np.float64), 2)


def test_plan_rejects_bad_row_pointers_before_touching_a_device():
    L = N.lib()
    plan = ctypes.c_void_p(None)
    nnz = ctypes.c_uint64(0)
    ix = np.array([1, 0, 1], np.uint32)
    dx = np.ones(3, np.float32)
    for ip, n, m in [(np.array([0, 2, 1], np.uint32), 2, 3),  # decreasing
                     (np.array([1, 2, 3], np.uint32), 2, 3),  # indptr[0] != 0
                     (np.array([0, 2, 3], np.uint32), 2, 2),  # indptr[n] > entries given
                     (np.array([0, 2, 3], np.uint32), 3, 3)]:  # wrong length
        rc = L.psq_fuzzy_plan_new(ip.ctypes.data, len(ip), ix.ctypes.data, dx.ctypes.data, m, n, 0,
                                  ctypes.byref(plan), ctypes.byref(nnz))
        assert rc == N.PSQ_E_INVALID
        assert not plan.value

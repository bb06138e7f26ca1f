"""CPU tests of the oracle: pinned against the reference's known answers, the
published SplitMix64 stream, float64 brute force and the reference's own
invariant and recall tests (crates/polariseq-core/src/nndescent.rs:1059-1199,
python/polariseq/tests/test_neighbors.py:33-51)."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from polariseq_amd import datasets as ds

KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
GOLD = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


# ── hashing (nndescent.rs:80-112) ──

def test_splitmix_matches_published_stream(orc):
    want = [int(h, 16) for h in KATS["splitmix64_seed0"]["outputs"]]
    got = [orc.splitmix((i * GOLD) & M64) for i in range(4)]
    assert got == want


def test_hash_unit_is_a_24_bit_uniform(orc):
    vals = np.array([orc.hash_unit(7, 3, b) for b in range(20000)])
    assert ((vals >= 0) & (vals < 1)).all()
    assert np.array_equal(vals * (1 << 24), np.floor(vals * (1 << 24)))  # on the 2^-24 grid
    hist, _ = np.histogram(vals, bins=20, range=(0, 1))
    chi2 = ((hist - 1000.0) ** 2 / 1000.0).sum()
    assert chi2 < 60, chi2  # 19 dof; p < 1e-5 beyond this
    # the key mixes seed, a and b: changing any changes the draw
    assert orc.hash_unit(7, 3, 5) != orc.hash_unit(8, 3, 5) != orc.hash_unit(7, 4, 5)


def test_hash_unit_composes_splitmix_as_in_the_source(orc):
    # nndescent.rs:93: splitmix(seed ^ splitmix(a) ^ splitmix(b * GOLD)) >> 40, / 2^24
    sm = orc.splitmix
    for seed, a, b in [(0, 0, 0), (5, 1, 2), (2**63 + 11, 77, 2**40 + 3)]:
        h = sm(seed ^ sm(a) ^ sm((b * GOLD) & M64))
        assert orc.hash_unit(seed, a, b) == (h >> 40) / float(1 << 24)


def test_hash_index_range_and_spread(orc):
    n = 97
    vals = [orc.hash_index(3, i, a, n) for i in range(50) for a in range(40)]
    assert min(vals) >= 0 and max(vals) < n
    assert len(set(vals)) == n


# ── sq_dist / dot (simd.rs:22-81) ──

def _tail_vecs(n):
    i = np.arange(n)
    a = (((i * 7) % 13) * 0.25 - 1.0).astype(np.float32)
    b = (((i * 5) % 11) * 0.5 - 2.0).astype(np.float32)
    return a, b


@pytest.mark.parametrize("n", KATS["simd_tails"]["lengths"])
def test_sq_dist_and_dot_match_f64_for_all_tail_lengths(orc, n):
    a, b = _tail_vecs(n)
    d64 = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    sq64 = float(((a.astype(np.float64) - b.astype(np.float64)) ** 2).sum())
    assert abs(float(orc.dot(a, b)) - d64) < 1e-3 * max(abs(d64), 1.0)
    assert abs(float(orc.sq_dist(a, b)) - sq64) < 1e-3 * max(sq64, 1.0)


def test_sq_dist_is_symmetric_and_repeatable(orc):
    rng = np.random.default_rng(0)
    for n in (3, 50, 101):
        a = rng.normal(size=n).astype(np.float32) * 10
        b = rng.normal(size=n).astype(np.float32)
        x = orc.sq_dist(a, b)
        assert x.view(np.uint32) == orc.sq_dist(b, a).view(np.uint32) == orc.sq_dist(a, b).view(np.uint32)


def test_sq_dist_follows_the_eight_lane_order(orc):
    # Values chosen so the summation order changes the f32 result: the 8-lane
    # order (lane l sums dims l, l+8, ...; reduce ((0+4)+(1+5))+((2+6)+(3+7));
    # then the sequential tail) must be reproduced bit for bit.
    rng = np.random.default_rng(1)
    hits = 0
    for trial in range(200):
        n = 50
        a = (rng.normal(size=n) * np.exp(rng.normal(size=n) * 3)).astype(np.float32)
        b = np.zeros(n, np.float32)
        t = (a - b).astype(np.float32) ** 2
        lanes = np.zeros(8, np.float32)
        for ch in range(n // 8):
            lanes = (lanes + t[ch * 8:(ch + 1) * 8]).astype(np.float32)
        red = np.float32(np.float32(np.float32(lanes[0] + lanes[4]) + np.float32(lanes[1] + lanes[5]))
                         + np.float32(np.float32(lanes[2] + lanes[6]) + np.float32(lanes[3] + lanes[7])))
        tail = np.float32(0)
        for i in range((n // 8) * 8, n):
            tail = np.float32(tail + t[i])
        want = np.float32(red + tail)
        seq = np.float32(0)
        for v in t:
            seq = np.float32(seq + v)
        got = orc.sq_dist(a, b)
        assert got.view(np.uint32) == want.view(np.uint32)
        hits += int(seq.view(np.uint32) != want.view(np.uint32))
    assert hits > 20  # the order is observable, so the check above is sensitive to it


# ── exact kNN oracle (doublets.rs:544-651 order, FP64) ──

def test_exact_oracle_line_points_kat(orc):
    kat = KATS["line_points"]
    x = np.array(kat["points"], np.float32)
    idx, _ = orc.knn_exact64(x, kat["k"])
    for r, want in kat["expect_rows"].items():
        assert list(idx[int(r)]) == want


def test_exact_oracle_four_cells_kat(orc):
    kat = KATS["four_cells"]
    x = np.array(kat["points"], np.float32)
    idx, d2 = orc.knn_exact64(x, kat["n_neighbors"] - 1)
    assert set(idx[0]) == set(kat["expect_row0_set"])
    assert np.allclose(np.sqrt(d2[0]), kat["expect_row0_dist"], atol=kat["tol"])


def _brute64(x, k, queries):
    x64 = x.astype(np.float64)
    out_i, out_d = [], []
    for q in queries:
        d = ((x64 - x64[q]) ** 2).sum(axis=1)
        order = np.lexsort((np.arange(len(x)), d))
        order = order[order != q][:k]
        out_i.append(order)
        out_d.append(d[order])
    return np.array(out_i, np.uint32), np.array(out_d)


@pytest.mark.parametrize("n,d,k", [(300, 6, 25), (500, 50, 15), (64, 3, 63), (40, 2, 5)])
def test_exact_oracle_matches_numpy_with_duplicates(orc, n, d, k):
    rng = np.random.default_rng(n + d)
    x = rng.random((n, d), dtype=np.float32)
    for i in range(0, n, 7):  # exact duplicates force ties broken by index
        x[(i + 1) % n] = x[i]
    x[: n // 4, 0] = np.round(x[: n // 4, 0] * 4) / 4  # and near-degenerate coordinates
    q = np.arange(0, n, 3, dtype=np.uint32)
    idx, d2 = orc.knn_exact64(x, k, q)
    bi, bd = _brute64(x, min(k, n - 1), q)
    assert np.array_equal(idx[:, : bi.shape[1]], bi)
    # numpy's pairwise float64 sum may round differently; the oracle's sum is sequential
    assert np.allclose(d2[:, : bd.shape[1]], bd, rtol=1e-12, atol=1e-12)


def test_exact_oracle_pads_when_fewer_than_k(orc):
    x = np.arange(8, dtype=np.float32).reshape(4, 2)
    idx, d2 = orc.knn_exact64(x, 5)
    assert (idx[:, 3:] == 0xFFFFFFFF).all() and np.isinf(d2[:, 3:]).all()
    assert not (idx[:, :3] == np.arange(4)[:, None]).any()


# ── the reference's exact path (neighbors.rs:101-241) ──

def test_reference_exact_path_kats(orc):
    g = orc.neighbors_exact(np.array(KATS["four_cells"]["points"], np.float32), 3)
    s, e = g["indptr"][0], g["indptr"][1]
    assert set(g["indices"][s:e]) == {1, 3} and np.allclose(g["data"][s:e], 1.0, atol=1e-5)
    g = orc.neighbors_exact(np.array(KATS["self_excluded"]["points"], np.float32), 2)
    for i in range(3):
        assert i not in g["indices"][g["indptr"][i]:g["indptr"][i + 1]]
    emb = (np.arange(30, dtype=np.float32) * 0.1).reshape(10, 3)
    g = orc.neighbors_exact(emb, 5)
    assert (np.diff(g["indptr"]) == 4).all()
    g = orc.neighbors_exact(np.array(KATS["cosine"]["points"], np.float32), 2, "cosine")
    assert g["indices"][0] == 1 and abs(g["data"][0]) < 1e-5


def _row(g, i):
    s, e = g["indptr"][i], g["indptr"][i + 1]
    return g["indices"][s:e].tolist(), g["data"][s:e].tolist()


def test_reference_exact_path_heap_traces(orc):
    # Hand traces of knn_for_one (neighbors.rs:151-223) on a line.
    # x = 0, -1, 1, 0.5; query 0 with k_with_self = 3: points 0, 1, 2 are pushed
    # (heap [(1,1), (0,0), (1,2)] after the sift-ups); point 3 (0.25) replaces the
    # root (1,1) and sifts below (1,2). Point 1 is evicted at the boundary tie and
    # point 2, the later index, survives; sorted: (0,0), (0.25,3), (1,2).
    x = np.array([[0.0], [-1.0], [1.0], [0.5]], np.float32)
    ix, dx = _row(orc.neighbors_exact(x, 3), 0)
    assert ix == [3, 2] and dx == [0.5, 1.0]
    # Duplicates x0 = x1: both distances 0 tie, the stable sort keeps the heap's
    # array order [(0,0), (0,1)] and entry 0 (point 0) is dropped as "self", so
    # row 1 keeps the query itself and row 0 keeps its duplicate. Query 2 pushes
    # (1,0), (1,1), then replaces the root (1,0) with itself: point 1 remains.
    x = np.array([[0.0, 0.0], [0.0, 0.0], [1.0, 0.0], [2.0, 0.0]], np.float32)
    g = orc.neighbors_exact(x, 2)
    assert _row(g, 0) == ([1], [0.0]) and _row(g, 1) == ([1], [0.0])
    assert _row(g, 2) == ([1], [1.0]) and _row(g, 3) == ([2], [1.0])


def test_reference_exact_path_cosine_zero_rows(orc):
    # cosine_dist (neighbors.rs:232-241) is 1.0 whenever a norm is 0.
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, 6)).astype(np.float32)
    x[[5, 17]] = 0.0
    g = orc.neighbors_exact(x, 8, "cosine")
    for i in (5, 17):
        _, dx = _row(g, i)
        assert dx == [1.0] * 7
    for i in range(40):
        ix, dx = _row(g, i)
        for j, v in zip(ix, dx):
            if j in (5, 17):
                assert v == 1.0


def test_reference_exact_path_without_ties_is_the_sorted_top(orc):
    # With distinct distances the heap holds the k smallest: the row is the query's
    # stable argsort of the reference's f32 distances (sq_dist, pinned by its own
    # tests) after the query itself, with square roots.
    rng = np.random.default_rng(11)
    x = rng.normal(size=(120, 9)).astype(np.float32)
    g = orc.neighbors_exact(x, 6)
    for i in range(0, 120, 7):
        d = np.array([orc.sq_dist(x[i], x[j]) for j in range(120)], np.float32)
        assert len(np.unique(d)) == 120
        order = np.argsort(d, kind="stable")[:6]
        assert order[0] == i
        ix, dx = _row(g, i)
        assert ix == order[1:].tolist()
        assert np.array_equal(np.array(dx, np.float32), np.sqrt(d[order[1:]]))
    # n_neighbors above n_obs: every other point, each row n - 1 long
    g = orc.neighbors_exact(x[:10], 25)
    assert (np.diff(g["indptr"]) == 9).all()


# ── the reference NN-Descent restated (nndescent.rs:1059-1199) ──

def _valid(g, n, k):
    ip, ix, dx = g["indptr"], g["indices"], g["data"]
    assert ip[0] == 0 and ip[-1] == len(ix)
    for i in range(n):
        row = ix[ip[i]:ip[i + 1]]
        assert len(row) == k, (i, len(row))
        assert i not in row
        assert len(set(row.tolist())) == k
        assert (np.diff(dx[ip[i]:ip[i + 1]]) >= 0).all()


def test_port_graph_structure_is_valid(orc):
    x = ds.clustered(2000, 8, 5, seed=1)
    g, _ = orc.nndescent_port(x, 10, seed=1)
    _valid(g, 2000, 10)


def test_port_recall_is_high_on_structured_data(orc):
    x = ds.blobs_20d(6000, 20, seed=9)
    g, _ = orc.nndescent_port(x, 15, seed=4)
    ei, _ = orc.knn_exact64(x, 15, np.arange(300, dtype=np.uint32))
    r = ds.recall(g["indices"].reshape(6000, 15)[:300], ei).mean()
    assert r >= 0.95, r


def test_port_gather_bound_does_not_change_the_graph(orc):
    x = ds.clustered(20000, 12, 15, seed=11)
    ref, _ = orc.nndescent_port(x, 10, seed=9, gather_max=2**62)
    for gm in (1, 64, 1000):
        g, _ = orc.nndescent_port(x, 10, seed=9, gather_max=gm)
        assert np.array_equal(g["indptr"], ref["indptr"])
        assert np.array_equal(g["indices"], ref["indices"])
        assert np.array_equal(g["data"].view(np.uint32), ref["data"].view(np.uint32))


def test_port_is_deterministic_and_thread_count_independent(orc):
    x = ds.clustered(3000, 10, 7, seed=5)
    a, _ = orc.nndescent_port(x, 12, seed=42, threads=1)
    b, _ = orc.nndescent_port(x, 12, seed=42, threads=8)
    c, _ = orc.nndescent_port(x, 12, seed=42, threads=3)
    for g in (b, c):
        assert np.array_equal(a["indices"], g["indices"])
        assert np.array_equal(a["data"].view(np.uint32), g["data"].view(np.uint32))


def test_port_tiny_inputs_work(orc):
    kat = KATS["tiny_inputs"]
    x = ds.clustered(kat["n"], kat["d"], 2, seed=0)
    g, _ = orc.nndescent_port(x, kat["k"])
    assert (np.diff(g["indptr"]) == kat["expect_per_row"]).all()


def test_port_recall_vs_exact_on_mixture(orc):
    # python/polariseq/tests/test_neighbors.py:19-39: 20k x 30, 25 centres x 5.
    rng = np.random.default_rng(0)
    centers = rng.normal(size=(25, 30)) * 5
    lab = rng.integers(0, 25, size=20000)
    x = (centers[lab] + rng.normal(size=(20000, 30))).astype(np.float32)
    g, _ = orc.nndescent_port(x, 15, seed=42)
    q = np.arange(0, 20000, 5, dtype=np.uint32)
    ei, _ = orc.knn_exact64(x, 15, q)
    r = ds.recall(g["indices"].reshape(20000, 15)[q], ei)
    assert r.mean() >= 0.97 and np.percentile(r, 5) >= 0.85, (r.mean(), np.percentile(r, 5))


# ── the GPU algorithm's CPU statement ──

@pytest.mark.parametrize("init", [0, 1])
def test_model_graph_structure_and_determinism(orc, init):
    x = ds.clustered(2000, 8, 5, seed=1)
    a, _ = orc.gpu_model_nndescent(x, 10, seed=1, init=init, threads=1)
    b, _ = orc.gpu_model_nndescent(x, 10, seed=1, init=init, threads=8)
    _valid(a, 2000, 10)
    assert np.array_equal(a["indices"], b["indices"])
    assert np.array_equal(a["data"].view(np.uint32), b["data"].view(np.uint32))


def test_model_distances_are_the_reference_sq_dist(orc):
    x = ds.mixture(1500, 50, 6, seed=2)
    g, _ = orc.gpu_model_nndescent(x, 15, seed=0)
    ip, ix, dx = g["indptr"], g["indices"], g["data"]
    for i in range(0, 1500, 37):
        for s in range(ip[i], ip[i + 1]):
            want = np.sqrt(orc.sq_dist(x[i], x[ix[s]]))
            assert dx[s].view(np.uint32) == np.float32(want).view(np.uint32)


def test_model_recall_matches_the_port_on_the_cpu_workload(orc):
    # 20k x 50, 20-component mixture, k=15, seed 0 (the CPU-runnable case).
    x = ds.mixture(20000, 50, 20, seed=0)
    q = np.random.default_rng(1).choice(20000, 4000, replace=False).astype(np.uint32)
    ei, _ = orc.knn_exact64(x, 15, q)
    gp, _ = orc.nndescent_port(x, 15, seed=0)
    gm, _ = orc.gpu_model_nndescent(x, 15, seed=0)
    rp = ds.recall(gp["indices"].reshape(-1, 15)[q], ei).mean()
    rm = ds.recall(gm["indices"].reshape(-1, 15)[q], ei).mean()
    assert rm >= rp - 0.002, (rm, rp)
    assert rm >= 0.95


def test_model_wider_rows_output_the_first_k_and_raise_recall(orc):
    # Rows of heap_k > k entries while iterating; the graph is each row's first k.
    x = ds.mixture(8000, 50, 12, seed=3)
    q = np.arange(0, 8000, 4, dtype=np.uint32)
    ei, _ = orc.knn_exact64(x, 10, q)
    g10, st10 = orc.gpu_model_nndescent(x, 10, seed=2, n_trees=4)
    g20, st20 = orc.gpu_model_nndescent(x, 10, seed=2, n_trees=4, heap_k=20)
    _valid(g20, 8000, 10)
    r10 = ds.recall(g10["indices"].reshape(-1, 10)[q], ei).mean()
    r20 = ds.recall(g20["indices"].reshape(-1, 10)[q], ei).mean()
    assert r20 > r10, (r20, r10)
    # candidates and leaf size follow the output k, slots and the stop rule the row size
    assert st20.max_candidates == st10.max_candidates == 20
    assert st20.leaf_size == st10.leaf_size == 10
    assert st20.slots == 128 and st10.slots == 64
    assert st20.stop_at == int(0.001 * 8000 * 20)

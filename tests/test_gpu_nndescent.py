"""GPU parity of NN-Descent.

* Bit-exact against the oracle's statement of the GPU algorithm
  (oracle/gpu_model.cpp): same indptr, indices and distance bits.
* The reference's own invariants (crates/polariseq-core/src/nndescent.rs:1059-1199):
  k per row, no self, no duplicates, sorted rows, determinism, the in-place form
  equal to the allocating form, tiny inputs.
* Recall at least the reference restatement's (oracle/nndescent_port.cpp) on the
  same input, against exact kNN.
* Distances are the reference's f32 sq_dist, square-rooted, bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def _same(g, m):
    assert np.array_equal(g["indptr"], m["indptr"]), "indptr differs"
    assert np.array_equal(g["indices"], m["indices"]), "indices differ"
    assert np.array_equal(g["data"].view(np.uint32), m["data"].view(np.uint32)), "distances differ"


@pytest.mark.parametrize("n,d,k,seed", [(5, 3, 10, 0), (100, 8, 10, 1), (3000, 50, 15, 2),
                                        (20000, 50, 15, 0), (5000, 20, 30, 3), (2000, 50, 1, 4),
                                        (4000, 50, 30, 5), (1000, 64, 7, 6),
                                        (3000, 128, 15, 8), (1500, 256, 20, 9)])
def test_bit_exact_against_the_oracle_random_start(orc, n, d, k, seed):
    x = ds.mixture(n, d, 8, seed=seed) if n > 20 else ds.clustered(n, d, 2, seed=0)
    g = P.nndescent_dense(x, k, seed=seed, init="random")
    m, _ = orc.gpu_model_nndescent(x, k, seed=seed, init=1)
    _same(g, m)


@pytest.mark.parametrize("n,d,k,seed,trees", [(12, 3, 4, 0, 3), (3000, 50, 15, 2, 0),
                                              (20000, 50, 15, 0, 0), (5000, 20, 30, 3, 6),
                                              (9000, 50, 10, 1, 12), (2500, 8, 16, 7, 5)])
def test_bit_exact_against_the_oracle_forest_start(orc, n, d, k, seed, trees):
    x = ds.mixture(n, d, 8, seed=seed) if n > 20 else ds.clustered(n, d, 2, seed=0)
    g = P.nndescent_dense(x, k, seed=seed, n_trees=trees)
    m, _ = orc.gpu_model_nndescent(x, k, seed=seed, n_trees=trees, init=0)
    _same(g, m)


def test_forest_with_duplicates_and_oversized_leaves(orc):
    # 400 copies of one point force hash-bit splits; leaf_size 40 exercises
    # groups of 64 lanes.
    x = ds.mixture(3000, 50, 5, seed=4)
    x[1000:1400] = x[1000]
    for ls in (0, 40):
        g = P.nndescent_dense(x, 12, seed=5, n_trees=4, leaf_size=ls)
        m, _ = orc.gpu_model_nndescent(x, 12, seed=5, n_trees=4, leaf_size=ls, init=0)
        _same(g, m)


def test_options_reach_the_kernels(orc):
    x = ds.mixture(6000, 50, 12, seed=9)
    kw = dict(max_iterations=3, max_candidates=20, delta=0.0)
    g = P.nndescent_dense(x, 12, seed=11, init="random", **kw)
    m, _ = orc.gpu_model_nndescent(x, 12, seed=11, init=1, **kw)
    _same(g, m)
    assert P.last_stats()["iterations"] == 3


@pytest.mark.parametrize("init,model_init", [("random", 1), ("rp_forest", 0)])
def test_wider_rows_bit_exact_against_the_oracle(orc, init, model_init):
    x = ds.mixture(6000, 50, 10, seed=4)
    g = P.nndescent_dense(x, 12, seed=3, init=init, n_trees=5, heap_k=24)
    st = P.last_stats()
    assert st["k"] == 12 and st["heap_k"] == 24
    m, _ = orc.gpu_model_nndescent(x, 12, seed=3, init=model_init, n_trees=5, heap_k=24)
    _same(g, m)


# Rows of 15, 20, 26 and 30 entries with leaves of at most 16 points take the
# leaf-pair kernels whose row length is fixed at compile time (d = 50 and any d).
@pytest.mark.parametrize("d,heap_k", [(50, 20), (50, 26), (50, 30), (20, 15), (20, 26), (33, 30)])
def test_leaf_pairs_of_fixed_row_lengths_bit_exact_against_the_oracle(orc, d, heap_k):
    x = ds.mixture(7000, d, 12, seed=heap_k)
    print(f"fixed-length leaf pairs d={d} heap_k={heap_k}", flush=True)
    g = P.nndescent_dense(x, 15, seed=5, n_trees=4, heap_k=heap_k, max_iterations=2)
    assert P.last_stats()["heap_k"] == heap_k
    m, _ = orc.gpu_model_nndescent(x, 15, seed=5, init=0, n_trees=4, heap_k=heap_k, max_iterations=2)
    _same(g, m)


@pytest.mark.parametrize("n,d,k,seed,init,heap_k", [(4000, 50, 65, 1, 0, 0), (5000, 20, 99, 2, 0, 0),
                                                     (3000, 50, 99, 3, 1, 0), (3000, 16, 150, 4, 0, 0),
                                                     (2500, 50, 200, 5, 1, 0), (4000, 50, 40, 6, 0, 130),
                                                     (3000, 50, 256, 7, 0, 0)])
def test_rows_wider_than_a_wave_bit_exact_against_the_oracle(orc, n, d, k, seed, init, heap_k):
    # rows of more than 64 entries (sampling with 2 or 4 forward vectors, the
    # LDS-sorted merge, one block per leaf for the leaf pairs); pp.neighbors with
    # n_neighbors = 100 asks for k = 99 (nndescent.rs accepts any k)
    x = ds.mixture(n, d, 8, seed=seed)
    print(f"wide rows n={n} d={d} k={k} init={init} heap_k={heap_k}", flush=True)
    g = P.nndescent_dense(x, k, seed=seed, init="rp_forest" if init == 0 else "random", n_trees=4,
                          heap_k=heap_k)
    m, _ = orc.gpu_model_nndescent(x, k, seed=seed, init=init, n_trees=4, heap_k=heap_k)
    _same(g, m)
    ip, ix = g["indptr"], g["indices"]
    assert (np.diff(ip) == k).all()
    for i in range(0, n, 97):
        row = ix[ip[i]:ip[i + 1]]
        assert i not in row and len(set(row.tolist())) == k


def test_structure_determinism_and_in_place_form():
    x = ds.clustered(2000, 8, 5, seed=1)
    a = P.nndescent_dense(x, 10, seed=1, init="random")
    b = P.nndescent_dense(x, 10, seed=1, init="random")
    assert np.array_equal(a["indices"], b["indices"])
    assert np.array_equal(a["data"].view(np.uint32), b["data"].view(np.uint32))
    n, k = 2000, 10
    ip, ix, dx = a["indptr"], a["indices"], a["data"]
    for i in range(n):
        row = ix[ip[i]:ip[i + 1]]
        assert len(row) == k and i not in row and len(set(row.tolist())) == k
        assert (np.diff(dx[ip[i]:ip[i + 1]]) >= 0).all()
    # nndescent.rs:1059-1079: dirty caller buffers give the same graph
    oip = np.full(n + 1, 7, np.uint32)
    oix = np.full(n * k, 9, np.uint32)
    odx = np.full(n * k, 1.0, np.float32)
    r = P.nndescent_dense(x, k, 0, 1, out=(oip, oix, odx), init="random")
    assert r["nnz"] == len(ix) and r["n_obs"] == n
    assert np.array_equal(oip, ip) and np.array_equal(oix, ix)
    assert np.array_equal(odx.view(np.uint32), dx.view(np.uint32))


def test_tiny_inputs_work():
    x = ds.clustered(5, 3, 2, seed=0)
    g = P.nndescent_dense(x, 10, init="random")
    assert (np.diff(g["indptr"]) == 4).all()


def test_one_point_is_the_reference_graph(orc):
    # nndescent.rs:727, 836-839: k = 1 and the random fill's only draw is the
    # point itself, so the graph is one self entry at distance 0.
    x = np.array([[0.5, -2.0, 3.0]], np.float32)
    want, _ = orc.nndescent_port(x, 15)
    for init in ("rp_forest", "random"):
        g = P.nndescent_dense(x, 15, init=init)
        assert g["indptr"].tolist() == want["indptr"].tolist() == [0, 1]
        assert g["indices"].tolist() == want["indices"].tolist() == [0]
        assert g["data"].tolist() == want["data"].tolist() == [0.0]
    m, _ = orc.gpu_model_nndescent(x, 15)
    assert m["indices"].tolist() == [0] and m["data"].tolist() == [0.0]


def test_calls_from_two_threads_give_the_graphs_of_single_calls():
    # ctypes releases the GIL, so the two threads' calls overlap on the one
    # device; the library holds the device for each whole call.
    import threading
    x = ds.mixture(60000, 50, 20, seed=4)
    y = ds.mixture(30000, 50, 10, seed=5)
    want = {"x": P.nndescent_dense(x, 15, seed=1), "y": P.nndescent_dense(y, 10, seed=2),
            "q": P.knn_exact(x, 15, np.arange(0, 60000, 7, dtype=np.uint32))}
    got, errors = {}, []

    def run(name):
        try:
            for r in range(3):
                if name == "x":
                    got[(name, r)] = P.nndescent_dense(x, 15, seed=1)
                elif name == "y":
                    got[(name, r)] = P.nndescent_dense(y, 10, seed=2)
                else:
                    got[(name, r)] = P.knn_exact(x, 15, np.arange(0, 60000, 7, dtype=np.uint32))
        except Exception as e:  # re-raised below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(nm,)) for nm in ("x", "y", "q")]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for (name, r), g in got.items():
        if name == "q":
            assert np.array_equal(g[0], want["q"][0]) and np.array_equal(g[1], want["q"][1])
        else:
            for key in ("indptr", "indices", "data"):
                assert np.array_equal(g[key], want[name][key]), (name, r, key)


def test_distances_are_the_reference_sq_dist(orc):
    x = ds.mixture(3000, 50, 6, seed=3)
    g = P.nndescent_dense(x, 15, init="random")
    ip, ix, dx = g["indptr"], g["indices"], g["data"]
    for i in range(0, 3000, 53):
        for s in range(ip[i], ip[i + 1]):
            want = np.float32(np.sqrt(orc.sq_dist(x[i], x[ix[s]])))
            assert dx[s].view(np.uint32) == want.view(np.uint32)


@pytest.mark.parametrize("init", ["rp_forest", "random"])
def test_recall_at_least_the_reference_restatement(orc, init, not_below):
    # 20k x 50 cells, 20-component mixture, k = 15, seed 0 (the CPU-runnable case).
    # Every cell is a query; three seeds of both searches.
    x = ds.mixture(20000, 50, 20, seed=0)
    ei, _ = orc.knn_exact64(x, 15)
    diffs = []
    for seed in (0, 1, 2):
        gp, _ = orc.nndescent_port(x, 15, seed=seed)
        g = P.nndescent_dense(x, 15, seed=seed, init=init)
        rp = ds.recall(gp["indices"].reshape(-1, 15), ei)
        rg = ds.recall(g["indices"].reshape(-1, 15), ei)
        assert rg.mean() >= 0.95
        diffs.append(rg - rp)
    not_below(np.concatenate(diffs))


def test_device_entry_and_workspace_reuse(orc):
    import torch
    x = ds.mixture(5000, 50, 9, seed=8)
    t = torch.from_numpy(x).cuda()
    host = P.nndescent_dense(x, 15, seed=2, init="random")
    for _ in range(2):  # the cached working set is reused, results unchanged
        ip, ix, dx, nnz = P.nndescent_device(t, 15, seed=2, init="random")
        assert nnz == len(host["indices"])
        assert np.array_equal(ix[:nnz].cpu().numpy().view(np.uint32), host["indices"])
        assert np.array_equal(dx[:nnz].cpu().numpy().view(np.uint32), host["data"].view(np.uint32))
    # a smaller problem after a larger one reuses the larger buffers
    small = P.nndescent_dense(x[:700], 10, seed=2, init="random")
    m, _ = orc.gpu_model_nndescent(x[:700], 10, seed=2, init=1)
    _same(small, m)


def test_device_budget_surfaces_as_an_error():
    from polariseq_amd import _native as N
    P.release_workspace()
    N.lib().psq_set_device_budget(1 << 20)
    try:
        with pytest.raises(N.PsqError, match="budget"):
            P.nndescent_dense(ds.mixture(20000, 50, 5, seed=0), 15, init="random")
    finally:
        N.lib().psq_set_device_budget(0)
    g = P.nndescent_dense(ds.mixture(500, 50, 5, seed=0), 15, init="random")
    assert len(g["indices"]) == 500 * 15


def test_trace_prints_the_stage_times(capfd, monkeypatch):
    # nndescent.rs:757-1001: POLARISEQ_TRACE (any value) prints per-stage times
    # to stderr; here one line per stage and iteration, then the total
    x = ds.mixture(3000, 50, 6, seed=2)
    monkeypatch.setenv("POLARISEQ_TRACE", "1")
    P.nndescent_dense(x, 10, seed=1, n_trees=3)
    it = P.last_stats()["iterations"]
    err = capfd.readouterr().err
    lines = [ln for ln in err.splitlines() if ln.startswith("[nndescent]")]
    assert lines[0].startswith("[nndescent] rp-forest build + leaf pairs (3 trees, leaf 10):")
    assert sum(": candidates:" in ln for ln in lines) == it
    assert sum("changed:" in ln for ln in lines) == it
    assert lines[-2].startswith("[nndescent] extract:") and lines[-1].startswith("[nndescent] total:")
    monkeypatch.delenv("POLARISEQ_TRACE")
    P.nndescent_dense(x, 10, seed=1, n_trees=3)
    assert "[nndescent]" not in capfd.readouterr().err

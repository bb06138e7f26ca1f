"""Launches of more blocks than one grid dimension holds.

A dispatch counts fewer than 2^32 work-items per grid dimension, so the
one-block-per-row kernels (sampling, joins, merge, leaf pairs, subtree finish,
exact-graph sorts) use a two-dimensional grid past 2^22 blocks. The reference
accepts any number of cells (crates/polariseq-core/src/nndescent.rs:699-752).

* The two-dimensional layout, forced at small sizes (PSQ_GRID_X_BLOCKS), gives
  the graphs of the one-dimensional one bit for bit, and the oracle's.
* A graph over 2^26 + 4096 cells (more rows than a 64-lane-per-row grid of one
  dimension can hold): rows of k distinct non-self neighbours, sorted, with the
  reference's sq_dist bit for bit, and recall against exact kNN for queries at
  both ends of the index range.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


class narrow_grid:
    def __init__(self, blocks: int):
        self.blocks = str(blocks)

    def __enter__(self):
        os.environ["PSQ_GRID_X_BLOCKS"] = self.blocks

    def __exit__(self, *exc):
        os.environ.pop("PSQ_GRID_X_BLOCKS", None)


def _same(a, b):
    for key in ("indptr", "indices"):
        assert np.array_equal(a[key], b[key]), key
    assert np.array_equal(a["data"].view(np.uint32), b["data"].view(np.uint32))


@pytest.mark.parametrize("precision,init,k", [("fp32", "rp_forest", 15), ("fp32", "random", 10),
                                               ("bf16", "rp_forest", 15)])
def test_two_dimensional_grids_give_the_same_graph(orc, precision, init, k):
    x = ds.mixture(20000, 50, 12, seed=3)
    print(f"2-D grid {precision} {init} k={k}", flush=True)
    want = P.nndescent_dense(x, k, seed=2, init=init, precision=precision, n_trees=6)
    with narrow_grid(64):  # 20000 rows -> 313 rows of 64 blocks
        got = P.nndescent_dense(x, k, seed=2, init=init, precision=precision, n_trees=6)
    _same(got, want)
    if precision == "fp32":
        m, _ = orc.gpu_model_nndescent(x, k, seed=2, n_trees=6, init=0 if init == "rp_forest" else 1)
        _same(got, m)


def test_two_dimensional_grids_in_the_exact_graph():
    x = ds.mixture(3000, 20, 6, seed=1)
    for nn in (15, 100):  # the 64-lane sort and the LDS sort
        want = P.neighbors_dense(x, nn)
        with narrow_grid(8):
            got = P.neighbors_dense(x, nn)
        _same(got, want)


@pytest.mark.timeout(600)
def test_more_cells_than_one_grid_dimension_holds(orc):
    import torch
    n, d, k = (1 << 26) + 4096, 8, 10
    print(f"large n={n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(0)
    centres = torch.randn((256, d), generator=g, device="cuda") * 6.0
    lab = torch.randint(0, 256, (n,), generator=g, device="cuda")
    x = centres[lab] + torch.randn((n, d), generator=g, device="cuda")
    del lab
    ip, ix, dx, nnz = P.nndescent_device(x, k, seed=1, n_trees=6, max_iterations=6)
    torch.cuda.synchronize()
    assert nnz == n * k
    st = P.last_stats()
    print(f"large n: {st['ms_total']:.0f} ms, {st['iterations']} iterations", flush=True)
    rng = np.random.default_rng(5)
    rows = np.concatenate([np.arange(0, 1024), np.arange(n - 2048, n),
                           rng.integers(0, n, 2048)]).astype(np.int64)
    ipr = ip.cpu().numpy().view(np.uint32)
    assert ipr[-1] == n * k and np.array_equal(ipr[rows], (rows * k).astype(np.uint32))
    sel = torch.from_numpy(rows).cuda()
    nb = ix.view(n, k)[sel].cpu().numpy().view(np.uint32).astype(np.int64)
    nd = dx.view(n, k)[sel].cpu().numpy()
    xs = x[sel].cpu().numpy()
    xn = x[torch.from_numpy(nb.ravel()).cuda()].cpu().numpy().reshape(len(rows), k, d)
    for r in range(len(rows)):
        assert rows[r] not in nb[r] and len(set(nb[r].tolist())) == k
        assert (np.diff(nd[r]) >= 0).all()
    for r in range(0, len(rows), 7):
        for s in range(k):
            want = np.float32(np.sqrt(orc.sq_dist(xs[r], xn[r, s])))
            assert nd[r, s].view(np.uint32) == want.view(np.uint32)
    # recall against exact kNN for the queries at both ends of the index range
    q = torch.from_numpy(rows[:3072].astype(np.int32)).cuda()
    ei, _ = P.knn_exact_device(x, k, q)  # self excluded
    ei = ei.cpu().numpy()
    rec = ds.recall(nb[:3072], ei)
    print(f"large n recall {rec.mean():.4f} (last rows {rec[1024:].mean():.4f})", flush=True)
    assert rec.mean() > 0.95 and rec[1024:].mean() > 0.95

"""GPU parity of the fuzzy connectivities (csrc/fuzzy_kernels.hip).

* Against the oracle (oracle/fuzzy_oracle.cpp) on the same graph: row pointers
  and column indices bit-exact; weights within WEIGHT_ULPS units in the last
  place of float32, and bit-exact for at least EXACT_FRAC of them. The one
  source of difference is the f64 exponential: the GPU's and the host libm's
  may differ in the last bit of a double, which moves a float32 weight only when
  the double lies within that bit of a float32 rounding boundary.
* The reference's invariants (crates/polariseq-core/src/umap.rs:1954-2000) at a
  size the oracle is not run on: symmetric with bitwise-equal weights, sorted
  rows, no self-loops.
* Row sizes of all three merge paths: a wave per row (<= 256 combined entries),
  a block per row in LDS (<= 8192) and the global sort for hubs named by more
  rows than that.
* The binding's forms: allocating, into caller arrays through ``alloc`` with
  its size check, and device tensors from NN-Descent left in HBM.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu

WEIGHT_ULPS = 2
EXACT_FRAC = 0.9999


def _same(got, want):
    assert np.array_equal(got["indptr"], want["indptr"]), "indptr differs"
    assert np.array_equal(got["indices"], want["indices"]), "indices differ"
    a = got["data"].view(np.int32).astype(np.int64)
    b = want["data"].view(np.int32).astype(np.int64)
    if len(a):
        ulps = np.abs(a - b)
        assert ulps.max() <= WEIGHT_ULPS, ulps.max()
        assert (ulps == 0).mean() >= EXACT_FRAC, (ulps == 0).mean()


@pytest.mark.parametrize("n,k,d,seed", [(1, 1, 4, 0), (7, 4, 3, 1), (3000, 15, 16, 2),
                                        (20000, 15, 50, 3), (5000, 30, 20, 4)])
def test_knn_graph_matches_the_oracle(orc, n, k, d, seed):
    x = ds.mixture(n, d, 8, seed=seed)
    g = orc.neighbors_exact(x, k + 1)
    want = orc.fuzzy_connectivities(g["indptr"], g["indices"], g["data"], n)
    print(f"fuzzy n={n} k={k} seed={seed}", flush=True)
    got = P.fuzzy_connectivities(g["indptr"], g["indices"], g["data"], n)
    _same(got, want)


def _hub_graph(n: int, seed: int):
    """Each row names 6 neighbours in ascending distance: node 0 (named by every
    other row: in-degree n - 1), node 2 for even rows, node 1 for every 20th row,
    the rest random; a few rows are empty, one names a neighbour twice."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i % 997 == 5:
            rows.append([])
            continue
        nb = []
        for hub, cond in ((0, True), (2, i % 2 == 0), (1, i % 20 == 0)):
            if cond and hub != i:
                nb.append(hub)
        while len(nb) < 6:
            j = int(rng.integers(0, n))
            if j != i and j not in nb:
                nb.append(j)
        rows.append(nb)
    rows[11] = [4, 4, 5]
    ip = np.zeros(n + 1, np.uint32)
    ip[1:] = np.cumsum([len(r) for r in rows])
    ix = np.array([j for r in rows for j in r], np.uint32)
    dx = np.concatenate([np.sort(rng.uniform(0.1, 3.0, len(r))).astype(np.float32)
                         for r in rows])
    return ip, ix, dx


@pytest.mark.parametrize("n", [600, 20000])
def test_hub_rows_take_every_merge_path(orc, n):
    ip, ix, dx = _hub_graph(n, seed=n)
    want = orc.fuzzy_connectivities(ip, ix, dx, n)
    print(f"fuzzy hubs n={n}", flush=True)
    got = P.fuzzy_connectivities(ip, ix, dx, n)
    _same(got, want)
    lens = np.diff(got["indptr"].astype(np.int64))
    if n == 20000:
        assert lens[0] > 8192 and lens[2] > 8192 and 256 < lens[1] <= 8192


def test_into_form_fills_the_callers_arrays(orc):
    x = ds.mixture(4000, 20, 8, seed=6)
    g = orc.neighbors_exact(x, 15)
    held = {}

    def alloc(n_rows, nnz):
        held["ip"] = np.zeros(n_rows + 1, np.uint32)
        held["ix"] = np.zeros(nnz, np.uint32)
        held["w"] = np.zeros(nnz, np.float32)
        return held["ip"], held["ix"], held["w"]

    print("fuzzy into", flush=True)
    nnz = P.fuzzy_connectivities_into(g["indptr"], g["indices"], g["data"], 4000, alloc)
    want = orc.fuzzy_connectivities(g["indptr"], g["indices"], g["data"], 4000)
    assert nnz == len(want["indices"])
    _same({"indptr": held["ip"], "indices": held["ix"], "data": held["w"]}, want)
    with pytest.raises(ValueError, match="alloc returned arrays of the wrong size"):
        P.fuzzy_connectivities_into(g["indptr"], g["indices"], g["data"], 4000,
                                    lambda r, e: (np.zeros(r + 1, np.uint32),
                                                  np.zeros(e + 1, np.uint32),
                                                  np.zeros(e, np.float32)))


def test_out_of_range_index_is_an_error():
    ip = np.array([0, 1, 2], np.uint32)
    with pytest.raises(ValueError, match="index"):
        P.fuzzy_connectivities(ip, np.array([1, 2], np.uint32), np.ones(2, np.float32), 2)


def test_device_form_on_the_nndescent_graph(orc):
    import torch
    x = ds.mixture(20000, 50, 8, seed=7)
    xt = torch.from_numpy(x).cuda()
    print("fuzzy device", flush=True)
    ip, ix, dx, nnz = P.nndescent_device(xt, 15, seed=7)
    cip, cix, cw, cnnz = P.fuzzy_connectivities_device(ip, ix, dx, 20000)
    torch.cuda.synchronize()
    h_ip = ip.cpu().numpy().view(np.uint32)
    want = orc.fuzzy_connectivities(h_ip, ix.cpu().numpy().view(np.uint32)[:nnz],
                                    dx.cpu().numpy()[:nnz], 20000)
    got = {"indptr": cip.cpu().numpy().view(np.uint32),
           "indices": cix.cpu().numpy().view(np.uint32)[:cnnz],
           "data": cw.cpu().numpy()[:cnnz]}
    _same(got, want)


def test_invariants_at_a_million_rows():
    """umap.rs:1954-2000 on a 1M-row graph (random neighbours, so in-degrees
    vary): the result equals its own transpose bit for bit, rows are sorted
    and free of self-loops."""
    import torch
    sp = pytest.importorskip("scipy.sparse")
    n, k = 1_000_000, 15
    rng = np.random.default_rng(11)
    ix = rng.integers(0, n - 1, (n, k), dtype=np.int64)
    ix += ix >= np.arange(n)[:, None]  # no self
    ix = np.sort(ix, axis=1).astype(np.uint32).ravel()
    dx = np.sort(rng.uniform(0.5, 4.0, (n, k)).astype(np.float32), axis=1).ravel()
    ip = (np.arange(n + 1, dtype=np.int64) * k).astype(np.uint32)
    print("fuzzy 1M", flush=True)
    cip, cix, cw, nnz = P.fuzzy_connectivities_device(
        torch.from_numpy(ip.view(np.int32)).cuda(), torch.from_numpy(ix.view(np.int32)).cuda(),
        torch.from_numpy(dx).cuda(), n)
    torch.cuda.synchronize()
    h_ip = cip.cpu().numpy().view(np.uint32).astype(np.int64)
    h_ix = cix.cpu().numpy().view(np.uint32)[:nnz].astype(np.int64)
    h_w = cw.cpu().numpy()[:nnz]
    assert h_ip[0] == 0 and h_ip[-1] == nnz
    rows = np.repeat(np.arange(n), np.diff(h_ip))
    assert not np.any(h_ix == rows)
    same_row = rows[1:] == rows[:-1]
    assert np.all(h_ix[1:][same_row] > h_ix[:-1][same_row])
    m = sp.csr_matrix((h_w.view(np.uint32).astype(np.int64) + 1, h_ix, h_ip), shape=(n, n))
    t = m.T.tocsr()
    t.sort_indices()
    assert np.array_equal(t.indptr, m.indptr) and np.array_equal(t.indices, m.indices)
    assert np.array_equal(t.data, m.data)


@pytest.mark.timeout(600)
def test_more_rows_than_the_narrow_partition_holds(orc):
    """A circulant kNN graph over 2^26 + 4096 rows (row i -> i+1 .. i+8 mod n,
    the same distances in every row), past the 64 super-buckets of 2^20 rows
    of the narrow transpose: every row's fuzzy union is row i +- s with the
    weight of offset s, which the oracle gives on a 64-row circulant graph."""
    import torch
    n, k = (1 << 26) + 4096, 8
    dist = (0.5 + 0.25 * np.arange(1, k + 1)).astype(np.float32)

    def circulant(nn):
        off = np.arange(1, k + 1)
        return ((np.arange(nn)[:, None] + off[None, :]) % nn).astype(np.uint32)

    ns = 64
    small = orc.fuzzy_connectivities(np.arange(ns + 1, dtype=np.uint32) * k, circulant(ns).ravel(),
                                     np.tile(dist, ns), ns)
    r = 32
    sl = slice(int(small["indptr"][r]), int(small["indptr"][r + 1]))
    want = {int(j) - r: w for j, w in zip(small["indices"][sl], small["data"][sl])}
    assert sorted(want) == [s for s in range(-k, k + 1) if s != 0]
    print(f"fuzzy large n={n}", flush=True)
    dev = torch.device("cuda:0")
    ar = torch.arange(n, dtype=torch.int64, device=dev)
    ix = ((ar[:, None] + torch.arange(1, k + 1, device=dev)[None, :]) % n).to(torch.int32).reshape(-1)
    ip = (torch.arange(n + 1, dtype=torch.int64, device=dev) * k).to(torch.int32)
    dx = torch.from_numpy(dist).to(dev).repeat(n)
    del ar
    cip, cix, cw, nnz = P.fuzzy_connectivities_device(ip, ix, dx, n)
    torch.cuda.synchronize()
    assert nnz == 2 * k * n
    rng = np.random.default_rng(3)
    rows = np.concatenate([np.arange(0, 1024), np.arange(n - 2048, n), rng.integers(0, n, 2048)])
    h_ip = cip.cpu().numpy().view(np.uint32).astype(np.int64)
    assert np.array_equal(h_ip[rows], rows * 2 * k)
    sel = (torch.from_numpy(rows).to(dev)[:, None] * (2 * k) + torch.arange(2 * k, device=dev)).reshape(-1)
    gi = cix[sel].cpu().numpy().view(np.uint32).astype(np.int64).reshape(len(rows), 2 * k)
    gw = cw[sel].cpu().numpy().reshape(len(rows), 2 * k)
    for t, i in enumerate(rows):
        assert (np.diff(gi[t]) > 0).all()
        for j, w in zip(gi[t], gw[t]):
            o = (j - i) % n
            o = o - n if o > n // 2 else o
            assert o in want and w.view(np.uint32) == want[o].view(np.uint32), (i, j, o)

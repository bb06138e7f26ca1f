"""GPU checks on the full-size workloads of the scope table's configurations.

* C2 (1.3M x 50, 60-component mixture, k = 15): the exact tiled kNN of a
  100,000-cell random subset, taken as its own dataset, equals the FP64 oracle
  on every row (indices bit-exact, distances within 1e-5 relative); the plain
  GPU NN-Descent's recall@15, over the same 100,000 cells as queries against the
  exact top-15 over all 1.3M cells, is not below the reference restatement's on
  the same input (paired over the queries).
* C3 (4.06M x 50, 150 components, 5% uniform outliers in a box 3x the data
  extent, k = 30, BF16 screen + FP32 re-rank of each row's top 60): every output
  distance is the exact f32 distance of its pair (FP64 relative error at the
  f32 rounding level on a 50,000-cell subset), rows are valid, and recall@30
  over 20,000 queries reaches 0.95.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def test_c2_exact_subset_and_recall_parity(orc, not_below):
    import torch
    n = 1_300_000
    x = ds.mixture(n, 50, 60, seed=0)
    sub = np.sort(np.random.default_rng(2).choice(n, 100_000, replace=False))
    # the subset as its own dataset: bit-exact against the FP64 oracle, every row
    xs = np.ascontiguousarray(x[sub])
    print("C2 exact subset", flush=True)
    idx, dist = P.knn_exact(xs, 15)
    ei, ed = orc.knn_exact64(xs, 15)
    assert np.array_equal(idx, ei), "exact kNN indices differ from the FP64 oracle"
    assert np.allclose(dist, np.sqrt(ed), rtol=1e-5, atol=0)
    # recall of the plain form and of the reference restatement, same queries
    print("C2 recall", flush=True)
    xt = torch.from_numpy(x).cuda()
    qt = torch.from_numpy(sub.astype(np.int32)).cuda()
    eq, _ = P.knn_exact_device(xt, 15, qt)
    eq = eq.cpu().numpy()
    ip, ix, dx, nnz = P.nndescent_device(xt, 15, seed=0)
    assert nnz == n * 15
    rg = ds.recall(ix[: n * 15].view(n, 15)[qt.long()].cpu().numpy(), eq)
    gp, _ = orc.nndescent_port(x, 15, seed=0)
    rp = ds.recall(gp["indices"].reshape(n, 15)[sub], eq)
    print(f"C2 recall@15 GPU {rg.mean():.5f} reference restatement {rp.mean():.5f}", flush=True)
    not_below(rg - rp)


def test_c3_precision_form_distances_and_recall():
    import torch
    n, k = 4_060_000, 30
    x = ds.mixture_with_outliers(n, 50, 150, seed=0, outlier_frac=0.05)
    xt = torch.from_numpy(x).cuda()
    print("C3 BF16 form", flush=True)
    ip, ix, dx, nnz = P.nndescent_device(xt, k, seed=0, precision="bf16")
    st = P.last_stats()
    assert st["heap_k"] == 2 * k and nnz == n * k
    ip = ip.cpu().numpy().view(np.uint32)
    assert (np.diff(ip) == k).all()
    rng = np.random.default_rng(777)
    sub = np.sort(rng.choice(n, 50_000, replace=False))
    subt = torch.from_numpy(sub).cuda().long()
    si = ix[: n * k].view(n, k)[subt].cpu().numpy().view(np.uint32).astype(np.int64)
    sd = dx[: n * k].view(n, k)[subt].cpu().numpy()
    # rows: no self, no duplicates, ascending
    assert not (si == sub[:, None]).any()
    srt = np.sort(si, axis=1)
    assert not (srt[:, 1:] == srt[:, :-1]).any()
    assert (np.diff(sd, axis=1) >= 0).all()
    # every distance is its pair's f32 distance: FP64 error at the f32 rounding level
    diff = x[si].astype(np.float64) - x[sub][:, None, :].astype(np.float64)
    d64 = np.sqrt((diff * diff).sum(-1))
    ok = d64 > 0
    rel = np.abs(sd[ok].astype(np.float64) - d64[ok]) / d64[ok]
    print(f"C3 output distance error vs FP64: max {rel.max():.2e} p99 {np.percentile(rel, 99):.2e}",
          flush=True)
    assert rel.max() < 1e-6
    # recall@30 against the exact top-30, 20k queries
    q = np.sort(rng.choice(n, 20_000, replace=False)).astype(np.int32)
    qt = torch.from_numpy(q).cuda()
    eq, _ = P.knn_exact_device(xt, k, qt)
    rec = ds.recall(ix[: n * k].view(n, k)[qt.long()].cpu().numpy(), eq.cpu().numpy())
    print(f"C3 recall@30 {rec.mean():.4f}", flush=True)
    assert rec.mean() >= 0.95

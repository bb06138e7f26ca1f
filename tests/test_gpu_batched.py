"""GPU checks of the batched form: per-sample exact kNN of many uneven samples
(cells plus simulated doublets) in one persistent launch per pass, tile shapes
from the autotuner. Each sample is bit-exact against the FP64 oracle run on that
sample alone (Scrublet's per-sample kNN, crates/polariseq-core/src/
doublets.rs:1085-1113)."""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def test_batched_samples_bit_exact_and_tuned_once(orc):
    samples = ds.batched_samples(n_samples=6, seed=3, lo=200, hi=3000, mu_log=6.5, sigma_log=0.9)
    pts = [p for p, _ in samples]
    off = np.concatenate([[0], np.cumsum([len(p) for p in pts])]).astype(np.uint64)
    x = np.ascontiguousarray(np.concatenate(pts, axis=0))
    idx, dist = P.knn_exact_batched(x, off, 15)
    rep = P.batched_report()
    assert rep["configs_tried"] >= 1 and rep["units"] > 0
    for s in range(len(pts)):
        a, b = int(off[s]), int(off[s + 1])
        ei, ed = orc.knn_exact64(pts[s], 15)
        assert np.array_equal(idx[a:b] - a, ei), f"sample {s}"
        assert np.allclose(dist[a:b], np.sqrt(ed), rtol=1e-5, atol=0)
    # the same shape class reuses the tuned tile and gives the same graph
    idx2, _ = P.knn_exact_batched(x, off, 15)
    assert np.array_equal(idx, idx2)
    assert P.batched_report()["tuned_now"] == 0


def test_batched_scrublet_width(orc):
    """d = 30, Scrublet's default number of principal components, k = 20."""
    samples = ds.batched_samples(n_samples=4, d=30, seed=8, lo=300, hi=2500, mu_log=6.5,
                                 sigma_log=0.8)
    pts = [p for p, _ in samples]
    off = np.concatenate([[0], np.cumsum([len(p) for p in pts])]).astype(np.uint64)
    x = np.ascontiguousarray(np.concatenate(pts, axis=0))
    idx, dist = P.knn_exact_batched(x, off, 20)
    for s in range(len(pts)):
        a, b = int(off[s]), int(off[s + 1])
        ei, ed = orc.knn_exact64(pts[s], 20)
        assert np.array_equal(idx[a:b] - a, ei), f"sample {s}"
        assert np.allclose(dist[a:b], np.sqrt(ed), rtol=1e-5, atol=0)


def test_batched_single_tiny_samples_and_generic_dims(orc):
    rng = np.random.default_rng(1)
    sizes = [5, 17, 1, 40, 3]
    pts = [rng.random((m, 7), dtype=np.float32) for m in sizes if m > 0]
    off = np.concatenate([[0], np.cumsum([len(p) for p in pts])]).astype(np.uint64)
    x = np.ascontiguousarray(np.concatenate(pts, axis=0))
    idx, dist = P.knn_exact_batched(x, off, 4)
    for s in range(len(pts)):
        a, b = int(off[s]), int(off[s + 1])
        ei, ed = orc.knn_exact64(pts[s], 4)
        got = idx[a:b].astype(np.int64) - a
        got[idx[a:b] == 0xFFFFFFFF] = 0xFFFFFFFF
        assert np.array_equal(got, ei.astype(np.int64)), f"sample {s}"

"""Oracle (test infrastructure only) for the conventions of pp.neighbors around the
graph step, restated from python/polariseq/__init__.py:2514-2600.

These lines are numpy in the reference itself, so the restatement is the same
numpy expressions; what pins them is numpy's own arithmetic (float32 pairwise
summation for the row norm, IEEE division, float64 squaring), checked
element for element against the GPU kernels. ``neighbors_graph`` restates the
control flow (k = n_neighbors - 1, the exact/approximate switch at 10,000
cells, the cosine handling) on top of callables for the two searches and the
fuzzy step, so a test can run it with the oracle's searches and compare with
the library's GPU pipeline.
"""

from __future__ import annotations

import numpy as np


def unit_rows(emb: np.ndarray) -> np.ndarray:
    """__init__.py:2540-2543: rows scaled to unit length (float32)."""
    norms = np.linalg.norm(emb, axis=1, keepdims=True)
    return np.ascontiguousarray(emb / np.maximum(norms, 1e-12), dtype=np.float32)


def cosine_from_unit_euclidean(d: np.ndarray) -> np.ndarray:
    """__init__.py:2551-2552: |a - b| on unit rows -> 1 - cos = |a - b|^2 / 2."""
    return (np.asarray(d, np.float32).astype(np.float64) ** 2 / 2.0).astype(np.float32)


def neighbors_graph(emb: np.ndarray, n_neighbors: int, metric: str, approximate, seed: int,
                    nndescent, exact, fuzzy) -> dict:
    """__init__.py:2514-2600 without AnnData. nndescent(search, k, seed) returns
    (indptr, indices, data) of the approximate graph; exact(emb, n_neighbors,
    metric) the exact one; fuzzy(indptr, indices, data, n_obs) the
    connectivities dict."""
    emb = np.ascontiguousarray(emb, dtype=np.float32)
    n_obs = emb.shape[0]
    k = max(1, n_neighbors - 1)
    use_approx = approximate if approximate is not None else (n_obs > 10_000)
    if metric not in ("euclidean", "cosine"):
        raise ValueError(f"metric must be 'euclidean' or 'cosine', not {metric!r}")
    if use_approx:
        search = unit_rows(emb) if metric == "cosine" else emb
        ip, ix, dx = nndescent(search, k, seed)
        if metric == "cosine":
            dx = cosine_from_unit_euclidean(dx)
        method = "nndescent"
    else:
        ip, ix, dx = exact(emb, n_neighbors, metric)
        method = "exact"
    conn = fuzzy(ip, ix, dx, n_obs)
    return {"distances": {"indptr": ip, "indices": ix, "data": dx}, "connectivities": conn,
            "method": method}

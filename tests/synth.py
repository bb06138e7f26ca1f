"""Synthetic inputs shared by the GPU tests (own recipes; the tests compare the
GPU with float64 references computed on the same arrays, so any matrix of the
right character serves).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg
import scipy.sparse as sp


def low_rank_counts(n: int = 4000, m: int = 1500, rank: int = 6, seed: int = 0,
                    density: float = 0.3) -> sp.csr_matrix:
    """n x m non-negative float32 CSR: `rank` well-separated directions
    (scales 600 * 0.65^j) plus N(0, 0.25) noise, absolute values kept on a
    random `density` share of the entries (the shape of a normalised count
    matrix with a few programs)."""
    rng = np.random.default_rng(seed)
    scale = 600.0 * 0.65 ** np.arange(rank)
    cells = rng.standard_normal((n, rank)) * scale
    genes = rng.standard_normal((rank, m))
    keep = rng.random((n, m)) < density
    x = cells @ genes + 0.5 * rng.standard_normal((n, m))
    return sp.csr_matrix(np.where(keep, np.abs(x), 0.0).astype(np.float32))


def subspace_angle_deg(a: np.ndarray, b: np.ndarray) -> float:
    """Largest principal angle, in degrees, between the column spans of a and b."""
    return float(np.degrees(scipy.linalg.subspace_angles(a, b)).max())


def normalize_log1p(x: sp.csr_matrix) -> sp.csr_matrix:
    """Per-cell total scaled to 1e4, then log1p (pp.normalize_total + pp.log1p),
    on a float64 copy, returned as float32 CSR."""
    x = x.tocsr(copy=True).astype(np.float64)
    tot = np.asarray(x.sum(axis=1)).ravel()
    f = np.zeros_like(tot)
    f[tot > 0] = 1e4 / tot[tot > 0]
    x = sp.csr_matrix(sp.diags(f) @ x)
    x.data = np.log1p(x.data)
    return sp.csr_matrix(x.astype(np.float32))

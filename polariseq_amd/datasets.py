"""Synthetic single-cell PCA embeddings: Gaussian mixtures shaped like a PCA spectrum.

No dataset can be fetched in this environment, so every workload is generated here
with ``numpy.random.default_rng(seed)``. The draw order is fixed and written out
below, and the FP32 cast comes last.

Common mixture with C components in D dimensions:

1. weights ``w ~ Dirichlet(1_C)``;
2. centres ``mu_c ~ N(0, 4^2 I_D)``, drawn as a (C, D) array;
3. labels ``c_i ~ Categorical(w)``, drawn with ``rng.choice(C, size=n, p=w)``;
4. noise ``z_i ~ N(0, I_D)``, drawn as an (n, D) array;
5. ``x_i = mu_{c_i} + sigma * z_i``, where ``sigma_j = 1/sqrt(j+1)`` for j = 0..D-1.

The sigma decay mimics the decaying variance of principal components.
"""

from __future__ import annotations

import numpy as np


def sigma(d: int) -> np.ndarray:
    return 1.0 / np.sqrt(np.arange(d, dtype=np.float64) + 1.0)


def mixture(n: int, d: int = 50, components: int = 20, seed: int = 0,
            shift: np.ndarray | None = None, rng: np.random.Generator | None = None,
            chunk: int = 1 << 20) -> np.ndarray:
    """n x d float32 cells from the common mixture (module docstring).

    Large n is generated in row chunks (same draw order as one call per chunk),
    so a 4M x 50 embedding never holds more than one float64 chunk at a time.
    """
    rng = rng if rng is not None else np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(components))
    mu = rng.normal(0.0, 4.0, size=(components, d))
    if shift is not None:
        mu = mu + shift[None, :]
    labels = rng.choice(components, size=n, p=w)
    s = sigma(d)
    out = np.empty((n, d), np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        z = rng.normal(size=(b - a, d))
        out[a:b] = (mu[labels[a:b]] + s * z).astype(np.float32)
    return out


def mixture_with_outliers(n: int, d: int = 50, components: int = 150, seed: int = 0,
                          outlier_frac: float = 0.05) -> np.ndarray:
    """4.06M-cell precision workload: mixture on (1 - f) of the cells, then f·n
    uniform outliers in a box centred on the data with 3x its per-dimension
    extent (max - min over the mixture cells). Outliers are appended last."""
    rng = np.random.default_rng(seed)
    n_out = int(round(outlier_frac * n))
    base = mixture(n - n_out, d, components, rng=rng)
    lo = base.min(axis=0).astype(np.float64)
    hi = base.max(axis=0).astype(np.float64)
    centre = 0.5 * (lo + hi)
    half = 1.5 * (hi - lo)
    u = rng.uniform(-1.0, 1.0, size=(n_out, d))
    outl = (centre + half * u).astype(np.float32)
    return np.concatenate([base, outl], axis=0)


def batched_samples(n_samples: int = 64, d: int = 50, components: int = 20, seed: int = 0,
                    lo: int = 2000, hi: int = 200000, doublet_factor: int = 2,
                    mu_log: float = 10.0, sigma_log: float = 1.2):
    """Per-sample kNN workload: sizes n_s = clip(round(exp(N(mu, sigma^2))), lo, hi).

    Each sample is an independent mixture with a shift ``delta_s ~ N(0, I_D)``,
    followed by ``doublet_factor * n_s`` simulated doublets. A doublet is the mean
    of a random pair of distinct cells of the sample. Returns the list of
    (points f32[(1 + doublet_factor) n_s, d], n_cells) in sample order.
    """
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(np.exp(rng.normal(mu_log, sigma_log, size=n_samples))), lo, hi)
    out = []
    for s in sizes.astype(np.int64):
        shift = rng.normal(size=d)
        cells = mixture(int(s), d, components, shift=shift, rng=rng)
        nd = doublet_factor * int(s)
        a = rng.integers(0, s, size=nd)
        b = rng.integers(0, s - 1, size=nd)
        b = b + (b >= a)  # distinct partner
        dbl = (0.5 * (cells[a].astype(np.float64) + cells[b].astype(np.float64))).astype(np.float32)
        out.append((np.concatenate([cells, dbl], axis=0), int(s)))
    return out


def blobs_20d(n: int = 6000, d: int = 20, seed: int = 9) -> np.ndarray:
    """The reference's structured-recall shape (nndescent.rs:1110-1120): 12 groups,
    coordinate j of group c is c*3*(j mod 3) + 4u with u ~ U[0,1). The
    reference draws u with ChaCha8, which is absent here; numpy draws it instead."""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) % 12).astype(np.float32)[:, None]
    j = (np.arange(d) % 3).astype(np.float32)[None, :]
    u = rng.random((n, d), dtype=np.float32)
    return (c * 3.0 * j + u * 4.0).astype(np.float32)


def clustered(n: int, d: int, n_clusters: int, seed: int) -> np.ndarray:
    """The reference's test generator shape (nndescent.rs:1012-1022): cluster i mod C
    at offset 10*C plus U[0,1) noise. numpy draws the noise (ChaCha8 is absent)."""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) % n_clusters).astype(np.float32)[:, None]
    return (c * 10.0 + rng.random((n, d), dtype=np.float32)).astype(np.float32)


def recall(approx_indices: np.ndarray, exact_indices: np.ndarray) -> np.ndarray:
    """Per-row |approx ∩ exact| / k (pyapi/tests/test_neighbors.py:26-29)."""
    a = np.asarray(approx_indices)
    e = np.asarray(exact_indices)
    k = e.shape[1]
    hits = np.zeros(a.shape[0], np.int64)
    for j in range(a.shape[1]):
        hits += (a[:, j:j + 1] == e).any(axis=1)
    return hits / k

"""Oracle (test infrastructure only) for the exact Gram PCA of a CSR matrix,
restated from crates/polariseq-core/src/pca.rs:396-535 and rsvd.rs:155-177,
241-283, 356-390, 439-456 (the route ``pca_sparse`` / ``pca_sparse_scaled``
take for at most 4,096 genes), in numpy for small matrices.

What pins it (tests/test_pca_oracle.py): the reference's own tests, restated
with numpy's generator in place of ChaCha8 (pca.rs:708-905): an exact thin SVD
of the densified, centred (and for the scaled form, densely standardised)
matrix, which is a separate textbook route to the same components, within the
reference's tolerances and tighter ones; the rank-two fixture; the sign rule
brute-forced on small columns with ties; the float32 summation order of the
projection on a case where any other order gives a different answer. The
eigendecomposition itself is LAPACK's (numpy.linalg.eigh), in place of the
reference's faer solver.
"""

from __future__ import annotations

import numpy as np


def dense_f64(indptr, indices, data, n: int, m: int) -> np.ndarray:
    """The CSR matrix as dense float64 (stored values of one position added)."""
    out = np.zeros((n, m), dtype=np.float64)
    rows = np.repeat(np.arange(n), np.diff(np.asarray(indptr, dtype=np.int64)))
    np.add.at(out, (rows, np.asarray(indices, dtype=np.int64)), np.asarray(data, dtype=np.float32))
    return out


def gene_stats(indptr, indices, data, n: int, m: int):
    """preprocess.rs:314-358: per-gene mean and ddof-1 variance in float64,
    var = max((q/n - mean^2) n / (n - 1), 0), 0 for fewer than 2 rows."""
    d = dense_f64(indptr, indices, data, n, m)
    s = d.sum(axis=0)
    q = (d * d).sum(axis=0)
    mean = s / n
    if n < 2:
        var = np.zeros(m)
    else:
        var = np.maximum((q / n - mean * mean) * n / (n - 1.0), 0.0)
    return mean, var


def scaled_values(indptr, indices, data, n: int, m: int, max_value=None) -> np.ndarray:
    """pca.rs:403-414, 450-460: stored value x of gene j -> float32(clip((x - mu_j)
    / sd_j) - clip(-mu_j / sd_j)), sd 0 -> 1, clip = max_value or none."""
    mean, var = gene_stats(indptr, indices, data, n, m)
    sd = np.sqrt(var)
    inv = np.where(sd == 0.0, 1.0, 1.0 / np.where(sd == 0.0, 1.0, sd))
    clip = np.inf if max_value is None else float(max_value)
    zero = np.clip(-mean * inv, -clip, clip)
    j = np.asarray(indices[: int(indptr[-1])], dtype=np.int64)
    x = np.asarray(data[: int(indptr[-1])], dtype=np.float32).astype(np.float64)
    z = np.clip((x - mean[j]) * inv[j], -clip, clip)
    return (z - zero[j]).astype(np.float32)


def column_means(indptr, indices, data, n: int, m: int) -> np.ndarray:
    """rsvd.rs:155-177: float64 column sums / n, as float32."""
    return (dense_f64(indptr, indices, data, n, m).sum(axis=0) / max(n, 1)).astype(np.float32)


def centred_gram(indptr, indices, data, n: int, m: int, center: bool):
    """pca.rs:494-510: C = S^T S (float64), C_ij -= (mu_i n) mu_j, read from the
    lower triangle (i >= j) as the reference's solver does; trace."""
    d = dense_f64(indptr, indices, data, n, m)
    c = d.T @ d
    mu = None
    if center:
        mu = column_means(indptr, indices, data, n, m)
        mu64 = mu.astype(np.float64)
        c = c - (mu64[:, None] * float(n)) * mu64[None, :]
    lower = np.tril(c)
    c = lower + np.tril(lower, -1).T
    return c, float(np.trace(c)), mu


def top_components(c: np.ndarray, k: int):
    """pca.rs:512-524: eigenvalues descending and V[i][j] = eigenvector j as float32."""
    vals, vecs = np.linalg.eigh(c)
    m = c.shape[0]
    order = [m - 1 - j for j in range(k)]
    return vals[order], np.ascontiguousarray(vecs[:, order].astype(np.float32))


def mu_b(mu: np.ndarray, v: np.ndarray) -> np.ndarray:
    """rsvd.rs:252-262: sum_j mu_j V[j] in float64, j in order, as float32."""
    acc = np.zeros(v.shape[1], dtype=np.float64)
    for j in range(v.shape[0]):
        acc = acc + float(mu[j]) * v[j].astype(np.float64)
    return acc.astype(np.float32)


def project(indptr, indices, values, n: int, v: np.ndarray, mub=None) -> np.ndarray:
    """rsvd.rs:265-280: out[i] = sum over row i's stored entries, in order, of
    float32(value * V[j]) added in float32; minus mu^T V. Vectorised over rows
    by entry position, which keeps each row's order."""
    k = v.shape[1]
    ip = np.asarray(indptr, dtype=np.int64)
    lens = np.diff(ip)
    acc = np.zeros((n, k), dtype=np.float32)
    vals = np.asarray(values, dtype=np.float32)
    idx = np.asarray(indices, dtype=np.int64)
    for p in range(int(lens.max()) if n else 0):
        rows = np.nonzero(lens > p)[0]
        e = ip[rows] + p
        acc[rows] = acc[rows] + vals[e][:, None] * v[idx[e]]
    if mub is not None:
        acc = acc - mub[None, :]
    return acc


def canonicalize_signs(u: np.ndarray) -> np.ndarray:
    """rsvd.rs:439-456: flip each column whose first largest-magnitude entry
    (over magnitudes > 0; row 0 if none) is negative."""
    u = u.copy()
    for j in range(u.shape[1]):
        a = np.abs(u[:, j])
        i = int(np.argmax(a)) if a.size and a.max() > 0 else 0
        if u[i, j] < 0:
            u[:, j] = -u[:, j]
    return u


def pca_gram(indptr, indices, data, n: int, m: int, k: int, center: bool = True,
             components: np.ndarray | None = None) -> dict:
    """pca.rs:487-535. ``components`` replaces the eigenvectors (float32 m x k),
    so that the projection, signs and centring can be checked on exactly the
    axes another implementation found."""
    c, tv, mu = centred_gram(indptr, indices, data, n, m, center)
    vals, v = top_components(c, k)
    if components is not None:
        v = np.ascontiguousarray(components, dtype=np.float32)
    mub = mu_b(mu, v) if center else None
    emb = canonicalize_signs(project(indptr, indices, data, n, v, mub))
    ratio = (np.maximum(vals, 0.0) / tv).astype(np.float32)
    return {"embedding": emb, "variance_ratio": ratio, "components": v, "eigenvalues": vals,
            "trace": tv}


def pca_scaled(indptr, indices, data, n: int, m: int, k: int, max_value=None,
               components: np.ndarray | None = None) -> dict:
    """pca.rs:396-431: Gram PCA of the scaled view, always centred."""
    s = scaled_values(indptr, indices, data, n, m, max_value)
    return pca_gram(indptr, indices, s, n, m, k, True, components)

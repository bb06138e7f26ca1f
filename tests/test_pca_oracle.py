"""CPU checks of the PCA oracle (oracle/pca.py) and of the PCA entry points'
argument checks (which return before any device call).

The oracle restates the reference's exact Gram PCA (crates/polariseq-core/src/pca.rs:487-535).
It is pinned here the way the reference pins its own Gram route
(pca.rs:708-905), with numpy's generator in place of ChaCha8:

* against an exact thin SVD of the densified, centred matrix, a separate
  textbook route to the same components (gram_pca_equals_exact_thin_svd);
* the scaled form against a dense standardise-then-SVD
  (scaled_pca_matches_dense_scale_then_pca), clipping at none, 3 and 10, with a
  nearly constant gene, an all-zero gene and outliers;
* the rank-two fixture (dense_pca_recovers_low_rank_structure,
  sparse_pca_matches_dense_on_same_data);
* the sign rule by its defining property on columns with ties;
* the projection's float32 order on a case every other order gets wrong;
* planted mistakes (no centring term, the last maximum instead of the first,
  a float64 projection) fail these checks.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pca as P
from polariseq_amd import _native as N
from polariseq_amd import pca as gp


def csr(dense):
    m = sp.csr_matrix(np.asarray(dense, dtype=np.float32))
    m.sort_indices()
    return m.indptr.astype(np.uint64), m.indices.astype(np.uint32), m.data.astype(np.float32)


def svd_pca(x, k, center=True):
    """Thin SVD of the dense matrix (float64): embedding U S and S^2 / total."""
    x = np.asarray(x, dtype=np.float64)
    xc = x - x.mean(axis=0) if center else x
    u, s, _ = np.linalg.svd(xc, full_matrices=False)
    total = float((s ** 2).sum())
    return u[:, :k] * s[:k], (s[:k] ** 2) / total


def cos_and_scale(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = []
    for j in range(a.shape[1]):
        na, nb = np.linalg.norm(a[:, j]), np.linalg.norm(b[:, j])
        out.append((abs(a[:, j] @ b[:, j]) / (na * nb), na / nb))
    return out


def counts_600x40(seed=3):
    rng = np.random.default_rng(seed)
    n, m = 600, 40
    keep = rng.random((n, m)) < 0.2
    vals = np.floor(rng.random((n, m)) * 6.0) + 1.0
    return np.where(keep, vals, 0.0).astype(np.float32)


def scaled_fixture(seed=11):
    rng = np.random.default_rng(seed)
    n, m = 500, 30
    dense = np.zeros((n, m), np.float32)
    keep = rng.random((n, m - 2)) < 0.25
    dense[:, : m - 2] = np.where(keep, np.floor(rng.random((n, m - 2)) * 5.0) + 1.0, 0.0)
    dense[:, m - 2] = np.where(np.arange(n) % 250 == 0, 0.0, 5.0 + rng.random(n) * 0.01)
    dense[::97, 3] = 400.0
    return dense


def scale_dense_reference(dense, max_value):
    """pca.rs:780-796: centre, divide by the ddof-1 sd (0 -> 1), clip."""
    x = dense.astype(np.float64)
    mu = x.mean(axis=0)
    var = ((x - mu) ** 2).sum(axis=0) / (x.shape[0] - 1)
    sd = np.where(var == 0.0, 1.0, np.sqrt(var))
    c = math.inf if max_value is None else float(max_value)
    return np.clip((x - mu) / sd, -c, c).astype(np.float32)


def rank_two_dense():
    """pca.rs:689-706."""
    d = np.zeros((50, 5), np.float32)
    t = np.arange(50, dtype=np.float32)
    d[:, 0] = t
    d[:, 1] = t * np.float32(-2.0) + (np.arange(50) % 3).astype(np.float32) * np.float32(0.001)
    d[:, 2] = np.sin(t * np.float32(0.7)).astype(np.float32) * np.float32(1e-6)
    return d


def test_gram_equals_exact_thin_svd():
    dense = counts_600x40()
    ip, ix, dx = csr(dense)
    n, m, k = 600, 40, 6
    g = P.pca_gram(ip, ix, dx, n, m, k, True)
    emb, vr = svd_pca(dense, k)
    assert np.max(np.abs(g["variance_ratio"] - vr)) < 1e-6  # the reference asserts 1e-4
    for c, scale in cos_and_scale(g["embedding"], emb):
        assert c > 1 - 1e-9 and abs(scale - 1) < 1e-5  # the reference asserts |cos| > 0.9999
    again = P.pca_gram(ip, ix, dx, n, m, k, True)
    assert np.array_equal(g["embedding"], again["embedding"])


def test_gram_without_centring_matches_uncentred_svd():
    dense = counts_600x40(5)
    ip, ix, dx = csr(dense)
    g = P.pca_gram(ip, ix, dx, 600, 40, 5, False)
    emb, vr = svd_pca(dense, 5, center=False)
    assert np.max(np.abs(g["variance_ratio"] - vr)) < 1e-6
    assert all(c > 1 - 1e-9 for c, _ in cos_and_scale(g["embedding"], emb))


def test_planted_missing_centring_term_is_caught(monkeypatch):
    dense = counts_600x40()
    ip, ix, dx = csr(dense)
    real = P.centred_gram

    def wrong(indptr, indices, data, n, m, center):
        c, _, _ = real(indptr, indices, data, n, m, False)  # drops C -= n mu mu^T
        return c, float(np.trace(c)), P.column_means(indptr, indices, data, n, m)

    monkeypatch.setattr(P, "centred_gram", wrong)
    g = P.pca_gram(ip, ix, dx, 600, 40, 6, True)
    _, vr = svd_pca(dense, 6)
    assert np.max(np.abs(g["variance_ratio"] - vr)) > 1e-3


@pytest.mark.parametrize("max_value", [None, 3.0, 10.0])
def test_scaled_matches_dense_scale_then_pca(max_value):
    dense = scaled_fixture()
    ip, ix, dx = csr(dense)
    n, m, k = 500, 30, 5
    g = P.pca_scaled(ip, ix, dx, n, m, k, max_value)
    emb, vr = svd_pca(scale_dense_reference(dense, max_value), k)
    assert np.max(np.abs(g["variance_ratio"] - vr)) < 1e-5  # the reference asserts 1e-4
    for c, scale in cos_and_scale(g["embedding"], emb):
        assert c > 1 - 1e-7 and abs(scale - 1) < 1e-4  # reference: 0.9999 and 1e-3


def test_scaled_values_of_constant_and_zero_genes():
    dense = scaled_fixture()
    ip, ix, dx = csr(dense)
    mean, var = P.gene_stats(ip, ix, dx, 500, 30)
    x = dense.astype(np.float64)
    assert np.allclose(var, x.var(axis=0, ddof=1), rtol=1e-10, atol=1e-12)
    assert var[29] == 0.0 and mean[29] == 0.0  # the all-zero gene: sd 0 -> divided by 1
    s = P.scaled_values(ip, ix, dx, 500, 30, 3.0)
    # stored value -> clip(z) - clip(zero): rebuild the dense scaled matrix and compare
    rebuilt = np.zeros((500, 30), np.float64)
    rows = np.repeat(np.arange(500), np.diff(ip.astype(np.int64)))
    rebuilt[rows, ix.astype(np.int64)] = s
    zero = np.clip(-mean / np.where(var == 0, 1.0, np.sqrt(var)), -3.0, 3.0)
    want = scale_dense_reference(dense, 3.0).astype(np.float64)
    assert np.max(np.abs(rebuilt + zero[None, :] - want)) < 1e-5


def test_rank_two_structure():
    d = rank_two_dense()
    ip, ix, dx = csr(d)
    g = P.pca_gram(ip, ix, dx, 50, 5, 2, True)
    assert g["variance_ratio"].sum() > 0.999
    assert g["variance_ratio"][0] > g["variance_ratio"][1]
    u = P.pca_gram(ip, ix, dx, 50, 5, 2, False)
    assert u["variance_ratio"].sum() > 0.95


def test_sign_rule_property_and_ties():
    rng = np.random.default_rng(0)
    for _ in range(200):
        u = rng.integers(-2, 3, size=(7, 3)).astype(np.float32)
        v = P.canonicalize_signs(u)
        for j in range(3):
            col = v[:, j]
            assert np.array_equal(col, u[:, j]) or np.array_equal(col, -u[:, j])
            a = np.abs(col)
            if a.max() > 0:
                first = int(np.nonzero(a == a.max())[0][0])
                assert col[first] > 0
            else:
                assert np.array_equal(col, u[:, j])
    tie = np.array([[-2.0], [2.0]], np.float32)
    assert np.array_equal(P.canonicalize_signs(tie), -tie)  # the first of the tie decides

    def last_max(u):  # planted mistake: the last maximum decides
        u = u.copy()
        a = np.abs(u[:, 0])
        i = len(a) - 1 - int(np.argmax(a[::-1]))
        if u[i, 0] < 0:
            u[:, 0] = -u[:, 0]
        return u

    assert not np.array_equal(last_max(tie), P.canonicalize_signs(tie))


def test_projection_sums_each_row_in_order_in_float32():
    # row 0 holds 1e8, 1, -1e8 (genes 0, 1, 2); V = 1: left to right in float32
    # gives (1e8 + 1) - 1e8 = 0; any float64 or reordered sum gives 1
    ip = np.array([0, 3], np.uint64)
    ix = np.array([0, 1, 2], np.uint32)
    dx = np.array([1e8, 1.0, -1e8], np.float32)
    v = np.ones((3, 1), np.float32)
    assert P.project(ip, ix, dx, 1, v)[0, 0] == 0.0
    exact = float(np.float64(1e8) + 1.0 - 1e8)
    assert exact == 1.0
    # product rounded to float32 before the add: 3 * float32(1/3) then + 0
    v3 = np.full((1, 1), np.float32(1.0 / 3.0))
    out = P.project(np.array([0, 1], np.uint64), np.array([0], np.uint32), np.array([3.0], np.float32), 1, v3)
    assert out[0, 0] == np.float32(np.float32(3.0) * np.float32(1.0 / 3.0))


def test_mu_b_and_centring_equal_dense_projection():
    dense = counts_600x40(9)
    ip, ix, dx = csr(dense)
    g = P.pca_gram(ip, ix, dx, 600, 40, 4, True)
    v = g["components"].astype(np.float64)
    want = (dense.astype(np.float64) - dense.astype(np.float64).mean(axis=0)) @ v
    got = g["embedding"].astype(np.float64)
    for j in range(4):  # equal up to the column sign, to float32 rounding
        s = np.sign(got[:, j] @ want[:, j])
        assert np.max(np.abs(got[:, j] - s * want[:, j])) < 1e-4 * np.max(np.abs(want[:, j]))


# ── argument checks of the C-ABI (they return before any device call) ──


def _abi(n, m, k, scale=0, max_value=math.inf):
    ip = np.zeros(n + 1, np.uint64)
    emb = np.zeros(max(n * k, 1), np.float32)
    vr = np.zeros(max(k, 1), np.float32)
    rc = N.lib().psq_pca_csr(ip.ctypes.data, None, None, n, m, k, 1, scale, max_value, emb.ctypes.data,
                             vr.ctypes.data, None, 0)
    return rc, N.lib().psq_last_error().decode()


def test_abi_rejects_bad_shapes_before_device_work():
    rc, msg = _abi(0, 5, 1)
    assert rc == N.PSQ_E_INVALID and msg == "input matrix has zero rows (0) or columns (5)"
    rc, msg = _abi(50, 5, 10)
    assert rc == N.PSQ_E_INVALID and msg == "n_components (10) exceeds min(n_obs, n_features) (5)"
    rc, _ = _abi(50, 5, 0)
    assert rc == N.PSQ_E_INVALID
    # above 32768 genes the randomized route: its basis of (k + 10) 2 columns is capped at 1024
    rc, msg = _abi(5000, 40000, 600)
    assert rc == N.PSQ_E_UNSUPPORTED and "Krylov" in msg
    rc, _ = _abi(10, 5, 2, scale=1, max_value=-1.0)
    assert rc == N.PSQ_E_INVALID
    rc, _ = _abi(10, 5, 2, scale=1, max_value=float("nan"))
    assert rc == N.PSQ_E_INVALID


def test_python_binding_raises_like_the_reference():
    m = sp.csr_matrix(np.eye(5, dtype=np.float32))
    with pytest.raises(RuntimeError, match="PCA failed: n_components \\(6\\) exceeds"):
        gp.pca_sparse(m, 6, True, 10, 4, 0)
    with pytest.raises(ValueError):
        gp.pca_sparse(m, 2, True, 10, 4, 0, out=np.zeros((5, 3), np.float32))
    with pytest.raises(ValueError):
        gp.pca_sparse(m.tocsc(), 2, True, 10, 4, 0)

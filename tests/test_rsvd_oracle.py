"""CPU checks of the randomized-PCA oracle (oracle/rsvd.py) and of the
randomized route's argument checks in the C-ABI (which return before any
device call).

The oracle restates the reference's block Krylov SVD
(crates/polariseq-core/src/rsvd.rs:476-566). It is pinned by:

* the ChaCha block function against the published keystreams of the all-zero
  key and nonce (RFC 7539 appendix A.1 vectors 1 and 2 for 20 rounds, the
  8-round keystream of the same key);
* the sketch's defining layout (one stream, row by row);
* the route's answer against an exact SVD of the densified, centred matrix at
  the reference's tolerances (rsvd.rs:1015-1068;
  python/polariseq/tests/test_pca.py:38-62), and to rounding where the Krylov
  basis spans the whole row or column space (with the last block cut to the
  width there);
* implicit centring against the explicitly centred matrix.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import rsvd as R
from polariseq_amd import _native as N
from polariseq_amd import pca as gp


def csr_parts(x):
    m = sp.csr_matrix(np.asarray(x, dtype=np.float32))
    m.sort_indices()
    return m.indptr.astype(np.uint64), m.indices.astype(np.uint32), m.data.astype(np.float32)


def structured_counts(n=4000, m=1500, rank=6, seed=0):
    """python/polariseq/tests/test_pca.py:25-35: sparse non-negative counts with
    ``rank`` well-separated programs (numpy's generator)."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, rank))
    v = rng.normal(size=(m, rank))
    s = np.array([600, 400, 250, 150, 100, 60.0])[:rank]
    x = (u * s) @ v.T + rng.normal(size=(n, m)) * 0.5
    return np.where(rng.random((n, m)) < 0.3, np.abs(x), 0.0).astype(np.float32)


def low_rank_counts(n, m, seed):
    """Two programs of weights 20 and 8 over a noise floor (the shape of
    rsvd.rs's low_rank_csr fixture), about 40% zeros."""
    rng = np.random.default_rng(seed)
    x = 20.0 * np.outer(rng.normal(size=n), rng.normal(size=m)) + 8.0 * np.outer(rng.normal(size=n),
                                                                                rng.normal(size=m))
    x = x + rng.normal(size=(n, m))
    return np.where(rng.random((n, m)) < 0.4, 0.0, np.abs(x)).astype(np.float32)


def exact(x, k):
    """Thin SVD (float64) of the matrix minus its float32 column means."""
    mu = (np.asarray(x, np.float64).sum(axis=0) / x.shape[0]).astype(np.float32).astype(np.float64)
    xc = np.asarray(x, np.float64) - mu[None, :]
    u, s, vt = np.linalg.svd(xc, full_matrices=False)
    return u, s, vt, float((xc * xc).sum())


def angles_deg(a, b):
    qa, _ = np.linalg.qr(np.asarray(a, np.float64))
    qb, _ = np.linalg.qr(np.asarray(b, np.float64))
    sv = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return np.degrees(np.arccos(np.clip(sv, -1.0, 1.0)))


# ── the sketch ──

CHACHA20_ZERO = (
    "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
    "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586",
    "9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
    "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f",
)
CHACHA8_ZERO = ("3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e"
                "984ce172b9216f419f445367456d5619314a42a3da86b001387bfdb80e0cfe42")


def keystream_hex(blocks):
    return blocks.astype("<u4").tobytes().hex()


def test_chacha20_keystream_of_the_zero_key():
    got = R.chacha_blocks([0] * 8, [0, 1], 20)
    assert keystream_hex(got[0]) == CHACHA20_ZERO[0]
    assert keystream_hex(got[1]) == CHACHA20_ZERO[1]


def test_chacha8_keystream_of_the_zero_key():
    assert keystream_hex(R.chacha_blocks([0] * 8, [0], 8)[0]) == CHACHA8_ZERO


def test_high_counter_word_is_used():
    a = R.chacha_blocks([1, 2, 3, 4, 5, 6, 7, 8], [5, 5 + (1 << 32)], 8)
    assert not np.array_equal(a[0], a[1])


def test_sketch_is_one_stream_filled_row_by_row():
    g = R.gaussian_rowmajor(7, 5, 11)
    assert g.dtype == np.float32 and g.shape == (7, 5)
    np.testing.assert_array_equal(R.gaussian_rowmajor(35, 1, 11).reshape(7, 5), g)
    np.testing.assert_array_equal(R.gaussian_rowmajor(1, 3, 11)[0], R.gaussian_rowmajor(1, 4, 11)[0, :3])
    np.testing.assert_array_equal(R.gaussian_rowmajor(7, 5, 11), g)
    assert not np.array_equal(R.gaussian_rowmajor(7, 5, 12), g)


def test_sketch_is_standard_normal():
    g = R.gaussian_rowmajor(400, 250, 0).astype(np.float64).ravel()
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01
    # the Box-Muller pair is uncorrelated
    assert abs(np.corrcoef(g[0::2], g[1::2])[0, 1]) < 0.01


def test_seed_key_is_pcg32_of_the_seed():
    # seed 0: the first PCG32 step from state 0 leaves the increment
    st = 11634580027462260723
    xs = (((st >> 18) ^ st) >> 27) & 0xFFFFFFFF
    rot = st >> 59
    assert R.seed_key(0)[0] == ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF
    assert R.seed_key(1) != R.seed_key(0)


# ── the route ──

def test_matches_exact_svd_at_the_reference_tolerances():
    # python/polariseq/tests/test_pca.py:38-62 (k = 10, oversample 10, n_iter 4)
    x = structured_counts()
    ip, ix, dx = csr_parts(x)
    r = R.rsvd(ip, ix, dx, x.shape[0], x.shape[1], 10, 10, 4, 0, True)
    u, s, _, total = exact(x, 10)
    emb = r["embedding"].astype(np.float64)
    assert (np.abs(emb.mean(axis=0)) / emb.std(axis=0)).max() < 1e-3
    np.testing.assert_allclose(r["variance_ratio"][:6], (s[:6] ** 2 / total), rtol=5e-3)
    np.testing.assert_allclose(np.linalg.norm(emb, axis=0)[:6], s[:6], rtol=5e-3)
    for k in (1, 2, 4, 6):
        assert angles_deg(emb[:, :k], u[:, :k] * s[:k]).max() < 1.0


def test_directions_match_exact_svd():
    # rsvd.rs:1015-1068 (500 x 80, k = 4, oversample 10, n_iter 4, seed 3)
    x = low_rank_counts(500, 80, 11)
    ip, ix, dx = csr_parts(x)
    r = R.rsvd(ip, ix, dx, 500, 80, 4, 10, 4, 3, True)
    u, s, _, _ = exact(x, 4)
    sv = r["singular_values"]
    emb = r["embedding"].astype(np.float64)
    for j in range(4):
        assert abs(sv[j] - s[j]) < (2e-3 if j < 2 else 1e-2) * s[j]
        uj = emb[:, j] / sv[j]
        assert abs(np.linalg.norm(uj) - 1.0) < 1e-3
        if j < 2:
            assert abs(uj @ u[:, j]) > 0.999


@pytest.mark.parametrize("n,m,k,over,it", [
    (60, 300, 5, 10, 4),   # width = n: the basis spans the row space
    (400, 30, 3, 5, 3),    # width = m = 30 with blocks of 8: the last block cut to 6 columns
    (50, 50, 7, 3, 9),     # width = n = m
])
def test_exact_where_the_basis_spans_the_space(n, m, k, over, it):
    x = low_rank_counts(n, m, n + m)
    ip, ix, dx = csr_parts(x)
    r = R.rsvd(ip, ix, dx, n, m, k, over, it, 5, True)
    assert r["width"] == min(n, m)
    u, s, vt, total = exact(x, k)
    np.testing.assert_allclose(r["singular_values"], s[:k], rtol=1e-10)
    np.testing.assert_allclose(r["variance_ratio"], (s[:k] ** 2 / total).astype(np.float32), rtol=1e-6)
    ref = u[:, :k] * s[:k]
    emb = r["embedding"].astype(np.float64)
    for j in range(k):
        assert np.abs(np.abs(emb[:, j]) - np.abs(ref[:, j])).max() < 1e-5 * np.abs(ref[:, j]).max()
        # components: the right singular vectors, signed with the embedding
        c = r["components"][:, j].astype(np.float64)
        assert abs(abs(c @ vt[j]) - 1.0) < 1e-6
        assert np.allclose(np.asarray(x, np.float64) @ c - (np.asarray(x, np.float64) @ c).mean(),
                           emb[:, j], atol=1e-3 * np.abs(emb[:, j]).max())


def test_implicit_centring_equals_the_centred_matrix():
    x = low_rank_counts(300, 90, 4)
    ip, ix, dx = csr_parts(x)
    r = R.rsvd(ip, ix, dx, 300, 90, 4, 6, 2, 9, True)
    mu = (x.astype(np.float64).sum(axis=0) / 300).astype(np.float32).astype(np.float64)
    xc = x.astype(np.float64) - mu[None, :]
    dense = sp.csr_matrix(xc)  # every entry stored, float64 -> float32 values
    r2 = R.rsvd(dense.indptr.astype(np.uint64), dense.indices.astype(np.uint32), dense.data.astype(np.float32),
                300, 90, 4, 6, 2, 9, False)
    np.testing.assert_allclose(r2["singular_values"], r["singular_values"], rtol=1e-5)
    for j in range(2):
        a, b = r["embedding"][:, j].astype(np.float64), r2["embedding"][:, j].astype(np.float64)
        assert np.abs(a - b).max() < 1e-4 * np.abs(a).max()


def test_widths_follow_the_reference():
    assert R.widths(1000, 5000, 50, 10, 1) == (60, 120)
    assert R.widths(100, 5000, 50, 10, 4) == (60, 100)
    assert R.widths(10, 10, 2, 100, 0) == (10, 10)


# ── argument checks of the randomized route (no device call) ──

def _call(method, n=100, m=40, k=3, over=10, it=1):
    ip = np.zeros(n + 1, dtype=np.uint64)
    emb = np.zeros((n, k), dtype=np.float32)
    ratio = np.zeros(k, dtype=np.float32)
    L = N.lib()
    return L.psq_pca_csr_ex(ip.ctypes.data, None, None, n, m, k, 1, 0, math.inf, method, over, it, 0,
                            emb.ctypes.data, ratio.ctypes.data, None, 0)


def test_capi_rejects_an_unknown_method():
    assert _call(7) == N.PSQ_E_INVALID
    assert b"method" in N.lib().psq_last_error()


def test_capi_rejects_a_basis_wider_than_1024():
    assert _call(N.PSQ_PCA_RANDOMIZED, n=5000, m=5000, k=300, over=10, it=3) == N.PSQ_E_UNSUPPORTED
    assert b"Krylov" in N.lib().psq_last_error()


def test_capi_gram_route_stops_at_32768_genes():
    assert _call(N.PSQ_PCA_GRAM, m=32769) == N.PSQ_E_UNSUPPORTED


def test_python_rejects_an_unknown_method():
    x = sp.csr_matrix(np.eye(5, dtype=np.float32))
    with pytest.raises(ValueError, match="method"):
        gp.pca_sparse(x, 2, True, 10, 1, 0, method="arpack")


def test_python_rejects_negative_sketch_options():
    # pyo3 converts usize / u64 and raises OverflowError on a negative value
    x = sp.csr_matrix(np.eye(5, dtype=np.float32))
    for kw in ({"oversample": -1}, {"n_iter": -2}, {"seed": -3}):
        args = {"oversample": 10, "n_iter": 1, "seed": 0}
        args.update(kw)
        with pytest.raises(OverflowError):
            gp.pca_sparse(x, 2, True, args["oversample"], args["n_iter"], args["seed"])

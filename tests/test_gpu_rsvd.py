"""GPU parity of the randomized PCA route (polariseq_amd/csrc/pca.hip,
rsvd_route) with its oracle (oracle/rsvd.py, the reference's block Krylov SVD,
crates/polariseq-core/src/rsvd.rs:476-566, with the reference's ChaCha8
sketch, rsvd.rs:414-435).

Both sides compute in float64 from the same sketch, so they agree to rounding:
singular values and variance ratios to 1e-6 relative, every embedding and
component column to 1e-5 of its largest entry (float32 output), on the
reference's default options and deeper Krylov bases, centred and not, the
scaled view, more than 32,768 genes (where the automatic choice takes this
route), a basis cut to the width, the bare sketch (n_iter 0), one component
without oversampling, and a rank-deficient matrix. The reference's
acceptance bound against the exact SVD (python/polariseq/tests/test_pca.py:38-62)
is checked on its own structured fixture, and a run repeats bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import rsvd as R
from polariseq_amd import pca as gp

pytestmark = pytest.mark.gpu


def parts(mat):
    mat = mat.tocsr()
    mat.sort_indices()
    return mat.indptr.astype(np.uint64), mat.indices.astype(np.uint32), mat.data.astype(np.float32)


def structured(n, m, rank, density, seed):
    rng = np.random.default_rng(seed)
    load = rng.gamma(1.0, 1.0, size=(rank, m))
    score = rng.gamma(0.7, 1.0, size=(n, rank))
    x = rng.poisson(score @ load * (3.0 / rank)) * (rng.random((n, m)) < density)
    return sp.csr_matrix(x.astype(np.float32))


def sparse_structured(n, m, rank, nnz_per_row, seed):
    """Wide matrices without a dense intermediate: each row draws its genes
    from one of ``rank`` programs (a gene pool each), counts 1..5."""
    rng = np.random.default_rng(seed)
    pools = [rng.choice(m, size=m // 4, replace=False) for _ in range(rank)]
    rows, cols, vals = [], [], []
    for i in range(n):
        pool = pools[int(rng.integers(rank))]
        c = np.unique(np.concatenate([rng.choice(pool, nnz_per_row), rng.integers(0, m, nnz_per_row // 4)]))
        rows.append(np.full(c.size, i))
        cols.append(c)
        vals.append(rng.integers(1, 6, c.size).astype(np.float32))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, m))


def run(mat, k, over, it, seed, center=True, scale=False, max_value=None, method="randomized"):
    n, m = mat.shape
    comp = np.zeros((m, k), np.float32)
    if scale:
        r = gp.pca_sparse_scaled(mat, k, max_value, seed, components=comp, method=method)
    else:
        r = gp.pca_sparse(mat, k, center, over, it, seed, components=comp, method=method)
    return r["embedding"], r["variance_ratio"], comp


def check(emb, vr, comp, want, tol=1e-5):
    np.testing.assert_allclose(vr, want["variance_ratio"], rtol=1e-6, atol=1e-12)
    s = want["singular_values"]
    np.testing.assert_allclose(np.linalg.norm(emb.astype(np.float64), axis=0), s, rtol=1e-5, atol=1e-6)
    for j in range(emb.shape[1]):
        if s[j] < 1e-6 * s[0]:
            continue  # a direction of the null space: any unit vector
        for got, ref in ((emb[:, j], want["embedding"][:, j]), (comp[:, j], want["components"][:, j])):
            scale = max(float(np.abs(ref).max()), 1e-30)
            assert np.abs(got.astype(np.float64) - ref).max() <= tol * scale, j


@pytest.mark.parametrize("n,m,k,over,it,seed,center", [
    (3000, 800, 10, 10, 1, 0, True),     # the reference's default options
    (3000, 800, 10, 10, 4, 0, True),     # pp.pca's n_iter
    (2500, 600, 20, 10, 2, 7, False),
    (1200, 500, 5, 3, 3, 42, True),
])
def test_randomized_matches_oracle(n, m, k, over, it, seed, center):
    print(f"rsvd n={n} m={m} k={k} over={over} it={it} seed={seed} center={center}", flush=True)
    mat = structured(n, m, 8, 0.2, seed + 1)
    ip, ix, dx = parts(mat)
    emb, vr, comp = run(mat, k, over, it, seed, center)
    check(emb, vr, comp, R.rsvd(ip, ix, dx, n, m, k, over, it, seed, center))
    rep = gp.pca_last_report()
    assert rep["method"] == "randomized" and rep["block"] == min((k + over) * (it + 1), n, m)


@pytest.mark.parametrize("k,over,it", [(1, 0, 0), (8, 5, 0)])
def test_bare_sketch_and_one_component(k, over, it):
    # n_iter 0: the basis is the sketch's block alone; k 1, oversample 0: one column
    print(f"rsvd k={k} over={over} it={it}", flush=True)
    mat = structured(1500, 700, 6, 0.2, 31)
    ip, ix, dx = parts(mat)
    emb, vr, comp = run(mat, k, over, it, 9)
    check(emb, vr, comp, R.rsvd(ip, ix, dx, 1500, 700, k, over, it, 9, True))


def test_scaled_view_above_32768_genes():
    print("rsvd scaled m=33000 auto", flush=True)
    mat = sparse_structured(1200, 33000, 5, 150, 4)
    ip, ix, dx = parts(mat)
    emb, vr, comp = run(mat, 6, 10, 1, 2, scale=True, max_value=10.0, method="auto")
    assert gp.pca_last_report()["method"] == "randomized"
    check(emb, vr, comp, R.rsvd_scaled(ip, ix, dx, 1200, 33000, 6, 10.0, 2))


def test_scaled_view():
    print("rsvd scaled", flush=True)
    mat = structured(2000, 400, 6, 0.25, 3)
    ip, ix, dx = parts(mat)
    emb, vr, comp = run(mat, 8, 10, 1, 5, scale=True, max_value=10.0)
    check(emb, vr, comp, R.rsvd_scaled(ip, ix, dx, 2000, 400, 8, 10.0, 5))


def test_more_than_32768_genes_takes_the_randomized_route():
    print("rsvd m=33000 auto", flush=True)
    mat = sparse_structured(1500, 33000, 6, 160, 8)
    ip, ix, dx = parts(mat)
    emb, vr, comp = run(mat, 12, 10, 4, 0, method="auto")
    assert gp.pca_last_report()["method"] == "randomized"
    check(emb, vr, comp, R.rsvd(ip, ix, dx, 1500, 33000, 12, 10, 4, 0, True))


def test_reference_acceptance_bound():
    # python/polariseq/tests/test_pca.py:38-62 on its structured fixture
    print("rsvd reference bound", flush=True)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(4000, 6))
    v = rng.normal(size=(1500, 6))
    x = (u * np.array([600, 400, 250, 150, 100, 60.0])) @ v.T + rng.normal(size=(4000, 1500)) * 0.5
    x = np.where(rng.random((4000, 1500)) < 0.3, np.abs(x), 0.0).astype(np.float32)
    emb, vr, _ = run(sp.csr_matrix(x), 10, 10, 4, 0)
    xc = x.astype(np.float64) - x.astype(np.float64).mean(axis=0)
    uu, s, _ = np.linalg.svd(xc, full_matrices=False)
    emb = emb.astype(np.float64)
    assert (np.abs(emb.mean(axis=0)) / emb.std(axis=0)).max() < 1e-3
    np.testing.assert_allclose(vr[:6], s[:6] ** 2 / (s ** 2).sum(), rtol=5e-3)
    np.testing.assert_allclose(np.linalg.norm(emb, axis=0)[:6], s[:6], rtol=5e-3)
    for kk in (1, 2, 4, 6):
        qa, _ = np.linalg.qr(emb[:, :kk])
        sv = np.linalg.svd(qa.T @ uu[:, :kk], compute_uv=False)
        assert np.degrees(np.arccos(np.clip(sv, -1, 1))).max() < 1.0


@pytest.mark.parametrize("n,m,k,over,it", [(60, 300, 5, 10, 4), (400, 30, 3, 5, 3)])
def test_basis_spanning_the_space_is_exact(n, m, k, over, it):
    print(f"rsvd spanning n={n} m={m}", flush=True)
    mat = structured(n, m, 4, 0.5, n + m)
    x = mat.toarray().astype(np.float64)
    emb, vr, comp = run(mat, k, over, it, 5)
    xc = x - (x.sum(axis=0) / n).astype(np.float32).astype(np.float64)
    _, s, _ = np.linalg.svd(xc, full_matrices=False)
    np.testing.assert_allclose(np.linalg.norm(emb.astype(np.float64), axis=0), s[:k], rtol=1e-5)
    np.testing.assert_allclose(vr, (s[:k] ** 2 / (s ** 2).sum()).astype(np.float32), rtol=1e-5)
    ip, ix, dx = parts(mat)
    check(emb, vr, comp, R.rsvd(ip, ix, dx, n, m, k, over, it, 5, True))


def test_rank_deficient_matrix():
    # rank 2 (no noise) with a 15-column sketch: the basis has null directions
    print("rsvd rank 2", flush=True)
    rng = np.random.default_rng(9)
    x = (np.outer(rng.integers(1, 5, 200), rng.integers(0, 3, 100))
         + np.outer(rng.integers(0, 4, 200), rng.integers(1, 3, 100))).astype(np.float32)
    emb, vr, comp = run(sp.csr_matrix(x), 5, 10, 1, 0)
    assert np.all(np.isfinite(emb)) and np.all(np.isfinite(vr)) and np.all(np.isfinite(comp))
    xc = x.astype(np.float64) - (x.astype(np.float64).sum(axis=0) / 200).astype(np.float32)
    _, s, _ = np.linalg.svd(xc, full_matrices=False)
    np.testing.assert_allclose(np.linalg.norm(emb[:, :2].astype(np.float64), axis=0), s[:2], rtol=1e-5)
    assert np.all(vr[2:] < 1e-10) and abs(float(vr[:2].sum()) - 1.0) < 1e-5


def test_repeats_bit_for_bit_and_device_entry_agrees():
    import torch

    print("rsvd repeat", flush=True)
    mat = structured(5000, 700, 8, 0.15, 21)
    a = run(mat, 15, 10, 2, 3)
    b = run(mat, 15, 10, 2, 3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    ip, ix, dx = parts(mat)
    dev = torch.device("cuda:0")
    comp = torch.zeros((700, 15), dtype=torch.float32, device=dev)
    e, r = gp.pca_csr_device(torch.from_numpy(ip.astype(np.int64)).to(dev),
                             torch.from_numpy(ix.astype(np.int32)).to(dev), torch.from_numpy(dx).to(dev), 700, 15,
                             components=comp, method="randomized", oversample=10, n_iter=2, seed=3)
    torch.cuda.synchronize()
    assert np.array_equal(e.cpu().numpy(), a[0]) and np.array_equal(r.cpu().numpy(), a[1])
    assert np.array_equal(comp.cpu().numpy(), a[2])

"""The acceptance checks of the reference's pp.pca tests
(python/polariseq/tests/test_pca.py:45-140), with their tolerances, run on the
GPU PCA over low-rank count matrices of our own (tests/synth.py):
  * against the exact float64 SVD of the centred matrix: centred scores, the
    leading variance ratios and score norms within 0.5%, the leading 1, 2, 4 and
    6-dimensional score subspaces within 1 degree (this exact route is then held
    to 1e-5 and 1e-3 degrees);
  * against scanpy's scale-then-PCA restated in float64 numpy (scanpy is not
    installed): ratios within 0.1%, the 6-dimensional subspace within 0.1 degree;
  * two runs bit-identical; the uncentred first component dominated by its mean;
    the input arrays left unchanged; the dense form equal to the sparse one.
And the point of a GPU-resident PCA: the embedding goes from the PCA to
NN-Descent without leaving HBM, and gives the same graph as the host route.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from polariseq_amd import pca as gp

from synth import low_rank_counts, normalize_log1p, subspace_angle_deg

pytestmark = pytest.mark.gpu


def test_matches_exact_svd():
    print("test_matches_exact_svd", flush=True)
    x = low_rank_counts()
    r = gp.pca_sparse(x, 10, True, 10, 4, 0)
    emb = r["embedding"].astype(np.float64)
    vr = r["variance_ratio"].astype(np.float64)
    xc = x.toarray().astype(np.float64)
    xc -= xc.mean(axis=0)
    u, s, _ = np.linalg.svd(xc, full_matrices=False)
    emb_ref = u[:, :10] * s[:10]
    vr_ref = s[:10] ** 2 / (s ** 2).sum()
    assert (np.abs(emb.mean(axis=0)) / emb.std(axis=0)).max() < 1e-3
    np.testing.assert_allclose(vr[:6], vr_ref[:6], rtol=5e-3)
    np.testing.assert_allclose(np.linalg.norm(emb, axis=0)[:6], s[:6], rtol=5e-3)
    for k in (1, 2, 4, 6):
        assert subspace_angle_deg(emb[:, :k], emb_ref[:, :k]) < 1.0
    # the exact route does better than the reference's bounds for its randomized one
    np.testing.assert_allclose(vr, vr_ref, rtol=1e-5)
    assert subspace_angle_deg(emb, emb_ref) < 1e-3


def test_is_bit_identical_across_runs():
    x = normalize_log1p(low_rank_counts(n=3000, m=2000, seed=7))
    a = gp.pca_sparse(x, 30, True, 10, 4, 7)
    b = gp.pca_sparse(x, 30, True, 10, 4, 7)
    np.testing.assert_array_equal(a["embedding"], b["embedding"])
    np.testing.assert_array_equal(a["variance_ratio"], b["variance_ratio"])


def test_uncentered_is_available():
    x = low_rank_counts(n=2500, m=1200, seed=2)
    pc1 = gp.pca_sparse(x, 5, False, 10, 4, 0)["embedding"][:, 0].astype(np.float64)
    assert abs(pc1.mean()) > pc1.std()


@pytest.mark.parametrize("max_value", [None, 10.0, 2.0])
def test_matches_scale_then_pca(max_value):
    print(f"test_matches_scale_then_pca max_value={max_value}", flush=True)
    x = normalize_log1p(low_rank_counts(n=3000, m=1200, seed=3))
    r = gp.pca_sparse_scaled(x, 10, max_value, 0)
    # scanpy's pp.scale (ddof 1, zero sd -> 1, clip) and a full SVD, in float64
    d = x.toarray().astype(np.float64)
    mu = d.mean(axis=0)
    sd = d.std(axis=0, ddof=1)
    sd[sd == 0] = 1.0
    z = (d - mu) / sd
    if max_value is not None:
        z = np.clip(z, -max_value, max_value)
    zc = z - z.mean(axis=0)
    u, s, _ = np.linalg.svd(zc, full_matrices=False)
    emb_ref = u[:, :10] * s[:10]
    vr_ref = s[:10] ** 2 / (s ** 2).sum()
    emb = r["embedding"].astype(np.float64)
    np.testing.assert_allclose(r["variance_ratio"][:6], vr_ref[:6], rtol=1e-3)
    assert subspace_angle_deg(emb[:, :6], emb_ref[:, :6]) < 0.1


def test_leaves_input_unchanged():
    x = low_rank_counts(n=1500, m=800, seed=4)
    before = (x.indptr.copy(), x.indices.copy(), x.data.copy())
    gp.pca_sparse_scaled(x, 5, 10.0, 0)
    assert np.array_equal(x.indptr, before[0]) and np.array_equal(x.indices, before[1])
    assert np.array_equal(x.data, before[2])


def test_embedding_stays_in_hbm_for_the_graph():
    """pca_csr_device -> nndescent_device on the device tensor equals the host
    route (pca_sparse -> nndescent_dense) bit for bit."""
    import torch

    import polariseq_amd as P
    x = normalize_log1p(low_rank_counts(n=12000, m=1000, rank=6, seed=5))
    dev = torch.device("cuda:0")
    ip = torch.from_numpy(x.indptr.astype(np.int64)).to(dev)
    ix = torch.from_numpy(x.indices.astype(np.int32)).to(dev)
    dx = torch.from_numpy(x.data.astype(np.float32)).to(dev)
    emb_d, _ = gp.pca_csr_device(ip, ix, dx, 1000, 50)
    g_ip, g_ix, g_dx, nnz = P.nndescent_device(emb_d, 15, seed=3)
    torch.cuda.synchronize()
    host = gp.pca_sparse(x, 50, True, 10, 4, 0)["embedding"]
    assert np.array_equal(emb_d.cpu().numpy(), host)
    want = P.nndescent_dense(host, 15, seed=3)
    assert np.array_equal(g_ip.cpu().numpy().view(np.uint32), want["indptr"])
    assert np.array_equal(g_ix[:nnz].cpu().numpy().view(np.uint32), want["indices"])
    assert np.array_equal(g_dx[:nnz].cpu().numpy(), want["data"])


def test_dense_input():
    """pca_dense (lib.rs:355-388): the dense form agrees with the sparse call on the
    same matrix bit for bit, and with the exact SVD at the reference's bounds
    (its dense route is a thin SVD, pca.rs:567-650)."""
    x = low_rank_counts(n=2000, m=600, seed=8)
    d = x.toarray()
    r = gp.pca_dense(d, 8, True, 0)
    s_ = gp.pca_sparse(x, 8, True, 10, 4, 0)
    assert np.array_equal(r["embedding"], s_["embedding"])
    xc = d.astype(np.float64) - d.astype(np.float64).mean(axis=0)
    u, s, _ = np.linalg.svd(xc, full_matrices=False)
    np.testing.assert_allclose(r["variance_ratio"], s[:8] ** 2 / (s ** 2).sum(), rtol=1e-5)
    assert subspace_angle_deg(r["embedding"].astype(np.float64), u[:, :8] * s[:8]) < 1e-3


def test_more_genes_than_the_reference_keeps_exact():
    """20,000 genes, past the reference's switch to randomized SVD (pca.rs:201,
    4,096): the exact route's embedding against the thin SVD of the centred
    matrix (from the eigenvectors of X_c X_c^T, the cells being fewer), held to
    the exact route's bounds."""
    x = low_rank_counts(n=2500, m=20000, rank=6, seed=11, density=0.05)
    r = gp.pca_sparse(x, 10, True, 10, 4, 0)
    xc = x.toarray().astype(np.float64)
    xc -= xc.mean(axis=0)
    w, u = np.linalg.eigh(xc @ xc.T)
    order = np.argsort(w)[::-1]
    s = np.sqrt(np.maximum(w[order], 0.0))
    emb_ref = u[:, order[:10]] * s[:10]
    vr_ref = s[:10] ** 2 / (s ** 2).sum()
    emb = r["embedding"].astype(np.float64)
    np.testing.assert_allclose(r["variance_ratio"], vr_ref, rtol=1e-5)
    assert subspace_angle_deg(emb[:, :6], emb_ref[:, :6]) < 1e-3

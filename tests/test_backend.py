"""The ComputeBackend mirror (polariseq_amd/backend.py, crates/polariseq-core/src/backend.rs:65-141):
names, defaults, and the trait's error texts for inputs rejected before any
device work (CPU); PCA and neighbours through the backend equal the direct calls
(GPU)."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from polariseq_amd import backend as B
from polariseq_amd import neighbors_dense, pca_sparse


def test_names_and_defaults():
    b = B.RocmBackend(0)
    assert b.kind() == "rocm" and b.name() == "rocm:0"
    assert B.PcaOptions() == B.PcaOptions(50, True, 10, 1, 0)  # pca.rs:73-83
    assert B.NeighborsOptions() == B.NeighborsOptions(15, "euclidean", 0)  # neighbors.rs:52-60


def test_errors_carry_the_trait_texts():
    b = B.RocmBackend(0)
    with pytest.raises(B.BackendError, match=r"^PCA failed: n_components \(6\) exceeds min\(n_obs, n_features\) \(5\)$"):
        b.pca(np.eye(5, dtype=np.float32), B.PcaOptions(n_components=6))
    with pytest.raises(B.BackendError, match=r"^PCA failed: input matrix has zero rows \(0\)"):
        b.pca(sp.csr_matrix((0, 4), dtype=np.float32), B.PcaOptions(n_components=1))
    with pytest.raises(B.BackendError, match=r"^neighbors failed: embedding has zero rows \(0\) or dims \(5\)$"):
        b.neighbors(np.zeros(0, np.float32), 0, 5)
    with pytest.raises(B.BackendError, match="not n_obs \\* n_dim"):
        b.neighbors(np.zeros(10, np.float32), 3, 4)


@pytest.mark.gpu
def test_backend_equals_direct_calls(orc):
    rng = np.random.default_rng(0)
    x = np.where(rng.random((1500, 300)) < 0.2, rng.poisson(3.0, (1500, 300)), 0).astype(np.float32)
    b = B.RocmBackend(0)
    r = b.pca(sp.csr_matrix(x), B.PcaOptions(n_components=20))
    d = pca_sparse(sp.csr_matrix(x), 20, True, 10, 1, 0)
    assert np.array_equal(r.embedding, d["embedding"]) and np.array_equal(r.variance_ratio, d["variance_ratio"])
    rd = b.pca(x, B.PcaOptions(n_components=20))  # the dense form of Matrix
    assert np.array_equal(rd.embedding, d["embedding"])
    assert (r.n_obs, r.n_components) == (1500, 20)
    for metric in ("euclidean", "cosine"):
        g = b.neighbors(r.embedding.reshape(-1), 1500, 20, B.NeighborsOptions(10, metric))
        want = neighbors_dense(r.embedding, 10, metric)
        assert np.array_equal(g.indptr, want["indptr"]) and np.array_equal(g.indices, want["indices"])
        assert np.array_equal(g.data, want["data"]) and g.n_obs == 1500
        # and the reference's exact path (neighbors.rs:101-241), restated by the oracle
        ref = orc.neighbors_exact(r.embedding, 10, metric)
        assert np.array_equal(g.indptr, ref["indptr"]) and np.array_equal(g.indices, ref["indices"])
        assert np.array_equal(g.data.view(np.uint32), ref["data"].view(np.uint32))

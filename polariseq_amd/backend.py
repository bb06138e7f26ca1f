"""A GPU implementation of the reference's backend trait
(crates/polariseq-core/src/backend.rs:65-141): ``ComputeBackend`` with ``kind``,
``name``, ``pca`` and ``neighbors``, the plug-in point the reference leaves for
other devices (its ``BackendKind`` lists no ROCm target, backend.rs:51-63).

``RocmBackend.pca`` takes a scipy CSR matrix or a dense 2-D array (``Matrix``'s
two forms, the dense one stored as CSR of its nonzeros) and runs the exact Gram
PCA; ``RocmBackend.neighbors`` takes a row-major embedding and runs the exact
kNN graph of ``neighbors()`` (crates/polariseq-core/src/neighbors.rs:101-147). Failures are
``BackendError`` with the trait's texts: ``"PCA failed: ..."`` and
``"neighbors failed: ..."`` (backend.rs:97-113).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import pca as _pca
from .neighbors import neighbors_dense


@dataclass
class PcaOptions:
    """pca.rs:56-83 (defaults of ``PcaOptions::default``)."""
    n_components: int = 50
    center: bool = True
    oversample: int = 10
    n_iter: int = 1
    seed: int = 0


@dataclass
class NeighborsOptions:
    """neighbors.rs:43-60."""
    n_neighbors: int = 15
    metric: str = "euclidean"
    seed: int = 0


@dataclass
class PcaResult:
    embedding: np.ndarray  # float32 (n_obs, n_components)
    variance_ratio: np.ndarray
    n_obs: int
    n_components: int


@dataclass
class NeighborGraph:
    """backend.rs:37-46: CSR distance graph, self excluded."""
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    n_obs: int = field(default=0)


class BackendError(RuntimeError):
    """backend.rs:97-113."""


class RocmBackend:
    """``ComputeBackend`` on one MI355X (HIP device ``device``)."""

    def __init__(self, device: int = 0):
        self.device = int(device)

    def kind(self) -> str:
        return "rocm"

    def name(self) -> str:
        return f"rocm:{self.device}"

    def pca(self, matrix, options: PcaOptions | None = None) -> PcaResult:
        o = options or PcaOptions()
        if isinstance(matrix, np.ndarray):
            import scipy.sparse as sp
            if matrix.ndim != 2:
                raise BackendError("PCA failed: a dense matrix must be 2-D")
            matrix = sp.csr_matrix(np.asarray(matrix, dtype=np.float32))
        try:
            r = _pca.pca_sparse(matrix, o.n_components, o.center, o.oversample, o.n_iter, o.seed,
                                device=self.device)
        except RuntimeError as e:  # already "PCA failed: ..."
            raise BackendError(str(e)) from e
        emb = r["embedding"]
        return PcaResult(emb, r["variance_ratio"], emb.shape[0], r["n_components"])

    def neighbors(self, embedding, n_obs: int, n_dim: int,
                  options: NeighborsOptions | None = None) -> NeighborGraph:
        o = options or NeighborsOptions()
        emb = np.asarray(embedding, dtype=np.float32).reshape(-1)
        if emb.size != n_obs * n_dim:
            raise BackendError(f"neighbors failed: embedding has {emb.size} values, "
                               f"not n_obs * n_dim = {n_obs * n_dim}")
        try:
            g = neighbors_dense(emb.reshape(n_obs, n_dim), o.n_neighbors, o.metric,
                                device=self.device)
        except RuntimeError as e:  # already "neighbors failed: ..."
            raise BackendError(str(e)) from e
        return NeighborGraph(g["indptr"], g["indices"], g["data"], g["n_obs"])

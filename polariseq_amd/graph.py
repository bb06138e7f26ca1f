"""The graph computation of ``pp.neighbors`` (python/polariseq/__init__.py:2514-2600)
without AnnData, on the GPU.

``neighbors_graph(embedding, n_neighbors, ...)`` follows the reference step for
step: ``n_pcs`` columns, ``k = max(1, n_neighbors - 1)`` neighbours searched
(n_neighbors counts the cell itself), the exact search up to 10,000 cells and
NN-Descent above unless ``approximate`` says otherwise, ``metric`` euclidean or
cosine (on the approximate path: NN-Descent on unit rows, distances converted
to 1 - cos), the same ``ValueError`` text, and the symmetric fuzzy
connectivities of the distance graph. Result arrays come from ``alloc(name,
shape, dtype)`` (the role of ``_results.new``: memory-mapped result files in a
project; plain arrays by default) under the reference's names
(``distances_indptr`` ... ``connectivities_data``); the approximate path
reserves its distance arrays at ``n_obs * k`` entries, as the reference does.

On the approximate path the embedding is uploaded once and every step runs on
the device: unit rows (``psq_unit_rows_device``), NN-Descent, the cosine
conversion (``psq_cosine_distances_device``) and the connectivities; the host
receives the finished arrays.
"""

from __future__ import annotations

import numpy as np

from . import _native as N
from .fuzzy import fuzzy_connectivities_device, fuzzy_connectivities_into
from .neighbors import _stream_ptr, neighbors_dense, nndescent_device

EXACT_MAX_CELLS = 10_000  # __init__.py:2522: exact for small data


def _check_device_rows(x):
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, d)")


def unit_rows_device(x, out=None):
    """Rows of a CUDA float32 (n, d) tensor scaled to unit length, bit for bit as
    numpy's ``emb / np.maximum(np.linalg.norm(emb, axis=1, keepdims=True), 1e-12)``."""
    import torch
    _check_device_rows(x)
    n, d = x.shape
    if out is None:
        out = torch.empty_like(x)
    N.check(N.lib().psq_unit_rows_device(x.data_ptr(), n, d, out.data_ptr(), _stream_ptr()))
    return out


def cosine_distances_device(dist, count: int | None = None):
    """In place: ``dist[:count] = float32(float64(dist)^2 / 2)`` (1 - cos from the
    Euclidean distance of unit rows)."""
    import torch
    if not (dist.is_cuda and dist.dtype == torch.float32 and dist.is_contiguous()):
        raise ValueError("dist must be a contiguous float32 CUDA tensor")
    c = dist.numel() if count is None else int(count)
    N.check(N.lib().psq_cosine_distances_device(dist.data_ptr(), c, _stream_ptr()))
    return dist


def _default_alloc(name: str, shape, dtype):
    return np.empty(shape, dtype)


def neighbors_graph(embedding, n_neighbors: int = 15, *, n_pcs: int | None = None,
                    metric: str = "euclidean", approximate: bool | None = None, seed: int = 42,
                    alloc=None, device: int = 0) -> dict:
    """``pp.neighbors``' distances and connectivities of an embedding (n_obs, n_dim),
    with its defaults (n_neighbors 15, seed 42; __init__.py:2477-2486).

    Returns ``{"distances": {indptr, indices, data}, "connectivities": {indptr,
    indices, data}, "uns": {n_neighbors, metric, method, connectivities}}``;
    ``indices`` / ``data`` are views of the allocated arrays cut to the entries
    written."""
    emb = np.asarray(embedding)
    if n_pcs is not None:
        emb = emb[:, :n_pcs]
    if emb.ndim != 2:
        raise ValueError("embedding must be a 2-D array")
    n_obs = emb.shape[0]
    k = max(1, int(n_neighbors) - 1)
    use_approx = approximate if approximate is not None else (n_obs > EXACT_MAX_CELLS)
    if metric not in ("euclidean", "cosine"):
        raise ValueError(f"metric must be 'euclidean' or 'cosine', not {metric!r}")
    new = alloc or _default_alloc
    emb = np.ascontiguousarray(emb, dtype=np.float32)
    conn: dict[str, np.ndarray] = {}

    def conn_alloc(n_rows: int, n_entries: int):
        conn["indptr"] = new("connectivities_indptr", (n_rows + 1,), np.uint32)
        conn["indices"] = new("connectivities_indices", (n_entries,), np.uint32)
        conn["data"] = new("connectivities_data", (n_entries,), np.float32)
        return conn["indptr"], conn["indices"], conn["data"]

    if use_approx:
        import torch
        x = torch.from_numpy(emb).to(torch.device("cuda", device))
        search = unit_rows_device(x) if metric == "cosine" else x
        cap = n_obs * max(1, min(k, n_obs - 1))
        d_ip = new("distances_indptr", (n_obs + 1,), np.uint32)
        d_ix = new("distances_indices", (cap,), np.uint32)
        d_dx = new("distances_data", (cap,), np.float32)
        ip, ix, dx, nnz = nndescent_device(search, k, 0, seed)
        del search, x
        if metric == "cosine":
            cosine_distances_device(dx, nnz)
        cip, cix, cw, cnnz = fuzzy_connectivities_device(ip, ix, dx, n_obs)
        torch.cuda.synchronize(device)
        d_ip[...] = ip.cpu().numpy().view(np.uint32)
        d_ix[:nnz] = ix[:nnz].cpu().numpy().view(np.uint32)
        d_dx[:nnz] = dx[:nnz].cpu().numpy()
        c_ip, c_ix, c_w = conn_alloc(n_obs, cnnz)
        c_ip[...] = cip.cpu().numpy().view(np.uint32)
        c_ix[...] = cix[:cnnz].cpu().numpy().view(np.uint32)
        c_w[...] = cw[:cnnz].cpu().numpy()
        method = "nndescent"
    else:
        result = neighbors_dense(emb, n_neighbors, metric, device=device)
        nnz = len(result["indices"])
        d_ip = new("distances_indptr", (n_obs + 1,), np.uint32)
        d_ix = new("distances_indices", (nnz,), np.uint32)
        d_dx = new("distances_data", (nnz,), np.float32)
        d_ip[...], d_ix[...], d_dx[...] = result["indptr"], result["indices"], result["data"]
        fuzzy_connectivities_into(np.ascontiguousarray(d_ip), np.ascontiguousarray(d_ix[:nnz]),
                                  np.ascontiguousarray(d_dx[:nnz]), n_obs, conn_alloc,
                                  device=device)
        method = "exact"
    return {
        "distances": {"indptr": d_ip, "indices": d_ix[:nnz], "data": d_dx[:nnz]},
        "connectivities": conn,
        "uns": {"n_neighbors": n_neighbors, "metric": metric, "method": method,
                "connectivities": True},
    }

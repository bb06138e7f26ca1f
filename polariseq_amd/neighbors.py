"""Python face of the neighbour-graph step, mirroring the reference binding.

* ``nndescent_dense`` mirrors ``crates/polariseq-python/src/lib.rs:503-566``: same
  positional arguments, ``out=(indptr, indices, dists)`` written in place with the
  same size check and ``ValueError`` text, the same returned dict.
* ``neighbors_dense`` mirrors ``crates/polariseq-python/src/lib.rs:471-501``
  (exact search, ``n_neighbors`` counting the cell itself).
* ``knn_exact`` is the exact tiled kNN with ties broken by index
  (``core/doublets.rs:544-651`` order).
* ``*_device`` variants take and return torch tensors on the GPU (embedding
  resident in HBM, graph left in HBM), for callers and benchmarks that keep data
  on the device. torch supplies the device memory and stream; it computes nothing.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def _model_handle(model):
    if model is None:
        return None
    h = getattr(model, "handle", model)
    return h.value if isinstance(h, ctypes.c_void_p) else h


def effective_k(n_obs: int, k: int) -> int:
    """nndescent.rs:648-650."""
    return int(N.lib().psq_effective_k(int(n_obs), int(k)))


def _as_f32_2d(embedding) -> np.ndarray:
    a = np.asarray(embedding)
    if a.ndim != 2:
        raise ValueError("embedding must be a 2-D array")
    # pybind/lib.rs:526-533: read in place when contiguous, else copy.
    return np.ascontiguousarray(a, dtype=np.float32)


def nndescent_dense(embedding, k: int, max_iterations: int = 0, seed: int = 0,
                    max_candidates: int = 0, n_trees: int = 0, delta: float = 0.001,
                    out=None, *, precision: str = "fp32", init: str = "rp_forest",
                    device: int = 0, leaf_size: int = 0, slots: int = 0, model=None,
                    heap_k: int = 0, reorder: str = "auto") -> dict:
    """Approximate kNN graph by NN-Descent on the GPU (pybind/lib.rs:503-566).

    ``max_iterations``, ``max_candidates`` and ``n_trees`` accept 0 for "auto".
    ``heap_k``: entries each row keeps while iterating (0 = k for fp32, 2k for
    bf16); the graph holds each row's first k. ``reorder``: "auto" (on for bf16
    with a forest start), "on" or "off": renumber the points in a locality order
    while iterating (the graph comes back in the caller's numbering).
    Returns ``{"indptr", "indices", "data", "n_obs"}``, or with ``out`` writes the
    three arrays in place and returns ``{"nnz", "n_obs"}``.
    """
    emb = _as_f32_2d(embedding)
    n_obs, n_dim = emb.shape
    opts = N.make_opts(k, max_iterations, seed, max_candidates, n_trees, delta, leaf_size,
                       precision, init, device, slots, _model_handle(model), heap_k, reorder)
    nnz = ctypes.c_uint64(0)
    L = N.lib()
    if out is not None:
        ip, ix, dx = out
        cap = n_obs * effective_k(n_obs, k) if n_obs > 0 else 0
        for a, dt in ((ip, np.uint32), (ix, np.uint32), (dx, np.float32)):
            if not isinstance(a, np.ndarray) or a.dtype != dt or a.ndim != 1 or not a.flags.c_contiguous \
                    or not a.flags.writeable:
                raise TypeError("out must be (uint32[n+1], uint32[cap], float32[cap]) writable "
                                "contiguous arrays")
        if ip.size != n_obs + 1 or ix.size < cap or dx.size < cap:
            raise ValueError(f"outputs must hold {n_obs + 1} row pointers and {cap} neighbours")
        N.check(L.psq_nndescent_into(emb.ctypes.data, n_obs, n_dim, ctypes.byref(opts),
                                     ip.ctypes.data, ip.size, ix.ctypes.data, dx.ctypes.data,
                                     min(ix.size, dx.size), ctypes.byref(nnz)))
        return {"nnz": int(nnz.value), "n_obs": n_obs}
    ke = effective_k(n_obs, k) if k > 0 else 0
    ip = np.empty(n_obs + 1, np.uint32)
    ix = np.empty(max(n_obs * ke, 1), np.uint32)
    dx = np.empty(max(n_obs * ke, 1), np.float32)
    N.check(L.psq_nndescent_into(emb.ctypes.data, n_obs, n_dim, ctypes.byref(opts),
                                 ip.ctypes.data, ip.size, ix.ctypes.data, dx.ctypes.data,
                                 n_obs * ke, ctypes.byref(nnz)))
    m = int(nnz.value)
    return {"indptr": ip, "indices": ix[:m], "data": dx[:m], "n_obs": n_obs}


def neighbors_dense(embedding, n_neighbors: int, metric: str = "euclidean", *,
                    device: int = 0) -> dict:
    """Exact kNN graph (pybind/lib.rs:474-501): n_neighbors counts the cell itself."""
    emb = _as_f32_2d(embedding)
    n_obs, n_dim = emb.shape
    m = {"euclidean": 0, "cosine": 1}.get(str(metric).lower())
    if m is None:
        raise RuntimeError(f"unknown metric: {metric}")
    kk = max(0, min(int(n_neighbors), n_obs) - 1)
    ip = np.empty(n_obs + 1, np.uint32)
    ix = np.empty(max(n_obs * kk, 1), np.uint32)
    dx = np.empty(max(n_obs * kk, 1), np.float32)
    nnz = ctypes.c_uint64(0)
    rc = N.lib().psq_neighbors_exact(emb.ctypes.data, n_obs, n_dim, int(n_neighbors), m,
                                     ip.ctypes.data, ix.ctypes.data, dx.ctypes.data,
                                     n_obs * kk, ctypes.byref(nnz), int(device))
    if rc != N.PSQ_OK:
        msg = N.lib().psq_last_error().decode()
        raise RuntimeError(f"neighbors failed: {msg}")
    c = int(nnz.value)
    return {"indptr": ip, "indices": ix[:c], "data": dx[:c], "n_obs": n_obs}


def neighbors_dense_device(x, n_neighbors: int, metric: str = "euclidean"):
    """``neighbors_dense`` on a contiguous float32 CUDA tensor (n_obs, n_dim): the
    graph stays on the device. Returns (indptr, indices, data) CUDA tensors (int32
    views of the u32 arrays, float32) with n_obs * (min(n_neighbors, n_obs) - 1)
    entries."""
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, d)")
    n_obs, n_dim = x.shape
    m = {"euclidean": 0, "cosine": 1}.get(str(metric).lower())
    if m is None:
        raise RuntimeError(f"unknown metric: {metric}")
    kk = max(0, min(int(n_neighbors), n_obs) - 1)
    ip = torch.empty(n_obs + 1, dtype=torch.int32, device=x.device)
    ix = torch.empty(max(n_obs * kk, 1), dtype=torch.int32, device=x.device)
    dx = torch.empty(max(n_obs * kk, 1), dtype=torch.float32, device=x.device)
    nnz = ctypes.c_uint64(0)
    rc = N.lib().psq_neighbors_exact_device(x.data_ptr(), n_obs, n_dim, int(n_neighbors), m,
                                            ip.data_ptr(), ix.data_ptr(), dx.data_ptr(),
                                            n_obs * kk, ctypes.byref(nnz), _stream_ptr())
    if rc != N.PSQ_OK:
        msg = N.lib().psq_last_error().decode()
        raise RuntimeError(f"neighbors failed: {msg}")
    c = int(nnz.value)
    return ip, ix[:c], dx[:c]


def knn_exact(data, k: int, queries=None, *, device: int = 0):
    """Exact k nearest (self excluded) by (FP64 distance, index).

    Returns (idx u32[n_q, k], dist f32[n_q, k]); missing entries are 0xFFFFFFFF / inf.
    """
    x = _as_f32_2d(data)
    n, d = x.shape
    if queries is None:
        q = None
        nq = n
    else:
        q = np.ascontiguousarray(queries, dtype=np.uint32)
        nq = q.size
    idx = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float32)
    N.check(N.lib().psq_knn_exact(x.ctypes.data, n, d, None if q is None else q.ctypes.data, nq,
                                  int(k), idx.ctypes.data, dist.ctypes.data, int(device)))
    return idx, dist


def knn_exact_batched(data, offsets, k: int, *, device: int = 0):
    """Exact kNN of every point within its own sample: samples are the row ranges
    [offsets[s], offsets[s+1]). Returns (idx u32[n, k] global row indices,
    dist f32[n, k]); one persistent launch per pass over all samples, tile shapes
    from the autotuner."""
    x = _as_f32_2d(data)
    n, d = x.shape
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    idx = np.empty((n, k), np.uint32)
    dist = np.empty((n, k), np.float32)
    N.check(N.lib().psq_knn_exact_batched(x.ctypes.data, n, d, off.ctypes.data, off.size - 1,
                                          int(k), idx.ctypes.data, dist.ctypes.data, int(device)))
    return idx, dist


def knn_exact_batched_device(x, offsets, k: int):
    """Device form of knn_exact_batched (x: CUDA float32 (n, d); offsets: host array)."""
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, d)")
    n, d = x.shape
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    idx = torch.empty((n, k), dtype=torch.int32, device=x.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=x.device)
    N.check(N.lib().psq_knn_exact_batched_device(x.data_ptr(), n, d, off.ctypes.data, off.size - 1,
                                                 int(k), idx.data_ptr(), dist.data_ptr(),
                                                 _stream_ptr()))
    return idx, dist


def batched_report() -> dict:
    r = N.BatchedReport()
    N.check(N.lib().psq_batched_last_report(ctypes.byref(r)))
    return r.as_dict()


# ── device-resident forms (torch tensors on the GPU) ──

def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def nndescent_device(x, k: int, max_iterations: int = 0, seed: int = 0, max_candidates: int = 0,
                     n_trees: int = 0, delta: float = 0.001, *, precision: str = "fp32",
                     init: str = "rp_forest", leaf_size: int = 0, slots: int = 0, out=None,
                     model=None, heap_k: int = 0, reorder: str = "auto"):
    """NN-Descent on a CUDA float32 tensor (n, d). Returns (indptr, indices, dists, nnz)
    as device tensors (indices/dists sized n * effective_k; valid up to nnz)."""
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, d)")
    n, d = x.shape
    ke = effective_k(n, k) if (n > 0 and k > 0) else 0
    if out is None:
        ip = torch.empty(n + 1, dtype=torch.int32, device=x.device)
        ix = torch.empty(max(n * ke, 1), dtype=torch.int32, device=x.device)
        dx = torch.empty(max(n * ke, 1), dtype=torch.float32, device=x.device)
    else:
        ip, ix, dx = out
    opts = N.make_opts(k, max_iterations, seed, max_candidates, n_trees, delta, leaf_size,
                       precision, init, x.device.index or 0, slots, _model_handle(model),
                       heap_k, reorder)
    nnz = ctypes.c_uint64(0)
    N.check(N.lib().psq_nndescent_device(x.data_ptr(), n, d, ctypes.byref(opts), ip.data_ptr(),
                                         ip.numel(), ix.data_ptr(), dx.data_ptr(),
                                         min(ix.numel(), dx.numel()), ctypes.byref(nnz),
                                         _stream_ptr()))
    return ip, ix, dx, int(nnz.value)


def knn_exact_device(x, k: int, queries=None):
    """Exact kNN on a CUDA float32 tensor; queries: CUDA int32 tensor of row indices."""
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, d)")
    n, d = x.shape
    nq = n if queries is None else queries.numel()
    idx = torch.empty((nq, k), dtype=torch.int32, device=x.device)
    dist = torch.empty((nq, k), dtype=torch.float32, device=x.device)
    qp = None if queries is None else queries.data_ptr()
    N.check(N.lib().psq_knn_exact_device(x.data_ptr(), n, d, qp, nq, int(k), idx.data_ptr(),
                                         dist.data_ptr(), _stream_ptr()))
    return idx, dist


def last_stats() -> dict:
    """Statistics of the last NN-Descent call on this thread."""
    return N.last_stats()


def release_workspace() -> None:
    N.check(N.lib().psq_release_workspace())

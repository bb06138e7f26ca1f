"""Fuzzy connectivities of a kNN distance graph, mirroring the reference binding.

* ``fuzzy_connectivities(indptr, indices, data, n_obs)`` mirrors
  ``crates/polariseq-python/src/lib.rs:568-588``: 1-D uint32 / uint32 / float32
  arrays in, ``{"indptr", "indices", "data"}`` out (a symmetric CSR, rows sorted
  by column, both directions stored).
* ``fuzzy_connectivities_into(indptr, indices, data, n_obs, alloc)`` mirrors
  ``lib.rs:590-642``: the graph is counted, ``alloc(n_rows, nnz)`` returns the
  ``(indptr, indices, weights)`` arrays to fill (memory-mapped result files, say),
  they are filled in place and ``nnz`` is returned. Same ``ValueError`` texts.
* ``fuzzy_connectivities_device`` takes and returns torch tensors on the GPU, for
  a graph that NN-Descent left in HBM.

The arithmetic runs in the HIP kernels of ``csrc/fuzzy_kernels.hip``; there is
no CPU path.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def _array(a, dtype, name: str) -> np.ndarray:
    # pyo3's PyReadonlyArray1<T> accepts only a 1-D array of exactly T.
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.ndim != 1:
        raise TypeError(f"argument '{name}': expected a 1-D numpy array of {np.dtype(dtype).name}")
    return np.ascontiguousarray(a)


def _check(rc: int) -> None:
    # The reference binding raises ValueError for its size errors, the 2^32
    # limit included (lib.rs:609-631).
    if rc in (N.PSQ_E_INVALID, N.PSQ_E_UNSUPPORTED):
        raise ValueError(N.lib().psq_last_error().decode("utf-8", "replace"))
    N.check(rc)


class _Plan:
    """FuzzyPlan (core/umap.rs:327-460) held on the GPU: built, counted, filled."""

    def __init__(self, indptr, indices, data, n_obs: int, device: int):
        ip = _array(indptr, np.uint32, "indptr")
        ix = _array(indices, np.uint32, "indices")
        dx = _array(data, np.float32, "data")
        n_obs = int(n_obs)
        if n_obs < 0:
            raise OverflowError("can't convert negative int to unsigned")
        if len(ip) != n_obs + 1:
            raise ValueError("indptr must have n_obs + 1 entries")
        self.n_obs = n_obs
        self.handle = ctypes.c_void_p(None)
        nnz = ctypes.c_uint64(0)
        m = min(len(ix), len(dx))
        _check(N.lib().psq_fuzzy_plan_new(ip.ctypes.data, len(ip), ix.ctypes.data if m else None,
                                          dx.ctypes.data if m else None, m, n_obs, int(device),
                                          ctypes.byref(self.handle), ctypes.byref(nnz)))
        self.nnz = int(nnz.value)

    def fill(self, indptr: np.ndarray, indices: np.ndarray, weights: np.ndarray) -> None:
        for a in (indptr, indices, weights):
            if not (a.flags.c_contiguous and a.flags.writeable):
                raise TypeError("alloc must return writeable contiguous arrays")
        _check(N.lib().psq_fuzzy_plan_fill(self.handle, indptr.ctypes.data, len(indptr),
                                           indices.ctypes.data if len(indices) else None,
                                           weights.ctypes.data if len(weights) else None,
                                           min(len(indices), len(weights))))

    def close(self) -> None:
        if self.handle:
            N.lib().psq_fuzzy_plan_free(self.handle)
            self.handle = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def fuzzy_connectivities(indptr, indices, data, n_obs: int, *, device: int = 0) -> dict:
    """Symmetrised fuzzy connectivities of a kNN distance graph (lib.rs:568-588,
    core/umap.rs:130-139): ``{"indptr", "indices", "data"}``."""
    with _Plan(indptr, indices, data, n_obs, device) as plan:
        out_ip = np.empty(plan.n_obs + 1, np.uint32)
        out_ix = np.empty(plan.nnz, np.uint32)
        out_w = np.empty(plan.nnz, np.float32)
        plan.fill(out_ip, out_ix, out_w)
    return {"indptr": out_ip, "indices": out_ix, "data": out_w}


def fuzzy_connectivities_into(indptr, indices, data, n_obs: int, alloc, *, device: int = 0) -> int:
    """The fuzzy connectivities written into arrays the caller allocates once
    their size is known (lib.rs:590-642): ``alloc(n_rows, nnz)`` returns
    ``(indptr, indices, weights)`` (uint32 ``n_rows + 1``, uint32 ``nnz``,
    float32 ``nnz``). Returns ``nnz``."""
    with _Plan(indptr, indices, data, n_obs, device) as plan:
        arrays = alloc(plan.n_obs, plan.nnz)
        try:
            out_ip, out_ix, out_w = arrays
        except (TypeError, ValueError) as e:
            raise TypeError("alloc must return a tuple of three arrays") from e
        for a, dt, name in ((out_ip, np.uint32, "indptr"), (out_ix, np.uint32, "indices"),
                            (out_w, np.float32, "weights")):
            if not isinstance(a, np.ndarray) or a.dtype != dt or a.ndim != 1:
                raise TypeError(f"alloc's {name} must be a 1-D numpy array of {np.dtype(dt).name}")
        if len(out_ip) != plan.n_obs + 1 or len(out_ix) != plan.nnz or len(out_w) != plan.nnz:
            raise ValueError("alloc returned arrays of the wrong size")
        plan.fill(out_ip, out_ix, out_w)
        return plan.nnz


def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def fuzzy_connectivities_device(indptr, indices, data, n_obs: int, out=None):
    """The fuzzy connectivities of a graph held in GPU memory (int32 or uint32
    ``indptr``/``indices`` tensors, float32 ``data``). Returns
    ``(indptr, indices, weights, nnz)`` as device tensors; ``indices`` and
    ``weights`` hold ``2 * indptr[n_obs]`` entries, valid up to ``nnz``."""
    import torch
    for t, name in ((indptr, "indptr"), (indices, "indices"), (data, "data")):
        if not (t.is_cuda and t.dim() == 1 and t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous 1-D CUDA tensor")
    if indptr.dtype not in (torch.int32, torch.uint32) or indices.dtype not in (torch.int32,
                                                                                 torch.uint32):
        raise ValueError("indptr and indices must be int32 or uint32 tensors")
    if data.dtype != torch.float32:
        raise ValueError("data must be a float32 tensor")
    n_obs = int(n_obs)
    if indptr.numel() != n_obs + 1:
        raise ValueError("indptr must have n_obs + 1 entries")
    if out is None:
        m = int(indptr[n_obs].item()) & 0xFFFFFFFF
        cap = max(2 * m, 1)
        ip = torch.empty(n_obs + 1, dtype=torch.int32, device=indptr.device)
        ix = torch.empty(cap, dtype=torch.int32, device=indptr.device)
        w = torch.empty(cap, dtype=torch.float32, device=indptr.device)
    else:
        ip, ix, w = out
    nnz = ctypes.c_uint64(0)
    _check(N.lib().psq_fuzzy_connectivities_device(
        indptr.data_ptr(), indices.data_ptr(), data.data_ptr(), n_obs, ip.data_ptr(),
        ix.data_ptr(), w.data_ptr(), min(ix.numel(), w.numel()), ctypes.byref(nnz),
        _stream_ptr()))
    return ip, ix, w, int(nnz.value)

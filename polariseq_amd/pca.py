"""PCA of a sparse matrix on the GPU: the reference binding's ``pca_sparse`` and
``pca_sparse_scaled`` (crates/polariseq-python/src/lib.rs:395-467), the calls
``pp.pca`` makes for sparse input (python/polariseq/__init__.py:2421-2447).

Both take the exact Gram route of the reference (crates/polariseq-core/src/pca.rs:487-535):
C = S^T S - n mu mu^T in float64, its top ``n_components`` eigenvectors, the
projection of the cells (float32, each row summed over its stored entries in
order), column signs fixed so that each column's largest-magnitude entry is
positive, and variance ratios max(lambda, 0) / trace(C). ``pca_sparse_scaled``
first applies ``pp.scale`` semantics to the stored values (pca.rs:396-480).

Where the reference differs:

- For fewer than 1,000 cells or genes, or n_components * 4 >= min(n, m), the
  reference's ``pca_sparse`` densifies and takes a thin SVD (pca.rs:143-146,
  261-263), whose column signs are arbitrary; here the Gram route runs at every
  size, so the columns carry canonical signs. The two agree to the tolerance
  the reference's own test states (pca.rs:861-905: variance ratios within 1e-4,
  |cos| > 0.9999 per column).
- More than 4,096 genes: the reference switches to randomized block Krylov SVD
  (``oversample``, ``n_iter`` and ``seed`` configure it, rsvd.rs:476-566); this
  build keeps the exact Gram route up to 32,768 genes (the Gram matrix is 8 GiB
  of float64 in HBM; its product costs n_obs x genes^2 flops), whose answer is
  the one the randomized route approximates (the reference's own test bounds it
  against the exact SVD: ratios within 5e-3, principal angles below 1 degree),
  and takes the reference's randomized route, with the reference's ChaCha8
  Gaussian sketch, above that. ``method="randomized"`` asks for the randomized
  route at any size, ``method="gram"`` for the exact one (at most 32,768 genes).

Errors follow the binding: ``RuntimeError("PCA failed: ...")`` for an empty
matrix or too many components (the ``PcaError`` cases), ``ValueError`` for an
``out`` array of the wrong shape or type.
"""

from __future__ import annotations

import math

import numpy as np

from . import _native as N


def _csr_parts(matrix):
    """(indptr u64, indices u32, data f32, n_obs, n_features) of a scipy CSR
    matrix or an (indptr, indices, data, (n_obs, n_features)) tuple."""
    if isinstance(matrix, tuple):
        indptr, indices, data, shape = matrix
    else:
        fmt = getattr(matrix, "format", None)
        if fmt != "csr":
            raise ValueError("matrix must be a scipy CSR matrix or an (indptr, indices, data, shape) tuple")
        indptr, indices, data, shape = matrix.indptr, matrix.indices, matrix.data, matrix.shape
    n, m = int(shape[0]), int(shape[1])
    indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float32)
    if indptr.shape != (n + 1,):
        raise ValueError(f"indptr must have n_obs + 1 = {n + 1} entries")
    nnz = int(indptr[-1]) if n + 1 > 0 else 0
    if indices.shape[0] < nnz or data.shape[0] < nnz:
        raise ValueError("indices and data must hold indptr[-1] entries")
    return indptr, indices, data, n, m


_METHODS = {"auto": N.PSQ_PCA_AUTO, "gram": N.PSQ_PCA_GRAM, "randomized": N.PSQ_PCA_RANDOMIZED}


def _method(method: str) -> int:
    if method not in _METHODS:
        raise ValueError(f"method must be one of {sorted(_METHODS)}, got {method!r}")
    return _METHODS[method]


def _unsigned(**values) -> None:
    """The binding's usize/u64 arguments: a negative value raises OverflowError, as
    pyo3's conversion does."""
    for name, v in values.items():
        if int(v) < 0 or int(v) >= 1 << 64:
            raise OverflowError(f"{name}: can't convert {int(v)} to an unsigned 64-bit integer")


def _run(matrix, n_components: int, center: bool, scale: bool, max_value, out, components, device: int,
         method: str = "auto", oversample: int = 10, n_iter: int = 1, seed: int = 0):
    indptr, indices, data, n, m = _csr_parts(matrix)
    k = int(n_components)
    if out is not None:
        if not (isinstance(out, np.ndarray) and out.dtype == np.float32 and out.shape == (n, k)
                and out.flags.c_contiguous and out.flags.writeable):
            raise ValueError(f"out must be a writable C-contiguous float32 array of shape ({n}, {k})")
        emb = out
    else:
        emb = np.empty((n, max(k, 0)), dtype=np.float32)
    if components is not None and not (isinstance(components, np.ndarray) and components.dtype == np.float32
                                       and components.shape == (m, k) and components.flags.c_contiguous):
        raise ValueError(f"components must be a C-contiguous float32 array of shape ({m}, {k})")
    ratio = np.empty(max(k, 0), dtype=np.float32)
    mv = math.inf if max_value is None else float(max_value)
    L = N.lib()
    meth = _method(method)
    _unsigned(oversample=oversample, n_iter=n_iter, seed=seed)
    rc = L.psq_pca_csr_ex(indptr.ctypes.data, indices.ctypes.data, data.ctypes.data, n, m, max(k, 0),
                          int(bool(center)), int(bool(scale)), mv, meth, int(oversample), int(n_iter),
                          int(seed), emb.ctypes.data, ratio.ctypes.data,
                          None if components is None else components.ctypes.data, int(device))
    if rc != N.PSQ_OK:
        msg = L.psq_last_error().decode("utf-8", "replace")
        raise N.PsqError(rc, f"PCA failed: {msg}")
    return {"embedding": None if out is not None else emb, "variance_ratio": ratio, "n_components": k}


def pca_sparse(matrix, n_components: int, center: bool, oversample: int, n_iter: int, seed: int,
               out=None, *, components=None, device: int = 0, method: str = "auto") -> dict:
    """lib.rs:395-436: PCA of a CSR matrix with implicit centering (``center``).
    Returns ``{"embedding": (n_obs, n_components) float32 or None when written
    into ``out``, "variance_ratio": float32[n_components], "n_components": k}``.
    ``components`` (optional, float32 (n_features, n_components)) receives the
    principal axes. ``oversample``, ``n_iter`` and ``seed`` configure the
    randomized route (more than 32,768 genes, or ``method="randomized"``)."""
    return _run(matrix, n_components, center, False, None, out, components, device, method, oversample, n_iter,
                seed)


def pca_sparse_scaled(matrix, n_components: int, max_value=None, seed: int = 0, out=None, *,
                      components=None, device: int = 0, method: str = "auto") -> dict:
    """lib.rs:438-467: PCA of the per-gene standardised matrix (``pp.scale``,
    ddof 1, a constant gene divided by 1, values clipped to +-max_value), centred,
    without densifying. The randomized route takes the reference's default
    options here (oversample 10, n_iter 1, pca.rs:73-83) and ``seed``."""
    if max_value is not None and not float(max_value) >= 0.0:
        raise ValueError("max_value must be >= 0")
    return _run(matrix, n_components, True, True, max_value, out, components, device, method, 10, 1, seed)


def pca_dense(data, n_components: int, center: bool, seed: int, *, components=None,
              device: int = 0) -> dict:
    """lib.rs:355-388: PCA of a dense float32 (n_obs, n_features) array, the call
    ``pp.pca`` makes for dense input (python/polariseq/__init__.py:2426-2434,
    2448-2453). The reference densely factorises (pca.rs:176-223, a thin SVD
    with arbitrary column signs, or randomized SVD when large); here the matrix
    goes through the same exact Gram route as the sparse calls, as the CSR of
    its nonzero entries, so the columns carry canonical signs."""
    import scipy.sparse as sp
    x = np.asarray(data, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("data must be a 2-D float32 array")
    return _run(sp.csr_matrix(x), n_components, center, False, None, None, components, device, "auto", 10, 1,
                seed)


def pca_last_report() -> dict:
    """Phase times (HIP events on the call's stream) and solver counts of this
    thread's last successful PCA call."""
    import ctypes
    r = N.PcaReport()
    N.check(N.lib().psq_pca_last_report(ctypes.byref(r)))
    return r.as_dict()


def pca_csr_device(indptr, indices, data, n_features: int, n_components: int, *, center: bool = True,
                   scale: bool = False, max_value=None, out=None, components=None, method: str = "auto",
                   oversample: int = 10, n_iter: int = 1, seed: int = 0):
    """The same computation on CUDA tensors already on the device (uint64
    indptr as int64, int32 indices, float32 data); returns (embedding,
    variance_ratio) tensors on the current stream."""
    import torch

    from .neighbors import _stream_ptr
    for t, dt in ((indptr, torch.int64), (indices, torch.int32), (data, torch.float32)):
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous() and t.dim() == 1):
            raise ValueError("indptr/indices/data must be contiguous 1-D CUDA tensors "
                             "(int64, int32, float32)")
    n = indptr.numel() - 1
    k = int(n_components)

    def _check_out(t, shape, name):
        # the C-ABI takes no buffer sizes: a wrong tensor would be written past its end
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                and tuple(t.shape) == shape and t.device == indptr.device):
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of shape {shape} "
                             "on the input's device")
    if out is None:
        out = torch.empty((n, k), dtype=torch.float32, device=indptr.device)
    else:
        _check_out(out, (n, k), "out")
    if components is not None:
        _check_out(components, (int(n_features), k), "components")
    ratio = torch.empty(k, dtype=torch.float32, device=indptr.device)
    mv = math.inf if max_value is None else float(max_value)
    _unsigned(oversample=oversample, n_iter=n_iter, seed=seed)
    L = N.lib()
    rc = L.psq_pca_csr_device_ex(indptr.data_ptr(), indices.data_ptr(), data.data_ptr(), n, int(n_features), k,
                                 int(bool(center)), int(bool(scale)), mv, _method(method), int(oversample),
                                 int(n_iter), int(seed), out.data_ptr(), ratio.data_ptr(),
                                 None if components is None else components.data_ptr(), _stream_ptr())
    if rc != N.PSQ_OK:
        msg = L.psq_last_error().decode("utf-8", "replace")
        raise N.PsqError(rc, f"PCA failed: {msg}")
    return out, ratio

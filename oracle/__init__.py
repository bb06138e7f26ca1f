"""Oracle: CPU restatements used only as checkers and as the timed CPU baseline.

Test infrastructure, not product. Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import this package. The product path
(``polariseq_amd``) never imports it and fails loudly when its HIP library is
missing.

Contents (each C function cites the reference file:line it restates):

* ``nndescent_port`` — the reference's CPU NN-Descent, step for step
  (``crates/polariseq-core/src/nndescent.rs:41-1003``). The CPU baseline, and the
  recall yardstick that the GPU graph must reach. The bit-level graph of the
  Rust build is parity unpinned, because Rust cannot be compiled here. The port is
  pinned by the reference's invariant and recall tests.
* ``knn_exact64`` — FP64 brute-force kNN ordered by (distance, index), self
  excluded (``core/doublets.rs:544-651``). Pinned by the reference KATs, by
  numpy float64 brute force and by duplicate points.
* ``neighbors_exact`` — the reference's exact path (``core/neighbors.rs:101-241``).
* ``gpu_model_nndescent`` — the GPU algorithm's documented semantics, step for
  step on the CPU. Its output must equal the HIP path's bit for bit.
* ``fuzzy_strengths`` / ``fuzzy_connectivities`` — membership strengths and the
  symmetrised fuzzy union of a kNN distance graph (``core/umap.rs:327-460``,
  ``488-537``, ``551-608``). Pinned by the umap-learn values the reference tests
  hold (``core/umap.rs:2261-2276``), by the sum-to-log2(k+1) property of the
  bandwidth search, and by an independent undirected-edge formulation (the
  reference's HashMap oracle, ``core/umap.rs:1891-1918``) in ``tests/``.
* scalar ``sq_dist`` / ``dot`` / ``splitmix`` / ``hash_unit`` / ``hash_index``
  (``core/simd.rs:22-81``, ``core/nndescent.rs:80-100``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# PSQ_ORACLE_LIB selects another build of the same sources (the sanitizer build)
_LIB_PATH = Path(os.environ.get("PSQ_ORACLE_LIB", str(_HERE / "libpsq_oracle.so")))
_lib = None


def build() -> Path:
    """Compile the oracle library with the committed Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _declare(_lib)
    return _lib


class PortStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint64),
        ("n_trees", ctypes.c_uint64),
        ("max_candidates", ctypes.c_uint64),
        ("leaf_size", ctypes.c_uint64),
        ("max_iterations", ctypes.c_uint64),
        ("stop_at", ctypes.c_uint64),
        ("seconds_forest", ctypes.c_double),
        ("seconds_iterations", ctypes.c_double),
        ("seconds_total", ctypes.c_double),
        ("changed", ctypes.c_uint64 * 64),
    ]


class ModelStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint64),
        ("n_trees", ctypes.c_uint64),
        ("max_candidates", ctypes.c_uint64),
        ("leaf_size", ctypes.c_uint64),
        ("max_iterations", ctypes.c_uint64),
        ("stop_at", ctypes.c_uint64),
        ("slots", ctypes.c_uint64),
        ("changed", ctypes.c_uint64 * 64),
        ("pairs", ctypes.c_uint64 * 64),
        ("proposals", ctypes.c_uint64 * 64),
    ]


_P = ctypes.c_void_p
_U64 = ctypes.c_uint64


def _declare(L: ctypes.CDLL) -> None:
    L.psq_o_sq_dist.restype = ctypes.c_float
    L.psq_o_sq_dist.argtypes = [_P, _P, _U64]
    L.psq_o_dot.restype = ctypes.c_float
    L.psq_o_dot.argtypes = [_P, _P, _U64]
    L.psq_o_sq_dist64.restype = ctypes.c_double
    L.psq_o_sq_dist64.argtypes = [_P, _P, _U64]
    L.psq_o_splitmix.restype = _U64
    L.psq_o_splitmix.argtypes = [_U64]
    L.psq_o_hash_unit.restype = ctypes.c_float
    L.psq_o_hash_unit.argtypes = [_U64, _U64, _U64]
    L.psq_o_hash_index.restype = _U64
    L.psq_o_hash_index.argtypes = [_U64, _U64, _U64, _U64]
    L.psq_o_effective_k.restype = _U64
    L.psq_o_effective_k.argtypes = [_U64, _U64]
    L.psq_o_knn_exact64.restype = ctypes.c_int
    L.psq_o_knn_exact64.argtypes = [_P, _U64, _U64, _P, _U64, _U64, _P, _P, ctypes.c_int]
    L.psq_o_neighbors_exact.restype = ctypes.c_int64
    L.psq_o_neighbors_exact.argtypes = [_P, _U64, _U64, _U64, ctypes.c_int, _P, _P, _P,
                                        ctypes.c_int]
    L.psq_o_nndescent_port.restype = ctypes.c_int64
    L.psq_o_nndescent_port.argtypes = [_P, _U64, _U64, _U64, _U64, _U64, ctypes.c_double,
                                       _U64, _U64, _U64, _U64, ctypes.c_int, _P, _P, _P,
                                       ctypes.POINTER(PortStats)]
    L.psq_o_gpu_model_nndescent.restype = ctypes.c_int64
    L.psq_o_gpu_model_nndescent.argtypes = [_P, _U64, _U64, _U64, _U64, _U64, ctypes.c_double,
                                            _U64, _U64, _U64, _U64, _U64, _U64, _P, ctypes.c_int, _P,
                                            _P, _P, ctypes.POINTER(ModelStats)]
    L.psq_o_fuzzy_strengths.restype = None
    L.psq_o_fuzzy_strengths.argtypes = [_P, _P, _U64, _P]
    L.psq_o_fuzzy_connectivities.restype = ctypes.c_int64
    L.psq_o_fuzzy_connectivities.argtypes = [_P, _P, _P, _U64, _P, _P, _P, _U64,
                                             ctypes.POINTER(_U64)]


def _f32(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float32)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def default_threads() -> int:
    return max(1, os.cpu_count() or 1)


def sq_dist(a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    assert a.shape == b.shape and a.ndim == 1
    return np.float32(lib().psq_o_sq_dist(_ptr(a), _ptr(b), a.size))


def dot(a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    assert a.shape == b.shape and a.ndim == 1
    return np.float32(lib().psq_o_dot(_ptr(a), _ptr(b), a.size))


def sq_dist64(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return float(lib().psq_o_sq_dist64(_ptr(a), _ptr(b), a.size))


def splitmix(x: int) -> int:
    return int(lib().psq_o_splitmix(x))


def hash_unit(seed: int, a: int, b: int) -> float:
    return float(lib().psq_o_hash_unit(seed, a, b))


def hash_index(seed: int, a: int, b: int, n: int) -> int:
    return int(lib().psq_o_hash_index(seed, a, b, n))


def effective_k(n_obs: int, k: int) -> int:
    return int(lib().psq_o_effective_k(n_obs, k))


def knn_exact64(data: np.ndarray, k: int, queries: np.ndarray | None = None,
                threads: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(idx u32[n_q,k], squared FP64 distances f64[n_q,k]), order (d64, idx), self excluded."""
    data = _f32(data)
    n, d = data.shape
    if queries is None:
        q = None
        n_q = n
    else:
        q = np.ascontiguousarray(queries, dtype=np.uint32)
        n_q = q.size
    idx = np.empty((n_q, k), np.uint32)
    d2 = np.empty((n_q, k), np.float64)
    rc = lib().psq_o_knn_exact64(_ptr(data), n, d, None if q is None else _ptr(q), n_q, k,
                                 _ptr(idx), _ptr(d2), threads or default_threads())
    if rc != 0:
        raise ValueError("knn_exact64: empty input or k == 0")
    return idx, d2


def neighbors_exact(emb: np.ndarray, n_neighbors: int, metric: str = "euclidean",
                    threads: int | None = None) -> dict:
    emb = _f32(emb)
    n, d = emb.shape
    kk = min(n_neighbors, n)
    cap = max(0, n * (kk - 1))
    ip = np.empty(n + 1, np.uint32)
    ix = np.empty(max(cap, 1), np.uint32)
    dx = np.empty(max(cap, 1), np.float32)
    m = {"euclidean": 0, "cosine": 1}[metric]
    nnz = lib().psq_o_neighbors_exact(_ptr(emb), n, d, n_neighbors, m, _ptr(ip), _ptr(ix),
                                      _ptr(dx), threads or default_threads())
    if nnz < 0:
        raise RuntimeError("neighbors failed")
    return {"indptr": ip, "indices": ix[:nnz], "data": dx[:nnz], "n_obs": n}


def nndescent_port(data: np.ndarray, k: int, max_iterations: int = 0, seed: int = 0,
                   max_candidates: int = 0, n_trees: int = 0, delta: float = 0.001,
                   leaf_size: int = 0, gather_max: int = 0, threads: int | None = None):
    """The reference CPU NN-Descent restated; returns (graph dict, PortStats)."""
    data = _f32(data)
    n, d = data.shape
    if n == 0 or d == 0 or k == 0:
        raise ValueError("empty data or k == 0")
    ke = effective_k(n, k)
    ip = np.empty(n + 1, np.uint32)
    ix = np.empty(n * ke, np.uint32)
    dx = np.empty(n * ke, np.float32)
    st = PortStats()
    nnz = lib().psq_o_nndescent_port(_ptr(data), n, d, k, max_candidates, max_iterations, delta,
                                     n_trees, leaf_size, gather_max, seed,
                                     threads or default_threads(), _ptr(ip), _ptr(ix), _ptr(dx),
                                     ctypes.byref(st))
    if nnz < 0:
        raise ValueError("invalid input")
    return {"indptr": ip, "indices": ix[:nnz], "data": dx[:nnz], "n_obs": n, "nnz": nnz}, st


def gpu_model_nndescent(data: np.ndarray, k: int, max_iterations: int = 0, seed: int = 0,
                        max_candidates: int = 0, n_trees: int = 0, delta: float = 0.001,
                        leaf_size: int = 0, slots: int = 0, init: int = 0,
                        init_graph: np.ndarray | None = None, threads: int | None = None,
                        split_data: np.ndarray | None = None, heap_k: int = 0):
    """The GPU algorithm's semantics on the CPU; returns (graph dict, ModelStats).

    ``init``: 0 = RP forest, 1 = random fill only, 2 = seed rows from ``init_graph``
    (u32[n, c] candidate indices, 0xFFFFFFFF = none), 3 = RP forest splitting the
    rows of ``split_data`` (f32[n, ds], e.g. learned codes), leaf pairs measured on
    ``data``. ``heap_k``: entries each row keeps while iterating (0 = k); the
    output holds each row's first k.
    """
    data = _f32(data)
    n, d = data.shape
    if n == 0 or d == 0 or k == 0:
        raise ValueError("empty data or k == 0")
    ke = effective_k(n, k)
    ip = np.empty(n + 1, np.uint32)
    ix = np.empty(n * ke, np.uint32)
    dx = np.empty(n * ke, np.float32)
    st = ModelStats()
    ig = None
    ig_cols = 0
    if init_graph is not None:
        ig = np.ascontiguousarray(init_graph, dtype=np.uint32)
        ig_cols = ig.shape[1]
    if split_data is not None:
        ig = np.ascontiguousarray(split_data, dtype=np.float32)
        ig_cols = ig.shape[1]
        init = 3
    nnz = lib().psq_o_gpu_model_nndescent(
        _ptr(data), n, d, k, max_candidates, max_iterations, delta, n_trees, leaf_size, slots,
        heap_k, seed, (init & 0xFF) | (ig_cols << 8), None if ig is None else _ptr(ig),
        threads or default_threads(), _ptr(ip), _ptr(ix), _ptr(dx), ctypes.byref(st))
    if nnz < 0:
        raise ValueError("invalid input")
    return {"indptr": ip, "indices": ix[:nnz], "data": dx[:nnz], "n_obs": n, "nnz": nnz}, st


def _u32(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.uint32)


def fuzzy_strengths(indptr: np.ndarray, dists: np.ndarray) -> np.ndarray:
    """Membership strengths of every entry of a kNN distance graph
    (``core/umap.rs:551-608`` applied row by row, ``core/umap.rs:343-354``)."""
    ip = _u32(indptr)
    dx = _f32(dists)
    n = len(ip) - 1
    out = np.zeros(max(int(ip[-1]), 1), np.float32)
    lib().psq_o_fuzzy_strengths(_ptr(ip), _ptr(dx), n, _ptr(out))
    return out[: int(ip[-1])]


def fuzzy_connectivities(indptr: np.ndarray, indices: np.ndarray, dists: np.ndarray,
                         n_obs: int) -> dict:
    """The symmetrised fuzzy connectivities (``core/umap.rs:130-139``): a dict
    with ``indptr``, ``indices`` and ``data`` (the weights)."""
    ip = _u32(indptr)
    ix = _u32(indices)
    dx = _f32(dists)
    if len(ip) != n_obs + 1:
        raise ValueError("indptr must have n_obs + 1 entries")
    if len(ix) and int(ix.max()) >= n_obs:
        raise ValueError("index out of range")
    cap = 2 * int(ip[-1])
    out_ip = np.zeros(n_obs + 1, np.uint32)
    out_ix = np.zeros(max(cap, 1), np.uint32)
    out_w = np.zeros(max(cap, 1), np.float32)
    need = _U64(0)
    nnz = lib().psq_o_fuzzy_connectivities(_ptr(ip), _ptr(ix), _ptr(dx), n_obs, _ptr(out_ip),
                                           _ptr(out_ix), _ptr(out_w), cap, ctypes.byref(need))
    if nnz < 0:
        raise ValueError("output capacity too small")
    return {"indptr": out_ip, "indices": out_ix[:nnz], "data": out_w[:nnz]}

"""ctypes binding of libpsq_knn.so (declarations in include/psq_knn.h).

There is no fallback: if the HIP library is missing or fails to load, every call
raises. Build it with ``__graft_entry__.build()`` (or ``make -C polariseq_amd/csrc``).
"""

from __future__ import annotations

import ctypes
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libpsq_knn.so"

PSQ_OK = 0
PSQ_E_INVALID = -1
PSQ_E_UNSUPPORTED = -2
PSQ_E_DEVICE = -3
PSQ_E_NOMEM = -4
PSQ_PCA_AUTO = 0
PSQ_PCA_GRAM = 1
PSQ_PCA_RANDOMIZED = 2

PRECISION = {"fp32": 0, "bf16": 1, "bf16_rerank": 1}
INIT = {"rp_forest": 0, "random": 1, "learned": 2}

# Every symbol include/psq_knn.h declares (checked by a CPU test).
EXPORTS = (
    "psq_nndescent_opts_default",
    "psq_effective_k",
    "psq_nndescent_into",
    "psq_nndescent_device",
    "psq_knn_exact",
    "psq_knn_exact_device",
    "psq_knn_exact_device_noalloc",
    "psq_neighbors_exact",
    "psq_neighbors_exact_device",
    "psq_last_stats",
    "psq_last_error",
    "psq_release_workspace",
    "psq_set_device_budget",
    "psq_encoder_create",
    "psq_encoder_destroy",
    "psq_encoder_reset",
    "psq_encoder_train",
    "psq_encoder_encode",
    "psq_encoder_params",
    "psq_encoder_set_params",
    "psq_encoder_grad",
    "psq_encoder_info",
    "psq_encoder_mine_ms",
    "psq_knn_exact_batched",
    "psq_knn_exact_batched_device",
    "psq_batched_last_report",
    "psq_fuzzy_connectivities_device",
    "psq_fuzzy_plan_new",
    "psq_fuzzy_plan_fill",
    "psq_fuzzy_plan_free",
    "psq_unit_rows_device",
    "psq_cosine_distances_device",
    "psq_pca_csr",
    "psq_pca_csr_device",
    "psq_pca_csr_ex",
    "psq_pca_csr_device_ex",
    "psq_pca_last_report",
)


class BatchedReport(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_uint32), ("tile", ctypes.c_uint32), ("seg_len", ctypes.c_uint32),
        ("tuned_now", ctypes.c_uint32), ("configs_tried", ctypes.c_uint32),
        ("units", ctypes.c_uint32), ("overflow_queries", ctypes.c_uint32), ("pad", ctypes.c_uint32),
        ("tune_ms", ctypes.c_double), ("ms_pass1", ctypes.c_double), ("ms_pass2", ctypes.c_double),
        ("ms_rerank", ctypes.c_double), ("ms_total", ctypes.c_double), ("pairs", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class PcaReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint64), ("block", ctypes.c_uint64), ("chunks", ctypes.c_uint64),
        ("rows_per_chunk", ctypes.c_uint64), ("ms_scale", ctypes.c_double),
        ("ms_gram", ctypes.c_double), ("ms_gemm", ctypes.c_double), ("ms_eigen", ctypes.c_double),
        ("ms_project", ctypes.c_double), ("ms_total", ctypes.c_double), ("method", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["method"] = {PSQ_PCA_GRAM: "gram", PSQ_PCA_RANDOMIZED: "randomized"}.get(d["method"], d["method"])
        return d


class EncoderTrainOpts(ctypes.Structure):
    _fields_ = [
        ("subset", ctypes.c_uint64),
        ("k_pos", ctypes.c_uint64),
        ("batch", ctypes.c_uint64),
        ("steps", ctypes.c_uint64),
        ("lr", ctypes.c_float),
        ("margin", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("anchors", ctypes.c_uint64),
    ]


class NndescentOpts(ctypes.Structure):
    """psq_nndescent_opts, mirroring NndescentOptions (nndescent.rs:41-76)."""

    _fields_ = [
        ("k", ctypes.c_uint64),
        ("max_candidates", ctypes.c_uint64),
        ("max_iterations", ctypes.c_uint64),
        ("delta", ctypes.c_double),
        ("n_trees", ctypes.c_uint64),
        ("leaf_size", ctypes.c_uint64),
        ("gather_max", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("init", ctypes.c_int32),
        ("model", ctypes.c_void_p),
        ("device", ctypes.c_int32),
        ("slots", ctypes.c_uint32),
        ("heap_k", ctypes.c_uint32),
        ("reorder", ctypes.c_uint32),
    ]


class NndescentStats(ctypes.Structure):
    _fields_ = [
        ("n_obs", ctypes.c_uint64),
        ("n_dim", ctypes.c_uint64),
        ("k", ctypes.c_uint64),
        ("max_candidates", ctypes.c_uint64),
        ("max_iterations", ctypes.c_uint64),
        ("n_trees", ctypes.c_uint64),
        ("leaf_size", ctypes.c_uint64),
        ("slots", ctypes.c_uint64),
        ("stop_at", ctypes.c_uint64),
        ("iterations", ctypes.c_uint64),
        ("heap_k", ctypes.c_uint64),
        ("reordered", ctypes.c_uint64),
        ("changed", ctypes.c_uint64 * 64),
        ("pairs", ctypes.c_uint64 * 64),
        ("cand_rows", ctypes.c_uint64 * 64),
        ("exact_pairs", ctypes.c_uint64 * 64),
        ("proposals", ctypes.c_uint64 * 64),
        ("ms_init", ctypes.c_double),
        ("ms_iterations", ctypes.c_double),
        ("ms_extract", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_join", ctypes.c_double * 64),
    ]

    def as_dict(self) -> dict:
        it = int(self.iterations)
        return {
            "n_obs": int(self.n_obs), "n_dim": int(self.n_dim), "k": int(self.k),
            "max_candidates": int(self.max_candidates), "max_iterations": int(self.max_iterations),
            "n_trees": int(self.n_trees), "leaf_size": int(self.leaf_size),
            "slots": int(self.slots), "stop_at": int(self.stop_at), "iterations": it,
            "heap_k": int(self.heap_k), "reordered": int(self.reordered),
            "changed": [int(v) for v in self.changed[:it]],
            "pairs": [int(v) for v in self.pairs[:it]],
            "cand_rows": [int(v) for v in self.cand_rows[:it]],
            "exact_pairs": [int(v) for v in self.exact_pairs[:it]],
            "proposals": [int(v) for v in self.proposals[:it]],
            "ms_init": float(self.ms_init), "ms_iterations": float(self.ms_iterations),
            "ms_extract": float(self.ms_extract), "ms_total": float(self.ms_total),
            "ms_join": [float(v) for v in self.ms_join[:it]],
        }


_lib = None


def lib() -> ctypes.CDLL:
    """Load libpsq_knn.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() "
                "or make -C polariseq_amd/csrc")
        L = ctypes.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    P, U64, I32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32
    L.psq_nndescent_opts_default.restype = None
    L.psq_nndescent_opts_default.argtypes = [ctypes.POINTER(NndescentOpts)]
    L.psq_effective_k.restype = U64
    L.psq_effective_k.argtypes = [U64, U64]
    L.psq_nndescent_into.restype = ctypes.c_int
    L.psq_nndescent_into.argtypes = [P, U64, U64, ctypes.POINTER(NndescentOpts), P, U64, P, P, U64,
                                     ctypes.POINTER(U64)]
    L.psq_nndescent_device.restype = ctypes.c_int
    L.psq_nndescent_device.argtypes = [P, U64, U64, ctypes.POINTER(NndescentOpts), P, U64, P, P,
                                       U64, ctypes.POINTER(U64), P]
    L.psq_knn_exact.restype = ctypes.c_int
    L.psq_knn_exact.argtypes = [P, U64, U64, P, U64, U64, P, P, I32]
    L.psq_knn_exact_device.restype = ctypes.c_int
    L.psq_knn_exact_device.argtypes = [P, U64, U64, P, U64, U64, P, P, P]
    L.psq_knn_exact_device_noalloc.restype = ctypes.c_int
    L.psq_knn_exact_device_noalloc.argtypes = [P, U64, U64, P, U64, U64, P, P, P]
    L.psq_neighbors_exact.restype = ctypes.c_int
    L.psq_neighbors_exact.argtypes = [P, U64, U64, U64, I32, P, P, P, U64, ctypes.POINTER(U64), I32]
    L.psq_neighbors_exact_device.restype = ctypes.c_int
    L.psq_neighbors_exact_device.argtypes = [P, U64, U64, U64, I32, P, P, P, U64,
                                             ctypes.POINTER(U64), P]
    L.psq_last_stats.restype = ctypes.c_int
    L.psq_last_stats.argtypes = [ctypes.POINTER(NndescentStats)]
    L.psq_last_error.restype = ctypes.c_char_p
    L.psq_last_error.argtypes = []
    L.psq_release_workspace.restype = ctypes.c_int
    L.psq_release_workspace.argtypes = []
    L.psq_set_device_budget.restype = None
    L.psq_set_device_budget.argtypes = [U64]
    F32, PP = ctypes.c_float, ctypes.POINTER(ctypes.c_void_p)
    L.psq_encoder_create.restype = ctypes.c_int
    L.psq_encoder_create.argtypes = [PP, U64, I32]
    L.psq_encoder_destroy.restype = None
    L.psq_encoder_destroy.argtypes = [P]
    L.psq_encoder_reset.restype = ctypes.c_int
    L.psq_encoder_reset.argtypes = [P, U64]
    L.psq_encoder_train.restype = ctypes.c_int
    L.psq_encoder_train.argtypes = [P, P, U64, U64, ctypes.POINTER(EncoderTrainOpts), P]
    L.psq_encoder_encode.restype = ctypes.c_int
    L.psq_encoder_encode.argtypes = [P, P, U64, U64, P, P]
    L.psq_encoder_params.restype = ctypes.c_int
    L.psq_encoder_params.argtypes = [P, P, U64, P]
    L.psq_encoder_set_params.restype = ctypes.c_int
    L.psq_encoder_set_params.argtypes = [P, P, U64, F32, F32]
    L.psq_encoder_grad.restype = ctypes.c_int
    L.psq_encoder_grad.argtypes = [P, P, U64, U64, P, U64, F32, P, ctypes.POINTER(ctypes.c_double)]
    L.psq_knn_exact_batched.restype = ctypes.c_int
    L.psq_knn_exact_batched.argtypes = [P, U64, U64, P, U64, U64, P, P, I32]
    L.psq_knn_exact_batched_device.restype = ctypes.c_int
    L.psq_knn_exact_batched_device.argtypes = [P, U64, U64, P, U64, U64, P, P, P]
    L.psq_batched_last_report.restype = ctypes.c_int
    L.psq_batched_last_report.argtypes = [ctypes.POINTER(BatchedReport)]
    L.psq_fuzzy_connectivities_device.restype = ctypes.c_int
    L.psq_fuzzy_connectivities_device.argtypes = [P, P, P, U64, P, P, P, U64, ctypes.POINTER(U64), P]
    L.psq_fuzzy_plan_new.restype = ctypes.c_int
    L.psq_fuzzy_plan_new.argtypes = [P, U64, P, P, U64, U64, I32, PP, ctypes.POINTER(U64)]
    L.psq_fuzzy_plan_fill.restype = ctypes.c_int
    L.psq_fuzzy_plan_fill.argtypes = [P, P, U64, P, P, U64]
    L.psq_fuzzy_plan_free.restype = None
    L.psq_fuzzy_plan_free.argtypes = [P]
    L.psq_unit_rows_device.restype = ctypes.c_int
    L.psq_unit_rows_device.argtypes = [P, U64, U64, P, P]
    L.psq_cosine_distances_device.restype = ctypes.c_int
    L.psq_cosine_distances_device.argtypes = [P, U64, P]
    D = ctypes.c_double
    L.psq_pca_csr.restype = ctypes.c_int
    L.psq_pca_csr.argtypes = [P, P, P, U64, U64, U64, I32, I32, D, P, P, P, I32]
    L.psq_pca_csr_device.restype = ctypes.c_int
    L.psq_pca_csr_device.argtypes = [P, P, P, U64, U64, U64, I32, I32, D, P, P, P, P]
    L.psq_pca_csr_ex.restype = ctypes.c_int
    L.psq_pca_csr_ex.argtypes = [P, P, P, U64, U64, U64, I32, I32, D, I32, U64, U64, U64, P, P, P, I32]
    L.psq_pca_csr_device_ex.restype = ctypes.c_int
    L.psq_pca_csr_device_ex.argtypes = [P, P, P, U64, U64, U64, I32, I32, D, I32, U64, U64, U64, P, P, P, P]
    L.psq_pca_last_report.restype = ctypes.c_int
    L.psq_pca_last_report.argtypes = [ctypes.POINTER(PcaReport)]
    L.psq_encoder_mine_ms.restype = ctypes.c_int
    L.psq_encoder_mine_ms.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    L.psq_encoder_info.restype = ctypes.c_int
    L.psq_encoder_info.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(U64)]


class PsqError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def check(rc: int) -> None:
    """Raise for a non-zero status: ValueError for invalid arguments (the
    reference binding's class for bad outputs, pybind/lib.rs:546-552),
    PsqError (a RuntimeError) otherwise."""
    if rc == PSQ_OK:
        return
    msg = lib().psq_last_error().decode("utf-8", "replace")
    if rc == PSQ_E_INVALID:
        raise ValueError(msg)
    if rc == PSQ_E_UNSUPPORTED:
        raise PsqError(rc, f"unsupported: {msg}")
    raise PsqError(rc, msg)


def make_opts(k: int, max_iterations: int = 0, seed: int = 0, max_candidates: int = 0,
              n_trees: int = 0, delta: float = 0.001, leaf_size: int = 0,
              precision: str = "fp32", init: str = "rp_forest", device: int = 0,
              slots: int = 0, model=None, heap_k: int = 0, reorder: str = "auto") -> NndescentOpts:
    o = NndescentOpts()
    lib().psq_nndescent_opts_default(ctypes.byref(o))
    o.k = int(k)
    o.max_iterations = int(max_iterations)
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.max_candidates = int(max_candidates)
    o.n_trees = int(n_trees)
    o.delta = float(delta)
    o.leaf_size = int(leaf_size)
    o.precision = PRECISION[precision]
    o.init = INIT[init]
    o.device = int(device)
    o.slots = int(slots)
    o.heap_k = int(heap_k)
    o.reorder = {"auto": 0, "on": 1, "off": 2}[reorder]
    o.model = model
    return o


def last_stats() -> dict:
    st = NndescentStats()
    check(lib().psq_last_stats(ctypes.byref(st)))
    return st.as_dict()

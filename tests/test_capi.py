"""CPU tests of the C-ABI boundary: the library loads, exports every symbol the
header declares, and turns bad input into error codes (never an abort) before it
touches a device — the errors the reference raises as panics
(crates/polariseq-core/src/nndescent.rs:708-719) or as ValueError
(crates/polariseq-python/src/lib.rs:546-552)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from polariseq_amd import _native as N
from polariseq_amd import neighbors as nb

ROOT = Path(__file__).resolve().parents[1]


def _header_functions() -> set[str]:
    text = (ROOT / "include" / "psq_knn.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(psq_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = _header_functions()
    assert declared == set(N.EXPORTS), declared ^ set(N.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_effective_k_matches_the_reference_rule(orc):
    for n in (1, 2, 3, 5, 16, 1000):
        for k in (1, 2, 4, 10, 15, 30, 999):
            want = max(1, min(k, n - 1))
            assert nb.effective_k(n, k) == want == orc.effective_k(n, k)


def test_opts_defaults_mirror_nndescent_options():
    o = N.NndescentOpts()
    N.lib().psq_nndescent_opts_default(ctypes.byref(o))
    assert (o.k, o.max_candidates, o.max_iterations, o.n_trees, o.leaf_size, o.gather_max,
            o.seed) == (15, 0, 0, 0, 0, 0, 0)
    assert o.delta == 0.001 and o.precision == 0 and o.init == 0


def _call_into(x, k, ip_len=None, cap=None, **kw):
    n, d = x.shape if x.size else (x.shape[0], x.shape[1] if x.ndim == 2 else 0)
    ke = max(1, min(k, max(n - 1, 0)))
    ip = np.zeros((n + 1) if ip_len is None else ip_len, np.uint32)
    c = n * ke if cap is None else cap
    ix = np.zeros(max(c, 1), np.uint32)
    dx = np.zeros(max(c, 1), np.float32)
    nnz = ctypes.c_uint64(0)
    o = N.make_opts(k, **kw)
    rc = N.lib().psq_nndescent_into(x.ctypes.data if x.size else None, n, d, ctypes.byref(o),
                                    ip.ctypes.data, ip.size, ix.ctypes.data, dx.ctypes.data, c,
                                    ctypes.byref(nnz))
    return rc, N.lib().psq_last_error().decode()


def test_invalid_inputs_return_error_codes_before_any_device_work():
    x = np.zeros((10, 4), np.float32)
    rc, msg = _call_into(np.zeros((0, 4), np.float32), 5)
    assert rc == N.PSQ_E_INVALID and "empty" in msg
    rc, msg = _call_into(x, 0)
    assert rc == N.PSQ_E_INVALID and "k must be" in msg
    rc, msg = _call_into(x, 5, ip_len=10)
    assert rc == N.PSQ_E_INVALID and "indptr" in msg
    rc, msg = _call_into(x, 5, cap=10)
    assert rc == N.PSQ_E_INVALID and "outputs must hold" in msg
    rc, msg = _call_into(x, 5, slots=48)
    assert rc == N.PSQ_E_INVALID and "slots" in msg


def test_unsupported_options_are_reported_not_ignored():
    x = np.zeros((400, 4), np.float32)
    rc, msg = _call_into(x, 300)
    assert rc == N.PSQ_E_UNSUPPORTED and "heap_k > 256" in msg
    rc, msg = _call_into(np.zeros((10, 300), np.float32), 5)
    assert rc == N.PSQ_E_UNSUPPORTED and "n_dim" in msg


def test_python_binding_raises_the_reference_value_error():
    x = np.zeros((10, 4), np.float32)
    ip = np.zeros(11, np.uint32)
    ix = np.zeros(20, np.uint32)
    dx = np.zeros(20, np.float32)
    with pytest.raises(ValueError, match=r"outputs must hold 11 row pointers and 50 neighbours"):
        nb.nndescent_dense(x, 5, out=(ip, ix, dx))
    with pytest.raises(ValueError, match="k must be"):
        nb.nndescent_dense(x, 0)


def test_exact_path_errors_mirror_the_reference():
    with pytest.raises(RuntimeError, match="neighbors failed"):
        nb.neighbors_dense(np.zeros((0, 3), np.float32), 5)
    with pytest.raises(RuntimeError, match="neighbors failed"):
        nb.neighbors_dense(np.zeros((4, 3), np.float32), 0)
    with pytest.raises(RuntimeError, match="unknown metric"):
        nb.neighbors_dense(np.zeros((4, 3), np.float32), 2, "manhattan")
    with pytest.raises(ValueError):
        nb.knn_exact(np.zeros((4, 3), np.float32), 0)
    with pytest.raises(ValueError, match="out of range"):
        nb.knn_exact(np.zeros((4, 3), np.float32), 2, queries=np.array([9], np.uint32))


def test_valid_call_without_a_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    rc, msg = _call_into(np.random.default_rng(0).random((50, 4), dtype=np.float32), 5,
                         init="random")
    assert rc in (N.PSQ_E_DEVICE, N.PSQ_E_INVALID) and msg

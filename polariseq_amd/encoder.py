"""Learned seeding encoder: a 50 -> 128 -> 32 MLP (GELU, BF16-rounded weights,
FP32 accumulation) trained on the GPU by hand-written forward and backward
kernels (polariseq_amd/csrc/mlp_kernels.hip) on triplets from the exact kNN of
a subset of the cells. Its 32-D codes are what the RP forest splits when
NN-Descent runs with init="learned"; leaf pairs are still measured exactly in
the original space. The reference has no learned component; this seeds the
start that crates/polariseq-core/src/nndescent.rs:767-817 builds from the raw
embedding."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

N_PARAMS = 128 * 50 + 128 + 32 * 128 + 32


class Encoder:
    """Owns the device parameters of one encoder."""

    def __init__(self, seed: int = 0, device: int = 0):
        h = ctypes.c_void_p()
        N.check(N.lib().psq_encoder_create(ctypes.byref(h), int(seed), int(device)))
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def close(self) -> None:
        if self._h:
            N.lib().psq_encoder_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, seed: int = 0) -> None:
        """Back to the untrained weights of ``Encoder(seed)`` (buffers kept)."""
        N.check(N.lib().psq_encoder_reset(self._h, int(seed)))

    def train(self, x, *, subset: int = 200_000, k_pos: int = 15, batch: int = 8192,
              steps: int = 100, lr: float = 3e-3, margin: float = 0.01, seed: int = 0,
              anchors: int = 4096) -> dict:
        """Train on a CUDA float32 tensor (n, 50) resident on the device: triplets
        from the exact top-k_pos (within an evenly spaced subset) of an anchor pool."""
        import torch
        _check_dev(x)
        o = N.EncoderTrainOpts(subset, k_pos, batch, steps, lr, margin, seed, anchors)
        N.check(N.lib().psq_encoder_train(self._h, x.data_ptr(), x.shape[0], x.shape[1],
                                          ctypes.byref(o),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return self.info()

    def info(self) -> dict:
        loss, ms, steps = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        N.check(N.lib().psq_encoder_info(self._h, ctypes.byref(loss), ctypes.byref(ms),
                                         ctypes.byref(steps)))
        mine = ctypes.c_double()
        N.check(N.lib().psq_encoder_mine_ms(self._h, ctypes.byref(mine)))
        return {"last_loss": loss.value, "train_ms": ms.value, "mine_ms": mine.value,
                "steps": int(steps.value)}

    def encode(self, x):
        """Codes (n, 32) float32 on the device for a CUDA tensor (n, 50)."""
        import torch
        _check_dev(x)
        z = torch.empty((x.shape[0], 32), dtype=torch.float32, device=x.device)
        N.check(N.lib().psq_encoder_encode(self._h, x.data_ptr(), x.shape[0], x.shape[1],
                                           z.data_ptr(),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return z

    def params(self) -> tuple[np.ndarray, float, float]:
        p = np.empty(N_PARAMS, np.float32)
        ms = np.empty(2, np.float32)
        N.check(N.lib().psq_encoder_params(self._h, p.ctypes.data, N_PARAMS, ms.ctypes.data))
        return p, float(ms[0]), float(ms[1])

    def set_params(self, p: np.ndarray, mu: float = 0.0, inv_scale: float = 1.0) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().psq_encoder_set_params(self._h, p.ctypes.data, N_PARAMS, mu, inv_scale))

    def grad(self, x, rows: np.ndarray, margin: float = 1.0) -> tuple[np.ndarray, float]:
        """Loss and parameter gradient of one triplet batch; rows = [anchors|positives|negatives]."""
        _check_dev(x)
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        B = rows.size // 3
        g = np.empty(N_PARAMS, np.float32)
        loss = ctypes.c_double()
        N.check(N.lib().psq_encoder_grad(self._h, x.data_ptr(), x.shape[0], x.shape[1],
                                         rows.ctypes.data, B, margin, g.ctypes.data,
                                         ctypes.byref(loss)))
        return g, loss.value


def _check_dev(x) -> None:
    import torch
    if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2 and x.is_contiguous()):
        raise ValueError("x must be a contiguous float32 CUDA tensor of shape (n, 50)")

"""Oracle (test infrastructure only): FP64 restatement of the learned encoder.

Mirrors polariseq_amd/csrc/mlp_kernels.hip: parameters are rounded to BF16
(round to nearest even) before every product, inputs are standardised as
(x - mu) * inv_scale in float32, and everything after that is float64:

    h_pre = W1 s + b1,  h = gelu(h_pre) = h_pre * Phi(h_pre),  z = W2 h + b2
    L = mean_b max(0, |z_a - z_p|^2 - |z_a - z_n|^2 + margin)

The gradient is derived by hand below and pinned by central finite differences
on the same FP64 function (tests/test_mlp_oracle.py).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, H, OUT = 50, 128, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + OUT * H
N_PARAMS = OFF_B2 + OUT


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 value (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(H, IN)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(OUT, H)
    b2 = p[OFF_B2:N_PARAMS]
    return W1, b1, W2, b2


def standardise(x: np.ndarray, mu: float, inv_scale: float) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    return ((x - np.float32(mu)) * np.float32(inv_scale)).astype(np.float64)


def gelu(u):
    return 0.5 * u * (1.0 + erf(u / np.sqrt(2.0)))


def gelu_grad(u):
    return 0.5 * (1.0 + erf(u / np.sqrt(2.0))) + u * np.exp(-0.5 * u * u) / np.sqrt(2.0 * np.pi)


def forward(p: np.ndarray, x: np.ndarray, mu: float, inv_scale: float, round_bf16: bool = True):
    q = bf16_round(p) if round_bf16 else np.asarray(p, np.float32)
    W1, b1, W2, b2 = unpack(q)
    s = standardise(x, mu, inv_scale)
    pre = s @ W1.T + b1
    h = gelu(pre)
    z = h @ W2.T + b2
    return s, pre, h, z


def loss_and_grad(p: np.ndarray, xa, xp, xn, mu: float, inv_scale: float, margin: float,
                  round_bf16: bool = True):
    """(loss, dL/dparams, per-triplet hinge terms). The gradient is taken with
    respect to the rounded parameters (the straight-through convention the GPU
    uses: the FP32 masters receive it)."""
    B = xa.shape[0]
    x = np.concatenate([xa, xp, xn], axis=0)
    q = bf16_round(p) if round_bf16 else np.asarray(p, np.float32)
    W1, b1, W2, b2 = unpack(q)
    s, pre, h, z = forward(p, x, mu, inv_scale, round_bf16)
    za, zp, zn = z[:B], z[B:2 * B], z[2 * B:]
    l = ((za - zp) ** 2).sum(1) - ((za - zn) ** 2).sum(1) + margin
    act = (l > 0).astype(np.float64)
    L = float((l * act).sum() / B)
    sc = (2.0 / B) * act[:, None]
    dz = np.concatenate([sc * (zn - zp), -sc * (za - zp), sc * (za - zn)], axis=0)
    dW2 = dz.T @ h
    db2 = dz.sum(0)
    dh = dz @ W2
    dpre = dh * gelu_grad(pre)
    dW1 = dpre.T @ s
    db1 = dpre.sum(0)
    g = np.zeros(N_PARAMS)
    g[OFF_W1:OFF_B1] = dW1.ravel()
    g[OFF_B1:OFF_W2] = db1
    g[OFF_W2:OFF_B2] = dW2.ravel()
    g[OFF_B2:] = db2
    return L, g, l


def loss_only(p64: np.ndarray, xa, xp, xn, mu, inv_scale, margin) -> float:
    """FP64 loss of unrounded float64 parameters (for finite differences)."""
    W1, b1, W2, b2 = unpack(p64)
    x = np.concatenate([xa, xp, xn], axis=0)
    s = standardise(x, mu, inv_scale)
    z = gelu(s @ W1.T + b1) @ W2.T + b2
    B = xa.shape[0]
    za, zp, zn = z[:B], z[B:2 * B], z[2 * B:]
    l = ((za - zp) ** 2).sum(1) - ((za - zn) ** 2).sum(1) + margin
    return float(np.maximum(l, 0).sum() / B)

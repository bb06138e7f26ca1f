"""CPU pinning of the encoder oracle: the hand-derived FP64 gradient against
central finite differences, and the BF16 rounding against known values."""

from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.0078125, 3.14159265, -2.5, 1e-3, 65504.0], np.float32)
    # 1.00390625 = 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.0
    want = np.array([1.0, 1.0, 1.0078125, 3.140625, -2.5, 0.00099945068359375, 65536.0],
                    np.float32)
    assert np.array_equal(M.bf16_round(x), want)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(scale=0.2, size=M.H * M.IN)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(scale=0.1, size=M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.normal(scale=0.2, size=M.OUT * M.H)
    p[M.OFF_B2:] = rng.normal(scale=0.1, size=M.OUT)
    B = 6
    xa, xp, xn = (rng.normal(size=(B, M.IN)).astype(np.float32) for _ in range(3))
    L, g, l = M.loss_and_grad(p, xa, xp, xn, 0.0, 1.0, 2.0, round_bf16=False)
    assert abs(L - M.loss_only(p, xa, xp, xn, 0.0, 1.0, 2.0)) < 1e-9 * max(1.0, L)
    assert (np.abs(l) > 1e-3).all()  # away from the hinge, so the loss is smooth here
    idx = np.concatenate([rng.choice(M.OFF_B1, 40, replace=False),
                          M.OFF_B1 + rng.choice(M.H, 10, replace=False),
                          M.OFF_W2 + rng.choice(M.OUT * M.H, 40, replace=False),
                          M.OFF_B2 + rng.choice(M.OUT, 8, replace=False)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss_only(p + e, xa, xp, xn, 0.0, 1.0, 2.0)
              - M.loss_only(p - e, xa, xp, xn, 0.0, 1.0, 2.0)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])

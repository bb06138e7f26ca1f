"""GPU checks of the learned encoder and the learned seeding.

* The hand-written HIP backward agrees with the FP64 oracle gradient
  (oracle/mlp_oracle.py, itself pinned by finite differences).
* Training lowers the triplet loss.
* Seeding: with the same codes, the GPU graph equals the oracle model's graph
  bit for bit (the forest splits the codes, leaf pairs are exact).
* What the trained encoder gains: recall after a fixed number of iterations,
  against the random start and against an untrained encoder.
"""

from __future__ import annotations

import numpy as np
import pytest

import polariseq_amd as P
from oracle import mlp_oracle as M
from polariseq_amd import datasets as ds

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.N_PARAMS, np.float32)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(scale=0.15, size=M.H * M.IN)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(scale=0.05, size=M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.normal(scale=0.1, size=M.OUT * M.H)
    p[M.OFF_B2:] = rng.normal(scale=0.05, size=M.OUT)
    return p


def test_backward_matches_the_fp64_oracle():
    import torch
    x = ds.mixture(5000, 50, 10, seed=1)
    xt = torch.from_numpy(x).cuda()
    p = _params(3)
    mu, inv = 0.1, 0.2
    enc = P.Encoder(seed=0)
    enc.set_params(p, mu, inv)
    rng = np.random.default_rng(5)
    B, margin = 256, 1.0
    cand = rng.integers(0, 5000, size=(3, 4 * B))
    _, _, l = M.loss_and_grad(p, x[cand[0]], x[cand[1]], x[cand[2]], mu, inv, margin)
    keep = np.flatnonzero(np.abs(l) > 0.05)[:B]  # away from the hinge
    assert keep.size == B
    rows = np.concatenate([cand[0][keep], cand[1][keep], cand[2][keep]]).astype(np.uint32)
    g_gpu, L_gpu = enc.grad(xt, rows, margin)
    L64, g64, _ = M.loss_and_grad(p, x[rows[:B]], x[rows[B:2 * B]], x[rows[2 * B:]], mu, inv,
                                  margin)
    assert abs(L_gpu - L64) <= 1e-4 * max(1.0, abs(L64))
    scale = np.abs(g64).max()
    err = np.abs(g_gpu - g64).max() / scale
    assert err < 2e-4, err
    # every parameter block gets a non-trivial gradient
    for a, b in [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.N_PARAMS)]:
        assert np.abs(g64[a:b]).max() > 0


@pytest.mark.parametrize("n", [5003, 16, 1])
def test_encode_matches_the_fp64_oracle_forward(n):
    """The matrix-core encode (k_encode_mfma) against the oracle's forward with
    the same BF16-rounded weights: FP32 rounding only. Row counts that are not
    a multiple of the 16-row tile exercise the tail."""
    import torch
    x = ds.mixture(max(n, 2), 50, 10, seed=6)[:n]
    p = _params(9)
    mu, inv = -0.05, 0.3
    enc = P.Encoder(seed=0)
    enc.set_params(p, mu, inv)
    z = enc.encode(torch.from_numpy(np.ascontiguousarray(x)).cuda()).cpu().numpy()
    _, _, _, z64 = M.forward(p, x, mu, inv)
    scale = np.abs(z64).max()
    err = np.abs(z - z64).max() / scale
    assert err < 1e-5, err
    # a hidden unit dropped or a transposed weight block would not pass
    q = p.copy()
    q[M.OFF_W2:M.OFF_B2] = q[M.OFF_W2:M.OFF_B2].reshape(M.OUT, M.H)[:, ::-1].ravel()
    _, _, _, zq = M.forward(q, x, mu, inv)
    assert np.abs(z - zq).max() / scale > 1e-2


def test_training_lowers_the_loss_and_is_deterministic():
    import torch
    x = ds.mixture(30000, 50, 20, seed=2)
    xt = torch.from_numpy(x).cuda()
    a = P.Encoder(seed=1)
    info0 = a.train(xt, subset=20000, steps=1, seed=7)
    info = a.train(xt, subset=20000, steps=300, seed=7)
    assert info["last_loss"] < 0.5 * info0["last_loss"], (info0, info)
    b = P.Encoder(seed=1)
    b.train(xt, subset=20000, steps=1, seed=7)
    b.train(xt, subset=20000, steps=300, seed=7)
    assert np.array_equal(a.params()[0], b.params()[0])


def test_learned_seeding_matches_the_oracle_on_the_same_codes(orc):
    import torch
    x = ds.mixture(8000, 50, 12, seed=3)
    xt = torch.from_numpy(x).cuda()
    enc = P.Encoder(seed=2)
    enc.train(xt, subset=8000, steps=100, seed=1)
    codes = enc.encode(xt).cpu().numpy()
    g = P.nndescent_dense(x, 15, seed=4, init="learned", model=enc, n_trees=6)
    m, _ = orc.gpu_model_nndescent(x, 15, seed=4, n_trees=6, split_data=codes)
    assert np.array_equal(g["indptr"], m["indptr"])
    assert np.array_equal(g["indices"], m["indices"])
    assert np.array_equal(g["data"].view(np.uint32), m["data"].view(np.uint32))


def test_trained_encoder_gain_over_random_start_and_untrained(orc):
    import torch
    n = 60000
    x = ds.mixture(n, 50, 40, seed=4)
    xt = torch.from_numpy(x).cuda()
    q = np.random.default_rng(0).choice(n, 3000, replace=False).astype(np.uint32)
    ei, _ = orc.knn_exact64(x, 15, q)

    def rec(g):
        return ds.recall(g["indices"].reshape(n, 15)[q], ei).mean()

    trained = P.Encoder(seed=5)
    trained.train(xt, subset=30000, steps=400, seed=2)
    untrained = P.Encoder(seed=5)
    one = dict(max_iterations=1, delta=0.0, n_trees=8)
    r_learned = rec(P.nndescent_dense(x, 15, seed=1, init="learned", model=trained, **one))
    r_untrained = rec(P.nndescent_dense(x, 15, seed=1, init="learned", model=untrained, **one))
    r_random = rec(P.nndescent_dense(x, 15, seed=1, init="random", **one))
    assert r_learned > r_random + 0.2, (r_learned, r_random)
    assert r_learned > r_untrained, (r_learned, r_untrained)
    full = rec(P.nndescent_dense(x, 15, seed=1, init="learned", model=trained, n_trees=8))
    assert full >= 0.95

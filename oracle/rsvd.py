"""Oracle (test infrastructure only) for the randomized route of sparse PCA:
block Krylov SVD with implicit centring, restated from
crates/polariseq-core/src/rsvd.rs:186-208 (the Frobenius norm), 218-335 (the
products with the centred matrix), 414-435 (the Gaussian sketch), 476-566 (the
algorithm) and pca.rs:149-166, 273-286 (embedding and variance ratio), in
numpy/scipy float64 for small matrices.

The sketch is the reference's: ``gaussian_rowmajor`` draws from rand_chacha's
ChaCha8Rng keyed by rand_core's ``seed_from_u64`` (both third-party crates,
rand_chacha 0.10.0 and rand 0.10.2 in the reference's Cargo.lock, absent here),
restated from their published algorithms: the ChaCha block function
(Bernstein 2008), a 64-bit block counter from 0 and stream 0, the key from
eight PCG32 outputs (O'Neill 2014, XSH-RR), a u64 as two consecutive words low
first, a uniform as the top 53 bits times 2^-53, Box-Muller in float64 rounded
to float32.

What pins it (tests/test_rsvd_oracle.py): the ChaCha block function against
the published ChaCha20 and ChaCha8 keystreams of the all-zero key; the
route's answer against an exact SVD of the densified, centred matrix, at the
tolerances of the reference's own tests (rsvd.rs:1015-1068,
python/polariseq/tests/test_pca.py:38-62) and exactly (to rounding) where the
Krylov basis spans the whole row space; implicit against explicit centring;
the cut last block against the same basis taken whole. Parity unpinned: the
PCG32 seeding and the word order of rand_chacha's buffer (no reference fixture
holds sketch values), so the sketch's identity with the reference's rests on
those published descriptions.
"""

from __future__ import annotations

import math

import numpy as np

from .pca import canonicalize_signs

_M64 = (1 << 64) - 1
_SIGMA = (0x61707865, 0x3320646E, 0x79622D32, 0x6B206574)  # "expand 32-byte k"


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def chacha_blocks(key, counters, rounds: int) -> np.ndarray:
    """ChaCha block function (Bernstein 2008) for each 64-bit block counter:
    state = constants, the 8 key words, counter low/high, stream 0, 0; ``rounds``
    rounds (column then diagonal quarter rounds), then the input added back.
    Returns uint32 (len(counters), 16)."""
    c = np.asarray(counters, dtype=np.uint64).ravel()
    st = np.zeros((16, c.size), dtype=np.uint32)
    for i, v in enumerate(_SIGMA):
        st[i] = v
    for i, v in enumerate(key):
        st[4 + i] = np.uint32(v)
    st[12] = (c & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    st[13] = (c >> np.uint64(32)).astype(np.uint32)
    x = st.copy()

    def quarter(a, b, cc, d):
        x[a] = x[a] + x[b]
        x[d] = _rotl(x[d] ^ x[a], 16)
        x[cc] = x[cc] + x[d]
        x[b] = _rotl(x[b] ^ x[cc], 12)
        x[a] = x[a] + x[b]
        x[d] = _rotl(x[d] ^ x[a], 8)
        x[cc] = x[cc] + x[d]
        x[b] = _rotl(x[b] ^ x[cc], 7)

    for _ in range(rounds // 2):
        quarter(0, 4, 8, 12)
        quarter(1, 5, 9, 13)
        quarter(2, 6, 10, 14)
        quarter(3, 7, 11, 15)
        quarter(0, 5, 10, 15)
        quarter(1, 6, 11, 12)
        quarter(2, 7, 8, 13)
        quarter(3, 4, 9, 14)
    return (x + st).T.copy()


def seed_key(seed: int) -> list[int]:
    """rand_core's seed_from_u64: the 32-byte seed from eight PCG32 outputs
    (state <- state * 6364136223846793005 + 11634580027462260723, then
    rotr((((state >> 18) ^ state) >> 27) as u32, state >> 59)), little-endian,
    i.e. eight key words."""
    st = int(seed) & _M64
    key = []
    for _ in range(8):
        st = (st * 6364136223846793005 + 11634580027462260723) & _M64
        xs = (((st >> 18) ^ st) >> 27) & 0xFFFFFFFF
        rot = st >> 59
        key.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    return key


def gaussian_rowmajor(rows: int, cols: int, seed: int) -> np.ndarray:
    """rsvd.rs:414-435: a rows x cols float32 matrix, filled row by row from
    Box-Muller pairs (r cos(2 pi u2), r sin(2 pi u2)), r = sqrt(-2 ln max(u1,
    1e-300)), in float64 with the C library's log/cos/sin, rounded to float32."""
    count = rows * cols
    pairs = (count + 1) // 2
    n_u64 = 2 * pairs
    words = chacha_blocks(seed_key(seed), np.arange((2 * n_u64 + 15) // 16, dtype=np.uint64), 8).ravel()
    words = words[: 2 * n_u64].astype(np.uint64)
    u64 = words[0::2] | (words[1::2] << np.uint64(32))
    uni = (u64 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    out = np.empty(2 * pairs, dtype=np.float32)
    tau = 2.0 * math.pi
    for p in range(pairs):
        u1 = max(float(uni[2 * p]), 1e-300)
        r = math.sqrt(-2.0 * math.log(u1))
        th = tau * float(uni[2 * p + 1])
        out[2 * p] = r * math.cos(th)
        out[2 * p + 1] = r * math.sin(th)
    return out[:count].reshape(rows, cols)


def widths(n: int, m: int, k: int, oversample: int, n_iter: int):
    """rsvd.rs:496-499: l = min(k + oversample, min(n, m)) (at least 1) and the
    basis width min(l (n_iter + 1), min(n, m))."""
    most = min(n, m)
    l = max(1, min(k + oversample, most))
    return l, min(l * (n_iter + 1), most)


def rsvd(indptr, indices, values, n: int, m: int, k: int, oversample: int, n_iter: int, seed: int,
         center: bool = True) -> dict:
    """rsvd.rs:476-566 with pca.rs:149-166, 273-286: Q_0 = qr(A' Omega),
    Q_d = qr(A' qr(A'^T Q_{d-1})) while d l < width, a last block that does not
    fit cut to the width (the reference's slice would overrun there), Q =
    qr([Q_0 .. Q_q]), B^T = A'^T Q = P S W^T, U = Q W[:, :k]; embedding U S with
    canonical column signs, ratio s^2 / ||A'||_F^2, components P[:, :k] with the
    embedding's signs. A' = S - 1 mu^T, mu the float32 column means."""
    import scipy.sparse as sp

    ip = np.asarray(indptr, dtype=np.int64)
    nnz = int(ip[-1])
    vals = np.asarray(values[:nnz], dtype=np.float32).astype(np.float64)
    a = sp.csr_matrix((vals, np.asarray(indices[:nnz], dtype=np.int64), ip), shape=(n, m))
    mu = None
    total = float(np.sum(vals * vals))
    if center:
        sums = np.bincount(np.asarray(indices[:nnz], dtype=np.int64), weights=vals, minlength=m)
        mu = (sums / max(n, 1)).astype(np.float32).astype(np.float64)
        total -= n * float(np.sum(mu * mu))

    def fwd(x):
        y = a @ x
        return y - (mu @ x)[None, :] if mu is not None else y

    def bwd(y):
        z = a.T @ y
        return z - np.outer(mu, y.sum(axis=0)) if mu is not None else z

    def qr(y):
        return np.linalg.qr(y)[0]

    l, width = widths(n, m, k, oversample, n_iter)
    omega = gaussian_rowmajor(m, l, seed).astype(np.float64)
    block = qr(fwd(omega))
    basis = [block]
    d = 1
    while d <= n_iter and d * l < width:
        block = qr(fwd(qr(bwd(block))))
        basis.append(block[:, : min(l, width - d * l)])
        d += 1
    q = qr(np.hstack(basis))
    p, s, wt = np.linalg.svd(bwd(q), full_matrices=False)
    u = q @ wt.T[:, :k]
    raw = (u * s[:k][None, :]).astype(np.float32)
    emb = canonicalize_signs(raw)
    flip = np.where((emb != raw).any(axis=0), -1.0, 1.0).astype(np.float32)
    comps = p[:, :k].astype(np.float32) * flip[None, :]
    ratio = (s[:k] * s[:k] / total).astype(np.float32)
    return {"embedding": emb, "variance_ratio": ratio, "components": comps, "singular_values": s[:k],
            "total": total, "width": width}


def rsvd_scaled(indptr, indices, data, n: int, m: int, k: int, max_value, seed: int) -> dict:
    """pca.rs:396-431 on the randomized route: the scaled view's stored values,
    always centred, the reference's default options (oversample 10, n_iter 1)."""
    from .pca import scaled_values

    s = scaled_values(indptr, indices, data, n, m, max_value)
    return rsvd(indptr, indices, s, n, m, k, 10, 1, seed, True)

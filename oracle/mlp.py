"""Oracle of the learned block preconditioner (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.  The learned preconditioner has no counterpart in the
reference: it replaces the block-Jacobi 'c' solve of the local GMRES
(djac_gen / djac_apply 'c', tt_linalg.f90:397-429 and :497-533), whose
blocks are
    B_b(i, j) = sum_{a1,a2} Phi1(b1,b1,a1) A(a1,i,j,a2) Phi2(b2,a2,b2),
one per pair b = (b1, b2), acting on the slice x(b1, :, b2) of a local
vector.  The network sees, per block, the three central diagonals of B_b and
the residual slice, all normalised; it predicts the normalised correction.

Conventions (shared with the device kernels):
  * s_b = mean_i |B_b(i, i)| (the block's scale), rn_b = ||r_b||_2;
  * features X(:, b) = [diag(B_b) / s_b, superdiag / s_b (last entry 0),
    subdiag / s_b (last entry 0), r_b / rn_b]   (4n = 256 rows for n = 64);
  * network: h1 = relu(W1^T X + b1), h2 = relu(W2^T h1 + b2),
    y = W3^T h2 + b3 + X[3n:] / X[:n] with W1 (256 x 256), W2 (256 x 256),
    W3 (256 x 64); the last term is the point-Jacobi solution of the block
    (residual over diagonal, 0 where the diagonal is 0), so the layers learn
    the correction the block's off-diagonals make to it; weights rounded to
    bfloat16 (round to nearest even), float32 arithmetic in training; at
    inference the layer inputs are rounded to bfloat16 too (the bf16 matrix
    cores, fp32 accumulation);
  * correction z_b = y_b * rn_b / s_b;
  * training target y*_b = (B_b^-1 r_b) * s_b / rn_b; loss
    L = sum (y - y*)^2 / (M * n) over a batch of M blocks; Adam.

Pinning: gradients are checked by central finite differences (float64) and
the features against explicit dense blocks built from djac_gen's block
definition (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

from .kernels import F

K_IN = 256
HID = 256
N_OUT = 64


def param_shapes():
    return [("W1", (K_IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
            ("W3", (HID, N_OUT)), ("b3", (N_OUT,))]


def n_params():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def unflatten(theta):
    """Flat parameter vector (column-major matrices, in param_shapes order)."""
    out, off = {}, 0
    for name, shape in param_shapes():
        sz = int(np.prod(shape))
        out[name] = np.reshape(theta[off:off + sz], shape, order="F")
        off += sz
    return out


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def block_diagonals(rx1, n, rx2, ra1, ra2, phi1, A, phi2):
    """Diagonal, super- and sub-diagonal of every 'c' block: arrays
    (n, rx1, rx2) each (tt_linalg.f90:412-416 block definition)."""
    P1 = F(phi1, rx1, rx1, ra1)
    AA = F(A, ra1, n, n, ra2)
    P2 = F(phi2, rx2, ra2, rx2)
    d1 = np.stack([np.diag(P1[:, :, a]) for a in range(ra1)], axis=1)   # (rx1, ra1)
    d2 = np.stack([np.diag(P2[:, a, :]) for a in range(ra2)], axis=1)   # (rx2, ra2)
    i = np.arange(n)
    diag = np.einsum("pa,aik,qk->ipq", d1, AA[:, i, i, :], d2)
    up = np.zeros((n, rx1, rx2))
    lo = np.zeros((n, rx1, rx2))
    if n > 1:
        i = np.arange(n - 1)
        up[:-1] = np.einsum("pa,aik,qk->ipq", d1, AA[:, i, i + 1, :], d2)
        lo[:-1] = np.einsum("pa,aik,qk->ipq", d1, AA[:, i + 1, i, :], d2)
    return diag, up, lo


def features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v):
    """Features X (4n x nb, blocks ordered b = b1 + rx1 * b2) and the output
    scale rn_b / s_b (nb) for the local vector v (rx1, n, rx2)."""
    diag, up, lo = block_diagonals(rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    V = F(v, rx1, n, rx2)
    nb = rx1 * rx2
    X = np.zeros((4 * n, nb))
    scale = np.zeros(nb)
    for b2 in range(rx2):
        for b1 in range(rx1):
            b = b1 + rx1 * b2
            s = np.mean(np.abs(diag[:, b1, b2]))
            r = V[b1, :, b2]
            rn = np.linalg.norm(r)
            s_inv = 1.0 / s if s > 0 else 0.0
            X[0:n, b] = diag[:, b1, b2] * s_inv
            X[n:2 * n, b] = up[:, b1, b2] * s_inv
            X[2 * n:3 * n, b] = lo[:, b1, b2] * s_inv
            X[3 * n:, b] = r / rn if rn > 0 else 0.0
            scale[b] = rn * s_inv
    return X, scale


def block_solve_targets(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v):
    """Normalised targets y*_b = (B_b^-1 r_b) s_b / rn_b (n x nb)."""
    P1 = F(phi1, rx1, rx1, ra1)
    AA = F(A, ra1, n, n, ra2)
    P2 = F(phi2, rx2, ra2, rx2)
    d1 = np.stack([np.diag(P1[:, :, a]) for a in range(ra1)], axis=1)
    d2 = np.stack([np.diag(P2[:, a, :]) for a in range(ra2)], axis=1)
    blk = np.einsum("pa,aijk,qk->pijq", d1, AA, d2)
    V = F(v, rx1, n, rx2)
    T = np.zeros((n, rx1 * rx2))
    for b2 in range(rx2):
        for b1 in range(rx1):
            B = blk[b1, :, :, b2]
            r = V[b1, :, b2]
            rn = np.linalg.norm(r)
            s = np.mean(np.abs(np.diag(B)))
            if rn > 0 and s > 0:
                T[:, b1 + rx1 * b2] = np.linalg.solve(B, r) * s / rn
    return T


def forward(theta, X, round_weights=True, dtype=np.float32, round_acts=False):
    """Network output y (64 x M) and the cache for backward.  With
    round_acts, every layer's input is rounded to bfloat16 before its
    product (the inference form on the bf16 matrix cores); the point-Jacobi
    term always uses the unrounded features."""
    p = unflatten(np.asarray(theta, dtype=np.float64))
    W = {k: (bf16_round(p[k]) if round_weights else p[k]).astype(dtype) for k in ("W1", "W2", "W3")}
    b = {k: p[k].astype(dtype) for k in ("b1", "b2", "b3")}
    X = X.astype(dtype)
    rnd = (lambda a: bf16_round(a).astype(dtype)) if round_acts else (lambda a: a)
    z1 = W["W1"].T @ rnd(X) + b["b1"][:, None]
    h1 = np.maximum(z1, 0)
    z2 = W["W2"].T @ rnd(h1) + b["b2"][:, None]
    h2 = np.maximum(z2, 0)
    y = W["W3"].T @ rnd(h2) + b["b3"][:, None] + jacobi_term(X)
    return y, dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, W=W)


def jacobi_term(X):
    """X[3n:] / X[:n] elementwise, 0 where the diagonal feature is 0."""
    n = N_OUT
    d = X[:n]
    r = X[3 * n:4 * n]
    out = np.zeros_like(r)
    nz = d != 0
    out[nz] = r[nz] / d[nz]
    return out


def loss(y, T):
    M = y.shape[1]
    return float(np.sum((y.astype(np.float64) - T) ** 2) / (M * y.shape[0]))


def backward(cache, y, T):
    """Gradient of loss(y, T) with respect to the flat parameters (the
    weight gradients are taken at the rounded weights)."""
    M, n = y.shape[1], y.shape[0]
    dy = (2.0 / (M * n)) * (y - T.astype(y.dtype))
    W = cache["W"]
    gW3 = cache["h2"] @ dy.T
    gb3 = dy.sum(axis=1)
    dh2 = W["W3"] @ dy
    dz2 = dh2 * (cache["z2"] > 0)
    gW2 = cache["h1"] @ dz2.T
    gb2 = dz2.sum(axis=1)
    dh1 = W["W2"] @ dz2
    dz1 = dh1 * (cache["z1"] > 0)
    gW1 = cache["X"] @ dz1.T
    gb1 = dz1.sum(axis=1)
    parts = dict(W1=gW1, b1=gb1, W2=gW2, b2=gb2, W3=gW3, b3=gb3)
    return np.concatenate([np.reshape(parts[k], -1, order="F") for k, _ in param_shapes()])


def adam_step(theta, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v

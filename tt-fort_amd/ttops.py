"""Host-side TT arithmetic for checking solver results in FP64: TT-matrix by
TT-vector products, differences, norms and rounding, on flat core arrays in
the reference's layout (core k of a TT vector is (r_{k-1}, n_k, r_k), of a
TT-matrix (ra_{k-1}, n_k, m_k, ra_k), column-major, concatenated as in
tt_adapt_als.f90:28-37).

These run in numpy on the host and are used for global residuals
||A x - y|| / ||y|| of the solvers' results; the exact difference is formed
in TT format and its norm taken after left-orthogonalisation, so small
residuals are not lost to cancellation.
"""
from __future__ import annotations

import numpy as np


def split(cr, r, n, m=None):
    """Flat cores -> list of arrays (r_{k-1}, n_k[, m_k], r_k)."""
    out, off = [], 0
    for k in range(len(n)):
        shape = (int(r[k]), int(n[k])) + ((int(m[k]),) if m is not None else ()) + (int(r[k + 1]),)
        sz = int(np.prod(shape))
        out.append(np.asarray(cr[off:off + sz]).reshape(shape, order="F"))
        off += sz
    return out


def join(cores):
    """List of cores -> (flat array, ranks)."""
    r = [cores[0].shape[0]] + [c.shape[-1] for c in cores]
    return np.concatenate([c.reshape(-1, order="F") for c in cores]), np.array(r, dtype=np.int64)


def matvec(crA, ra, n, m, cr, rx):
    """y = A x with ranks ra_k * rx_k (cores of y: (ra rx, n, ra' rx'))."""
    A = split(crA, ra, n, m)
    X = split(cr, rx, m)
    out = []
    for a, x in zip(A, X):
        y = np.einsum("aijb,cjd->acibd", a, x, optimize=True)
        out.append(y.reshape(a.shape[0] * x.shape[0], a.shape[1], a.shape[3] * x.shape[2], order="F"))
    return out


def sub(c1, c2):
    """Cores of x1 - x2 (block-diagonal interior cores, ranks add)."""
    d = len(c1)
    out = []
    for k, (a, b) in enumerate(zip(c1, c2)):
        if d == 1:
            out.append(a - b)
            continue
        if k == 0:
            out.append(np.concatenate([a, -b], axis=2))
        elif k == d - 1:
            out.append(np.concatenate([a, b], axis=0))
        else:
            z = np.zeros((a.shape[0] + b.shape[0], a.shape[1], a.shape[2] + b.shape[2]), dtype=np.result_type(a, b))
            z[:a.shape[0], :, :a.shape[2]] = a
            z[a.shape[0]:, :, a.shape[2]:] = b
            out.append(z)
    return out


def norm(cores):
    """Frobenius norm of a TT vector by left-orthogonalisation."""
    R = np.ones((1, 1), dtype=cores[0].dtype)
    for c in cores:
        c = np.einsum("ab,bic->aic", R, c)
        r0, nn, r1 = c.shape
        # reduced QR: R is min(r0 nn, r1) x r1 and carries the norm onwards
        _, R = np.linalg.qr(c.reshape(r0 * nn, r1, order="F"))
    return float(np.linalg.norm(R))


def residual(crA, ra, n, crX, rx, crY, ry):
    """||A x - y|| / ||y|| computed in FP64 in TT format."""
    Ax = matvec(crA, ra, n, n, crX, rx)
    Y = split(crY, ry, n)
    return norm(sub(Ax, Y)) / norm(Y)


def round_tt(cores, eps, rmax=None):
    """TT rounding (right-to-left QR, then left-to-right truncated SVD) to
    relative Frobenius accuracy eps and ranks at most rmax."""
    d = len(cores)
    cores = [c.copy() for c in cores]
    for k in range(d - 1, 0, -1):
        r0, nn, r1 = cores[k].shape
        Q, R = np.linalg.qr(cores[k].reshape(r0, nn * r1).T)
        cores[k] = Q.T.reshape(-1, nn, r1)
        cores[k - 1] = np.einsum("aib,cb->aic", cores[k - 1], R)
    tol = eps / np.sqrt(max(d - 1, 1)) * np.linalg.norm(cores[0])
    for k in range(d - 1):
        r0, nn, r1 = cores[k].shape
        U, s, Vt = np.linalg.svd(cores[k].reshape(r0 * nn, r1), full_matrices=False)
        tail = np.sqrt(np.cumsum(s[::-1] ** 2))[::-1]
        keep = max(1, int(np.sum(tail > tol)))
        if rmax is not None:
            keep = min(keep, rmax)
        cores[k] = U[:, :keep].reshape(r0, nn, keep)
        cores[k + 1] = np.einsum("ab,bic->aic", s[:keep, None] * Vt[:keep], cores[k + 1])
    return cores


def block_eig_residuals(crA, ra, n, cr, ry, lam):
    """FP64 eigen-residuals ||A x_b - lambda_b x_b|| / ||x_b|| of a block TT
    whose last core carries the block index (ranks ry, ry[-1] = B), as
    tt_eigb returns it.

    The difference A x_b - lambda_b x_b is formed exactly in TT format
    (cores [A x | x] block-diagonal, the last one [A x_b ; -lambda_b x_b]) and
    its norm taken by left-orthogonalisation, so residuals far below |lambda|
    are not lost to cancellation.  The left sweep is shared by all B
    eigenpairs; only the last core differs.  Also returns the Gram matrix
    X^T X of the block (from the last core, exact when the other cores are
    left-orthogonal, which the norms do not assume)."""
    d = len(n)
    A = split(crA, ra, n, n)
    X = split(cr, ry, n)
    B = X[-1].shape[2]
    R = np.ones((1, 1))
    Rx = np.ones((1, 1))   # left-orthogonalisation of x alone, for the Gram matrix and norms
    for k in range(d - 1):
        a, x = A[k], X[k]
        ax = np.einsum("aijb,cjd->acibd", a, x, optimize=True)
        ax = ax.reshape(a.shape[0] * x.shape[0], a.shape[1], a.shape[3] * x.shape[2], order="F")
        r0a, nn, r1a = ax.shape
        r0x, _, r1x = x.shape
        if k == 0:
            core = np.concatenate([ax, x], axis=2)
        else:
            core = np.zeros((r0a + r0x, nn, r1a + r1x))
            core[:r0a, :, :r1a] = ax
            core[r0a:, :, r1a:] = x
        c = np.einsum("ab,bic->aic", R, core, optimize=True)
        _, R = np.linalg.qr(c.reshape(c.shape[0] * nn, c.shape[2], order="F"))
        cx = np.einsum("ab,bic->aic", Rx, x, optimize=True)
        _, Rx = np.linalg.qr(cx.reshape(cx.shape[0] * nn, cx.shape[2], order="F"))
    a, x = A[-1], X[-1]
    res = np.zeros(B)
    gram_core = np.einsum("ab,bic->aic", Rx, x, optimize=True).reshape(-1, B, order="F")
    G = gram_core.T @ gram_core
    for b in range(B):
        xb = x[:, :, b:b + 1]
        ax = np.einsum("aijb,cjd->acibd", a, xb, optimize=True)
        ax = ax.reshape(a.shape[0] * xb.shape[0], a.shape[1], 1, order="F")
        last = np.concatenate([ax, -lam[b] * xb], axis=0)
        v = np.einsum("ab,bic->aic", R, last, optimize=True)
        res[b] = np.linalg.norm(v) / np.sqrt(G[b, b])
    return res, G

"""CPU pins of the oracle pieces behind the AMEn path: the Dirichlet
Laplacian TT-matrix against its dense Kronecker sum, the tensor-DST exact
solve against a dense solve, the pivoted incomplete Cholesky (duchol) on
exactly low-rank Gram matrices (where its output is fixed by linear
algebra), and the block-Jacobi blocks against the dense local matrix."""
import numpy as np
import pytest

import mpo
from oracle import kernels as K
from oracle import operators as O


def test_laplace_tt_matches_kronecker_sum():
    d, n = 3, 5
    crA, ra, nn = mpo.laplace_dd(d, n)
    cores, off = [], 0
    for k in range(d):
        sz = ra[k] * n * n * ra[k + 1]
        cores.append(crA[off:off + sz])
        off += sz
    Ad = O.mpo_to_dense(cores, ra, nn, nn)
    L = (n + 1) ** 2 * (2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1))
    I = np.eye(n)
    ref = np.zeros((n ** d, n ** d))
    for k in range(d):   # site 1 fastest: kron(site d, ..., site 1)
        mats = [L if j == k else I for j in range(d)]
        t = mats[d - 1]
        for j in range(d - 2, -1, -1):
            t = np.kron(t, mats[j])
        ref += t
    assert np.allclose(Ad, ref)


def test_dst_solve_matches_dense():
    d, n = 3, 6
    crA, ra, nn = mpo.laplace_dd(d, n)
    cores, off = [], 0
    for k in range(d):
        sz = ra[k] * n * n * ra[k + 1]
        cores.append(crA[off:off + sz])
        off += sz
    Ad = O.mpo_to_dense(cores, ra, nn, nn)
    rng = np.random.default_rng(0)
    y = rng.standard_normal(n ** d)
    x = O.laplace_dst_solve(d, n, y)
    assert np.allclose(x, np.linalg.solve(Ad, y), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_duchol_recovers_low_rank_gram(trans):
    """For items spanning an exactly rank-3 Gram matrix G, three pivoted
    Cholesky steps reproduce G = du diag(lam) du^T, with du an orthonormal
    basis of range(G) and lam its nonzero eigenvalues."""
    rng = np.random.default_rng(3)
    m, n, r = 9, 14, 3
    Am = rng.standard_normal((m, r)) @ rng.standard_normal((r, n))   # m x n, items = columns
    dA = Am if trans == "N" else Am.T                               # 'T': n x m, items = rows
    du, lam, real = K.duchol(trans, m, n, dA.reshape(-1, order="F"), 3)
    G = Am.T @ Am
    assert real == 3
    assert np.allclose(du.T @ du, np.eye(3), atol=1e-12)
    assert np.allclose(du @ np.diag(lam) @ du.T, G, atol=1e-9 * np.abs(G).max())
    ev = np.sort(np.linalg.eigvalsh(G))[-3:]
    assert np.allclose(np.sort(lam), ev, rtol=1e-10)


def test_duchol_zero_matrix_stops():
    du, lam, real = K.duchol("N", 4, 5, np.zeros(20), 3)
    assert real == 1 and np.all(du == 0)


@pytest.mark.parametrize("ptype", ["c", "l", "r"])
def test_jacobi_blocks_are_inverse_diagonal_blocks(ptype):
    """'c' blocks invert the (i, j) blocks of B restricted to the diagonal
    interface entries; with diagonal Phi1 and Phi2 ('c'), diagonal Phi2 ('l')
    or diagonal Phi1 ('r') these are exact diagonal blocks of the local matrix."""
    rng = np.random.default_rng(11)
    rx1, n, rx2, ra1, ra2 = 3, 4, 2, 2, 3
    P1 = rng.standard_normal((rx1, rx1, ra1))
    P2 = rng.standard_normal((rx2, ra2, rx2))
    if ptype in "cr":
        P1 = np.stack([np.diag(np.diag(P1[:, :, a])) for a in range(ra1)], axis=2)
    if ptype in "cl":
        P2 = np.stack([np.diag(np.diag(P2[:, a, :])) for a in range(ra2)], axis=1)
    A = rng.standard_normal((ra1, n, n, ra2))
    A[0, :, :, 0] += 5 * np.eye(n)
    phi1, Af, phi2 = (K._flat(v) for v in (P1, A, P2))
    B = K.d2d_fullmat(rx1, n, rx2, ra1, ra2, phi1, Af, phi2).reshape(rx1 * n * rx2, -1, order="F")
    J = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, Af, phi2)
    x = rng.standard_normal(rx1 * n * rx2)
    y = K.djac_apply(ptype, rx1, n, rx2, J, B @ x)
    # P B x == x when B is block diagonal in P's block structure
    Bm = B.copy()
    idx = np.arange(rx1 * n * rx2)
    b1, i, b2 = idx % rx1, (idx // rx1) % n, idx // (rx1 * n)
    if ptype == "c":
        same = (b1[:, None] == b1[None, :]) & (b2[:, None] == b2[None, :])
    elif ptype == "l":
        same = b2[:, None] == b2[None, :]
    else:
        same = b1[:, None] == b1[None, :]
    Bm[~same] = 0.0
    y = K.djac_apply(ptype, rx1, n, rx2, J, Bm @ x)
    assert np.allclose(y, x, atol=1e-10)

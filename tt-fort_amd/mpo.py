"""TT-matrix (MPO) builders for the benchmark Hamiltonians.

Cores are returned in the layout the C ABI takes (ra_{k-1}, n_k, m_k, ra_k),
column-major, concatenated as crA with offsets compute_ps
(tt_adapt_als.f90:28-37).  Site 1 is the fastest-varying index of a full
vector, as in the reference's fixtures (Heisen_20.sdv).
"""
from __future__ import annotations

import numpy as np

_I2 = np.eye(2)
_SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
_SP = np.array([[0.0, 1.0], [0.0, 0.0]])
_SM = np.array([[0.0, 0.0], [1.0, 0.0]])
_PX = np.array([[0.0, 1.0], [1.0, 0.0]])
_PZ = np.array([[1.0, 0.0], [0.0, -1.0]])


def _cores_from_bond_matrix(d, W, first_row, last_col):
    cores, ra = [], [1]
    for k in range(d):
        if k == 0:
            Wk = W[first_row:first_row + 1, :] if d > 1 else W[first_row:first_row + 1, last_col:last_col + 1]
        elif k == d - 1:
            Wk = W[:, last_col:last_col + 1]
        else:
            Wk = W
        cores.append(np.transpose(Wk, (0, 2, 3, 1)).reshape(-1, order="F"))   # (a1, i, j, a2)
        ra.append(Wk.shape[1])
    return cores, np.array(ra, dtype=np.int64)


def xxz(d, J=1.0, delta=1.0):
    """Open XXZ chain H = sum J (Sx Sx + Sy Sy + delta Sz Sz), S = sigma/2
    (rank-5 MPO).  J=1, delta=1 is the Heisenberg XXX chain of the
    reference fixture Heisen_20.sdv.  Returns (crA, ra, n)."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0] = _I2
    W[1, 0] = _SP
    W[2, 0] = _SM
    W[3, 0] = _SZ
    W[4, 1] = 0.5 * J * _SM
    W[4, 2] = 0.5 * J * _SP
    W[4, 3] = J * delta * _SZ
    W[4, 4] = _I2
    cores, ra = _cores_from_bond_matrix(d, W, 4, 0)
    return np.concatenate(cores), ra, np.full(d, 2, dtype=np.int64)


def tfim(d, g=1.0, J=1.0):
    """Transverse-field Ising chain H = -J sum Z_i Z_{i+1} - g sum X_i
    (rank-3 MPO).  Returns (crA, ra, n)."""
    W = np.zeros((3, 3, 2, 2))
    W[0, 0] = _I2
    W[1, 0] = _PZ
    W[2, 0] = -g * _PX
    W[2, 1] = -J * _PZ
    W[2, 2] = _I2
    cores, ra = _cores_from_bond_matrix(d, W, 2, 0)
    return np.concatenate(cores), ra, np.full(d, 2, dtype=np.int64)


def tfim_schroedinger(d, g=1.0, J=1.0):
    """A = -i H for the transverse-field Ising chain (the reference's time
    stepping convention: dY/dt = A Y with the factor i carried by A, as in
    test_ksl.dat).  Returns (crA complex128, ra, n)."""
    crH, ra, n = tfim(d, g, J)
    # -i scales the operator: put it on the first core only
    sz0 = int(ra[0] * n[0] * n[0] * ra[1])
    crA = crH.astype(np.complex128)
    crA[:sz0] *= -1j
    return crA, ra, n


def product_state_padded(d, r, index=0, n=2):
    """The product state e_index x ... x e_index as a TT with ranks
    min(r, n^k, n^(d-k)), zero-padded beyond the first rank component (the
    fixed-rank start of a KSL run).  Returns (cr complex128, ry)."""
    ry = [1] + [int(min(r, n ** k, n ** (d - k))) for k in range(1, d)] + [1]
    cores = []
    for k in range(d):
        c = np.zeros((ry[k], n, ry[k + 1]), dtype=np.complex128)
        c[0, index, 0] = 1.0
        cores.append(c.reshape(-1, order="F"))
    return np.concatenate(cores), np.array(ry, dtype=np.int64)


def random_block_tt(d, n, r0, B, seed):
    """Initial block TT: N(0,1) cores of rank r0 from PCG64(seed), last rank B."""
    rng = np.random.default_rng(seed)
    ry = [1] + [r0] * (d - 1) + [B]
    crY = np.concatenate([rng.standard_normal(ry[k] * int(n[k]) * ry[k + 1]) for k in range(d)])
    return crY, ry


def laplace_dd(d, n):
    """Dirichlet Laplacian on the interior grid of [0,1]^d with n points per
    direction, h = 1/(n+1): A = sum_k I x .. x L_k x .. x I with
    L = tridiag(-1, 2, -1) / h^2 (rank-2 TT-matrix).  The a == 1 case of the
    diffusion operator of the linear-solver workload.  Returns (crA, ra, nn)."""
    L = (n + 1) ** 2 * (2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1))
    W = np.zeros((2, 2, n, n))
    W[0, 0] = np.eye(n)
    W[1, 0] = L
    W[1, 1] = np.eye(n)
    cores, ra = _cores_from_bond_matrix(d, W, 1, 0)
    return np.concatenate(cores), ra, np.full(d, n, dtype=np.int64)


def ones_tt(d, n):
    """The all-ones vector as a rank-1 TT.  Returns (cr, r)."""
    return np.ones(d * n), np.ones(d + 1, dtype=np.int64)


def random_tt(d, n, r, seed):
    """N(0,1) cores of rank r (boundary ranks 1) from PCG64(seed).  Returns (cr, rr)."""
    rng = np.random.default_rng(seed)
    rr = [1] + [r] * (d - 1) + [1]
    cr = np.concatenate([rng.standard_normal(rr[k] * n * rr[k + 1]) for k in range(d)])
    return cr, np.array(rr, dtype=np.int64)

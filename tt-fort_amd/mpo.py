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


def maxvol(A, tol=1.05, maxit=200):
    """Rows of the tall n x r matrix A whose r x r submatrix has locally
    maximal volume: start from the LU pivot rows, then swap a row in while
    some coefficient of A expressed in the chosen rows exceeds tol in modulus
    (Goreinov, Oseledets, Savostyanov, Tyrtyshnikov, Zamarashkin, "How to find
    a good submatrix", 2010)."""
    import scipy.linalg as sl
    n, r = A.shape
    if n <= r:
        return np.arange(n)
    P, _, _ = sl.lu(A)
    idx = np.argmax(P, axis=0)[:r].copy()
    for _ in range(maxit):
        B = np.linalg.solve(A[idx].T, A.T).T          # n x r with B[idx] = I
        i, j = np.unravel_index(np.argmax(np.abs(B)), B.shape)
        if abs(B[i, j]) <= tol:
            break
        idx[j] = i
    return idx


def tt_cross(f, n, r, sweeps=4, seed=0):
    """TT approximation of a function on a tensor grid by alternating
    maxvol cross interpolation (Oseledets, Tyrtyshnikov, "TT-cross
    approximation of multidimensional arrays", 2010) at fixed rank r.
    f takes an (npoints, d) integer index array and returns npoints values.
    Returns the list of cores (r_{k-1}, n_k, r_k); the last core carries the
    values, the others interpolate on the selected index sets."""
    d = len(n)
    rng = np.random.default_rng(seed)
    rk = [1] + [int(min(r, np.prod(n[:k]), np.prod(n[k:]))) for k in range(1, d)] + [1]
    J = [None] * (d + 1)
    J[d] = np.zeros((1, 0), dtype=np.int64)
    for k in range(d - 1, 0, -1):
        J[k] = np.stack([rng.integers(0, n[l], rk[k]) for l in range(k, d)], axis=1)
    I = [None] * (d + 1)
    I[0] = np.zeros((1, 0), dtype=np.int64)

    def fiber(k):
        Ik, Jk = I[k], J[k + 1]
        a, b = Ik.shape[0], Jk.shape[0]
        idx = np.empty((a, n[k], b, d), dtype=np.int64)
        idx[:, :, :, :k] = Ik[:, None, None, :]
        idx[:, :, :, k] = np.arange(n[k])[None, :, None]
        idx[:, :, :, k + 1:] = Jk[None, None, :, :]
        return np.asarray(f(idx.reshape(-1, d)), dtype=np.float64).reshape(a, n[k], b)

    def left_step(k):
        C = fiber(k).reshape(rk[k] * n[k], -1)
        Q, _ = np.linalg.qr(C)
        sel = maxvol(Q)
        rk[k + 1] = len(sel)
        ii, xx = np.unravel_index(sel, (rk[k], n[k]))
        I[k + 1] = np.concatenate([I[k][ii], xx[:, None]], axis=1)
        return Q, sel

    for _ in range(sweeps):
        for k in range(d - 1):
            left_step(k)
        for k in range(d - 1, 0, -1):
            C = fiber(k).reshape(rk[k], n[k] * rk[k + 1])
            Q, _ = np.linalg.qr(C.T)
            sel = maxvol(Q)
            rk[k] = len(sel)
            xx, jj = np.unravel_index(sel, (n[k], rk[k + 1]))
            J[k] = np.concatenate([xx[:, None], J[k + 1][jj]], axis=1)
    cores = []
    for k in range(d - 1):
        Q, sel = left_step(k)
        cores.append(np.linalg.solve(Q[sel].T, Q.T).T.reshape(rk[k], n[k], rk[k + 1]))
    cores.append(fiber(d - 1).reshape(rk[d - 1], n[d - 1], 1))
    return cores


def diffusion_7pt(d, n, seed=0, coef_rank=3, amp=0.5):
    """Variable-coefficient diffusion -div(a grad u) on the interior grid of
    [0,1]^d (n points per direction, h = 1/(n+1), Dirichlet boundary), the
    7-point (2d+1) stencil with half-point coefficients averaged from the
    neighbouring grid points:
        (A u)_p = sum_k [a_{p+e_k/2} (u_p - u_{p+e_k}) + a_{p-e_k/2} (u_p - u_{p-e_k})] / h^2,
        a(x) = 1 + amp * tanh(sum_k c_k x_k),  c_k ~ U(-1, 1) from PCG64(seed).
    The coefficient on the (n+2)^d grid including the boundary points is
    represented by its rank-`coef_rank` TT cross approximation a_TT with cores
    G_k; the operator is then exactly
        A = sum_k (x)_{l != k} diag(G_l interior) (x) T(avg G_k)
    with T(g) = D^T diag(g) D / h^2 (D the (n+1) x n difference), written as
    a TT-matrix of rank 2 * coef_rank (a "not yet / already differentiated"
    flag times the coefficient bond).
    Returns (crA, ra, nn, info) with info = dict(c, coef_cores, coef_maxerr)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, d)
    h = 1.0 / (n + 1)
    xg = np.arange(n + 2) * h                  # boundary points included

    def coef(idx):
        return 1.0 + amp * np.tanh(xg[idx] @ c)

    G = tt_cross(coef, [n + 2] * d, coef_rank, seed=seed)
    test = rng.integers(0, n + 2, (4096, d))
    v = np.ones((test.shape[0], 1))
    for k in range(d):
        v = np.einsum("pa,apb->pb", v, G[k][:, test[:, k], :])
    maxerr = float(np.abs(v[:, 0] - coef(test)).max())
    Dm = np.zeros((n + 1, n))
    for j in range(n + 1):
        if j < n:
            Dm[j, j] = 1.0
        if j >= 1:
            Dm[j, j - 1] = -1.0
    cores, ra = [], [1]
    for k in range(d):
        Gk = G[k]
        r0, r1 = Gk.shape[0], Gk.shape[2]
        # bond state (alpha, flag): flag 1 = no direction differentiated yet,
        # flag 0 = one already differentiated; first core starts at flag 1,
        # last core ends at flag 0
        W = np.zeros((2 * r0, 2 * r1, n, n))
        for a in range(r0):
            for b in range(r1):
                g = Gk[a, :, b]
                Dg = np.diag(g[1:n + 1])
                gh = 0.5 * (g[:-1] + g[1:])                  # n+1 half points
                T = Dm.T @ (gh[:, None] * Dm) / h ** 2
                W[2 * a + 1, 2 * b + 1] = Dg
                W[2 * a + 1, 2 * b] = T
                W[2 * a, 2 * b] = Dg
        if k == 0:
            W = W[1:2, :, :, :]                         # alpha_0 = 0, flag 1
        if k == d - 1:
            W = W[:, 0:1, :, :]                         # alpha_d = 0, flag 0
        cores.append(np.transpose(W, (0, 2, 3, 1)).reshape(-1, order="F"))
        ra.append(W.shape[1])
    info = dict(c=c, coef_cores=G, coef_maxerr=maxerr)
    return np.concatenate(cores), np.array(ra, dtype=np.int64), np.full(d, n, dtype=np.int64), info


def ones_tt(d, n):
    """The all-ones vector as a rank-1 TT.  Returns (cr, r)."""
    return np.ones(d * n), np.ones(d + 1, dtype=np.int64)


def random_tt(d, n, r, seed):
    """N(0,1) cores of rank r (boundary ranks 1) from PCG64(seed).  Returns (cr, rr)."""
    rng = np.random.default_rng(seed)
    rr = [1] + [r] * (d - 1) + [1]
    cr = np.concatenate([rng.standard_normal(rr[k] * n * rr[k + 1]) for k in range(d)])
    return cr, np.array(rr, dtype=np.int64)

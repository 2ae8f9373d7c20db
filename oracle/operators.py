"""Operators and TT helpers for the oracle (TEST INFRASTRUCTURE).

* `read_sdv` restates the SDV reader (ttio.f90:196-297; header ttio.f90:10-17)
  so the reference's own fixtures (Heisen_20.sdv) can be loaded.
* `heisenberg_mpo`, `tfim_mpo` build TT-matrix cores from the Hamiltonian
  formulas used by the benchmark configurations.  The Heisenberg builder is
  pinned against the reference fixture Heisen_20.sdv (tests).
* `mpo_apply_full` applies a TT-matrix to a full vector (small d) and
  `mpo_to_dense` expands it, for exact-diagonalisation pins.

Layout conventions (the reference's, tt.f90:17; test_eigb.f90:57-63):
TT core k is (r_{k-1}, n_k, r_k) column-major; TT-matrix core k is
(ra_{k-1}, n_k, m_k, ra_k) column-major with n the row (output) index and m
the column (input) index; site 1 is the fastest-varying index of a full
vector.
"""
from __future__ import annotations

import struct

import numpy as np

SX = np.array([[0.0, 0.5], [0.5, 0.0]])
SY_IM = np.array([[0.0, -0.5], [0.5, 0.0]])   # S^y = i * SY_IM
SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])        # S^+
SM = np.array([[0.0, 0.0], [1.0, 0.0]])        # S^-
PX = np.array([[0.0, 1.0], [1.0, 0.0]])        # Pauli X
PZ = np.array([[1.0, 0.0], [0.0, -1.0]])       # Pauli Z
I2 = np.eye(2)


def read_sdv(path):
    """Read an SDV TT file (ttio.f90:196-297).

    Returns dict(l, m, n (array), r (array of length d+1), cores (list of
    column-major flat arrays (r_{k-1}, n_k, r_k)), dtype).
    Header (128 bytes): 'TT      ', ver(2) int32, inf(4) int32, 64-byte
    comment, i(8) int32 (ttio.f90:10-17); then l, m int32; n(l:m) int32;
    r(l-1:m) int32; then the cores, float64 (dtt) or complex128 (ztt).
    """
    b = open(path, "rb").read()
    if b[:2] != b"TT":
        raise ValueError("not a TT file")
    ver = struct.unpack("<2i", b[8:16])
    if ver[0] != 1:
        raise ValueError("unsupported version %r" % (ver,))
    off = 128
    l, m = struct.unpack("<2i", b[off:off + 8]); off += 8
    d = m - l + 1
    n = np.frombuffer(b, dtype="<i4", count=d, offset=off).astype(int); off += 4 * d
    r = np.frombuffer(b, dtype="<i4", count=d + 1, offset=off).astype(int); off += 4 * (d + 1)
    sz = int(np.sum(r[:-1] * n * r[1:]))
    rest = len(b) - off
    if rest == 8 * sz:
        dt = np.float64
    elif rest == 16 * sz:
        dt = np.complex128
    else:
        raise ValueError("core payload size %d does not match %d elements" % (rest, sz))
    x = np.frombuffer(b, dtype=dt, count=sz, offset=off).copy()
    cores, p = [], 0
    for k in range(d):
        c = r[k] * n[k] * r[k + 1]
        cores.append(x[p:p + c])
        p += c
    return dict(l=l, m=m, n=n, r=r, cores=cores, dtype=dt)


def flatten_cores(cores):
    """Concatenate cores into the flat crA / crY layout (compute_ps offsets,
    tt_adapt_als.f90:28-37)."""
    return np.concatenate([np.asarray(c).reshape(-1, order="F") for c in cores])


def _mpo_from_w(d, W, wfirst_row, wlast_col):
    """Cores (ra_{k-1}, n, m, ra_k) from an operator-valued bond matrix W
    (R x R x 2 x 2: W[a1, a2] is a 2x2 operator, row=output, col=input)."""
    R = W.shape[0]
    cores, ra = [], [1]
    for k in range(d):
        if d == 1:
            Wk = W[wfirst_row:wfirst_row + 1, wlast_col:wlast_col + 1]
        elif k == 0:
            Wk = W[wfirst_row:wfirst_row + 1, :]
        elif k == d - 1:
            Wk = W[:, wlast_col:wlast_col + 1]
        else:
            Wk = W
        a1, a2 = Wk.shape[:2]
        c = np.transpose(Wk, (0, 2, 3, 1))     # (a1, i, j, a2)
        cores.append(c.reshape(-1, order="F"))
        ra.append(a2)
    return cores, np.array(ra)


def xxz_bond_matrix(J=1.0, delta=1.0):
    """Lower-triangular W of H = sum J (Sx Sx + Sy Sy + delta Sz Sz)."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0] = I2
    W[1, 0] = SP
    W[2, 0] = SM
    W[3, 0] = SZ
    W[4, 1] = 0.5 * J * SM
    W[4, 2] = 0.5 * J * SP
    W[4, 3] = J * delta * SZ
    W[4, 4] = I2
    return W


def heisenberg_mpo(d, J=1.0, delta=1.0):
    """Open XXZ chain MPO (rank 5), the Hamiltonian of test_eigb's fixture
    Heisen_20.sdv for J=1, delta=1.  Returns (cores, ra, n) with n = m = 2."""
    cores, ra = _mpo_from_w(d, xxz_bond_matrix(J, delta), 4, 0)
    return cores, ra, np.full(d, 2)


def tfim_mpo(d, g=1.0, J=1.0):
    """Transverse-field Ising H = -J sum Z_i Z_{i+1} - g sum X_i (rank 3)."""
    W = np.zeros((3, 3, 2, 2))
    W[0, 0] = I2
    W[1, 0] = PZ
    W[2, 0] = -g * PX
    W[2, 1] = -J * PZ
    W[2, 2] = I2
    cores, ra = _mpo_from_w(d, W, 2, 0)
    return cores, ra, np.full(d, 2)


def mpo_apply_full(cores, ra, n, m, v):
    """Apply a TT-matrix to a full vector (site 1 fastest)."""
    d = len(cores)
    T = np.asarray(v).reshape([1] + [int(mm) for mm in m], order="F")   # (a0, j1..jd)
    # T axes: (a, done_1..done_{k}, j_{k+1}..j_d)
    for k in range(d):
        A = np.reshape(cores[k], (ra[k], n[k], m[k], ra[k + 1]), order="F")
        # contract a_{k} (axis 0) and j_{k+1} (axis k+1)
        T = np.tensordot(A, T, axes=([0, 2], [0, k + 1]))   # (i, a2, done..., rest...)
        # move i to position k+1 and a2 to front
        T = np.moveaxis(T, 1, 0)                            # (a2, i, done..., rest...)
        T = np.moveaxis(T, 1, k + 1)                        # (a2, done..., i, rest...)
    return T.reshape(-1, order="F")


def mpo_to_dense(cores, ra, n, m):
    """Dense matrix of a TT-matrix (small d only)."""
    N = int(np.prod(m))
    M = int(np.prod(n))
    out = np.empty((M, N), dtype=np.result_type(*cores))
    e = np.zeros(N, dtype=out.dtype)
    for j in range(N):
        e[:] = 0
        e[j] = 1
        out[:, j] = mpo_apply_full(cores, ra, n, m, e)
    return out


def heisenberg_sparse(d, J=1.0, delta=1.0):
    """Sparse Hamiltonian from the formula (site 1 fastest), for ED pins."""
    import scipy.sparse as sp
    ops = {}

    def site_op(op, k):
        key = (id(op), k)
        if key not in ops:
            left = sp.identity(2 ** k, format="csr")
            right = sp.identity(2 ** (d - k - 1), format="csr")
            # site 1 fastest => kron(right, op, left)
            ops[key] = sp.kron(sp.kron(right, sp.csr_matrix(op)), left, format="csr")
        return ops[key]

    H = sp.csr_matrix((2 ** d, 2 ** d))
    for k in range(d - 1):
        H = H + J * 0.5 * (site_op(SP, k) @ site_op(SM, k + 1) + site_op(SM, k) @ site_op(SP, k + 1))
        H = H + J * delta * site_op(SZ, k) @ site_op(SZ, k + 1)
    return H


def random_tt(rng, n, r, dtype=np.float64):
    """Random N(0,1) cores with ranks r (length d+1)."""
    cores = []
    for k in range(len(n)):
        c = rng.standard_normal(r[k] * n[k] * r[k + 1])
        if np.issubdtype(dtype, np.complexfloating):
            c = c + 1j * rng.standard_normal(c.size)
        cores.append(c.astype(dtype))
    return cores


def tt_full(cores, n, r):
    """Full tensor (site 1 fastest) of a TT; the last rank index, if > 1, is
    kept as the slowest index (block TT of tt_eigb)."""
    d = len(cores)
    T = np.reshape(cores[0], (r[0] * n[0], r[1]), order="F")
    for k in range(1, d):
        C = np.reshape(cores[k], (r[k], n[k] * r[k + 1]), order="F")
        T = np.reshape(T @ C, (-1, r[k + 1]), order="F")
    return np.reshape(T, (-1, r[d]), order="F") if r[d] > 1 else T.reshape(-1, order="F")


def laplace_dst_solve(d, n, y):
    """Exact solution of the Dirichlet Laplacian system (interior grid of
    [0,1]^d, n points per direction, h = 1/(n+1)) for a full right-hand side
    y (site 1 fastest), through the tensor sine transform: the 1D matrix
    tridiag(-1, 2, -1)/h^2 has eigenvectors S_jk = sqrt(2/(n+1))
    sin(pi (j+1)(k+1)/(n+1)) and eigenvalues (2 - 2 cos(pi (k+1)/(n+1)))/h^2."""
    k = np.arange(1, n + 1)
    S = np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(k, k) / (n + 1))
    lam = (n + 1) ** 2 * (2.0 - 2.0 * np.cos(np.pi * k / (n + 1)))
    T = np.reshape(y, (n,) * d, order="F")
    for ax in range(d):
        T = np.moveaxis(np.tensordot(S, T, axes=(1, ax)), 0, ax)
    den = np.zeros((n,) * d)
    for ax in range(d):
        sh = [1] * d
        sh[ax] = n
        den = den + lam.reshape(sh)
    T = T / den
    for ax in range(d):
        T = np.moveaxis(np.tensordot(S, T, axes=(1, ax)), 0, ax)
    return np.reshape(T, -1, order="F")

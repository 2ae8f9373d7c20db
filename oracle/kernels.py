"""Oracle restatement of the tt-fort local-operator kernels (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.  Every routine follows the reference's own GEMM / transpose
sequence on column-major (Fortran-order) flat arrays, so that intermediate
layouts match the reference exactly.  Arrays are 1-D numpy arrays holding the
column-major data of the reference's Fortran arrays.

Pinning (see tests/test_oracle_kernels.py):
  * bfun3 / phi / full-matrix kernels are checked against the index
    definitions written in the reference comments (ttals.f90:18-29, :159-170,
    :213-224) evaluated by an independent Kronecker-product construction of
    the local matrix, against each other (phi_left of x with y == y' B x), and
    against the exact Heisenberg spectrum from the reference's own MPO
    fixture Heisen_20.sdv.
  * the bfun3 halves (dbfun3_right, dbfun3_left_intended) by contracting
    their open interface with phi1 / phi2 back into the Kronecker-pinned
    bfun3, with enough rows that the contraction is injective.
  * chop rules are checked against hand-worked singular-value cases; the
    residual-rank scan against its closed form (one more than the largest
    candidate above the bound) and against the residuals themselves.
  * QR/SVD against orthogonality and reconstruction invariants.
  * GMRES against dense solves; block-Jacobi against block inverses.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "F", "dtransp", "dperm1324", "drow_add", "drow_cut",
    "dbfun3", "zbfun3", "zbfun3_transp", "dbfun3_right", "dbfun3_left_intended",
    "dphi_left", "dphi_right", "zphi_left", "zphi_right",
    "dphi2_left", "dphi2_right", "d2d_fullmat", "local_matrix_kron",
    "my_chop3", "chop", "residual_rank", "amen_truncation_rank", "dqr", "djac_gen", "djac_apply", "dgmresr_hh",
    "bfun3_flops",
]


def F(a, *shape):
    """View flat column-major data `a` as a Fortran array of `shape`."""
    return np.reshape(a, shape, order="F")


def _flat(a):
    return np.reshape(a, -1, order="F")


def dtransp(n, m, a):
    """matrix_util.f90:152-170 / tt_linalg.f90:82-99: (n x m) -> (m x n)."""
    return _flat(F(a, n, m).T)


def dperm1324(n1, n2, n3, n4, a):
    """tt_linalg.f90:101-119: B(n1,n3,n2,n4) = A(n1,n2,n3,n4)."""
    return _flat(np.transpose(F(a, n1, n2, n3, n4), (0, 2, 1, 3)))


def drow_add(m, n, k, A, B):
    """tt_linalg.f90:54-69: the k rows of B (k, n) appended under A (m, n),
    giving (m + k, n)."""
    return _flat(np.vstack([F(A, m, n), F(B, k, n)]))


def drow_cut(m, n, k, A):
    """tt_linalg.f90:71-79: the first k rows of A (m, n), packed (k, n)."""
    return _flat(F(A, m, n)[:k, :])


def bfun3_flops(rx1, m, rx2, ry1, n, ry2, ra1, ra2, complex_=False):
    """Algorithmic flops of one bfun3 application (the three reference GEMMs,
    ttals.f90:20, :24, :29); x4 for complex arithmetic."""
    f = (2.0 * rx1 * m * rx2 * ra2 * ry2
         + 2.0 * ra1 * n * m * ra2 * ry2 * rx1
         + 2.0 * ry1 * n * ry2 * rx1 * ra1)
    return 4.0 * f if complex_ else f


# ---------------------------------------------------------------- bfun3 family
def dbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x):
    """ttals.f90:9-32 (== tt_linalg.f90:127-159 bind(c), :185-207 dbfun32).

    y(c1,i,c2) = sum phi1(c1,b1,a1) A(a1,i,j,a2) phi2(b2,a2,c2) x(b1,j,b2)
    """
    res1 = F(x, rx1 * m, rx2) @ F(phi2, rx2, ra2 * ry2)            # :20
    res1 = dtransp(rx1, m * ra2 * ry2, _flat(res1))                  # :22
    res2 = F(A, ra1 * n, m * ra2) @ F(res1, m * ra2, ry2 * rx1)      # :24
    res2 = dtransp(ra1 * n * ry2, rx1, _flat(res2))                  # :26
    y = F(phi1, ry1, rx1 * ra1) @ F(res2, rx1 * ra1, n * ry2)        # :29
    return _flat(y)


def zbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x):
    """ttals.f90:34-57: complex bfun3, no conjugation anywhere."""
    return dbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2,
                  np.asarray(phi1, complex), np.asarray(A, complex),
                  np.asarray(phi2, complex), np.asarray(x, complex))


def zbfun3_transp(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x):
    """ttals.f90:85-108: y = B^H x for the local matrix B of zbfun3.

    x has the output shape (ry1,n,ry2); y has the input shape (rx1,m,rx2).
    """
    res1 = F(x, ry1, n * ry2).conj().T @ F(phi1, ry1, rx1 * ra1)       # :95
    res1 = dtransp(n, ry2 * rx1 * ra1, _flat(res1))                   # :98
    res2 = F(res1, ry2 * rx1, ra1 * n) @ F(A, ra1 * n, m * ra2)        # :100
    res2 = dtransp(ry2, rx1 * m * ra2, _flat(res2))                   # :103
    y = F(res2, rx1 * m, ra2 * ry2) @ F(phi2, rx2, ra2 * ry2).T        # :105
    return _flat(y).conj()                                            # :107


def dbfun3_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2, A, phi2, x):
    """tt_linalg.f90:210-240: the first two bfun3 stages only.

    Output layout (rx1*ra1, n, ry2) = (b1,a1,i,c2).
    """
    res1 = F(x, rx1 * m, rx2) @ F(phi2, rx2, ra2 * ry2)                # :223
    res1 = dtransp(rx1, m * ra2 * ry2, _flat(res1))                    # :228
    res2 = F(A, ra1 * n, m * ra2) @ F(res1, m * ra2, ry2 * rx1)        # :232
    return dtransp(ra1 * n * ry2, rx1, _flat(res2))                    # :237


def dbfun3_left_intended(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, x):
    """The contraction tt_linalg.f90:242-256 describes:
    y(c1,i,b2,a2) = sum phi1(c1,b1,a1) A(a1,i,j,a2) x(b1,j,b2)."""
    t = np.einsum("cba,aijk,bjd->cidk",
                  F(phi1, ry1, rx1, ra1), F(A, ra1, n, m, ra2), F(x, rx1, m, rx2))
    return _flat(t)


# ---------------------------------------------------------------- phi updates
def dphi_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi2_old, A, x, y):
    """tt_linalg.f90:259-284 (== ttals.f90:150-173).

    phi2(b1,a1,c1) = sum phi2_old(b2,a2,c2) x(b1,j,b2) A(a1,i,j,a2) y(c1,i,c2)
    Output layout (rx1, ra1, ry1).
    """
    res1 = F(x, rx1 * m, rx2) @ F(phi2_old, rx2, ra2 * ry2)            # :270
    res1 = dtransp(rx1, m * ra2 * ry2, _flat(res1))                    # :272
    res2 = F(A, ra1 * n, m * ra2) @ F(res1, m * ra2, ry2 * rx1)        # :275
    res1 = dtransp(ra1, n * ry2 * rx1, _flat(res2))                    # :277
    phi2 = F(y, ry1, n * ry2) @ F(res1, n * ry2, rx1 * ra1)            # :279
    return dtransp(ry1, rx1 * ra1, _flat(phi2))                        # :281


def dphi_left(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1_old, A, x, y):
    """tt_linalg.f90:287-312 (== ttals.f90:203-227).

    phi1(c2,b2,a2) = sum phi1_old(c1,b1,a1) y(c1,i,c2) A(a1,i,j,a2) x(b1,j,b2)
    Output layout (ry2, rx2, ra2).
    """
    res1 = F(phi1_old, ry1, rx1 * ra1).T @ F(y, ry1, n * ry2)          # :298
    res1 = dtransp(rx1, ra1 * n * ry2, _flat(res1))                    # :300
    res2 = F(res1, ra1 * n, ry2 * rx1).T @ F(A, ra1 * n, m * ra2)      # :303
    res1 = dtransp(ry2, rx1 * m * ra2, _flat(res2))                    # :305
    phi1 = F(res1, rx1 * m, ra2 * ry2).T @ F(x, rx1 * m, rx2)          # :307
    return dtransp(ra2, ry2 * rx2, _flat(phi1))                        # :309


def zphi_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi2_old, A, x, y):
    """ttals.f90:176-200: complex phi_right with conj(y) (:186)."""
    return dphi_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2,
                      np.asarray(phi2_old, complex), np.asarray(A, complex),
                      np.asarray(x, complex), np.conj(np.asarray(y, complex)))


def zphi_left(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1_old, A, x, y):
    """ttals.f90:229-256: complex phi_left with conj(y) (:239)."""
    return dphi_left(rx1, m, rx2, ry1, n, ry2, ra1, ra2,
                     np.asarray(phi1_old, complex), np.asarray(A, complex),
                     np.asarray(x, complex), np.conj(np.asarray(y, complex)))


def dphi2_right(rx1, rx2, ry1, n, ry2, phi2_old, x, y):
    """tt_linalg.f90:315-327: phi2(b1,c1) = sum x(b1,j,b2) phi2_old(b2,c2) y(c1,j,c2).

    Output layout (rx1, ry1).
    """
    res1 = F(x, rx1 * n, rx2) @ F(phi2_old, rx2, ry2)                  # :324
    phi2 = F(res1, rx1, n * ry2) @ F(y, ry1, n * ry2).T                # :326
    return _flat(phi2)


def dphi2_left(rx1, rx2, ry1, n, ry2, phi1_old, x, y):
    """tt_linalg.f90:329-340: phi1(c2,b2) = sum phi1_old(c1,b1) y(c1,i,c2) x(b1,i,b2).

    Output layout (ry2, rx2).
    """
    res1 = F(phi1_old, ry1, rx1).T @ F(y, ry1, n * ry2)                # :337
    phi1 = F(res1, rx1 * n, ry2).T @ F(x, rx1 * n, rx2)                # :339
    return _flat(phi1)


# ---------------------------------------------------------------- full matrix
def d2d_fullmat(rx1, m, rx2, ra1, ra2, phi1, A, phi2):
    """tt_eigb.f90:88-105: explicit local matrix B((c1,i,c2),(b1,j,b2)).

    Returned as a dense (N x N) Fortran-order matrix, N = rx1*m*rx2, with rows
    indexing y and columns indexing x.
    """
    res1 = F(phi1, rx1 * rx1, ra1) @ F(A, ra1, m * m * ra2)            # :97
    res2 = dperm1324(rx1, rx1, m, m * ra2, _flat(res1))                # :99
    phi2t = dtransp(rx2, ra2 * rx2, phi2)                              # :101
    res1 = F(res2, rx1 * m * rx1 * m, ra2) @ F(phi2t, ra2, rx2 * rx2)  # :102
    B = dperm1324(rx1 * m, rx1 * m, rx2, rx2, _flat(res1))             # :103
    N = rx1 * m * rx2
    return F(B, N, N)


def local_matrix_kron(ry1, n, ry2, rx1, m, rx2, ra1, ra2, phi1, A, phi2):
    """Independent construction of the local matrix used to pin the kernels:
    B = sum_{a1,a2} phi2[:, a2, :]^T (x) A[a1,:,:,a2] (x) phi1[:, :, a1]
    (Kronecker order matches the column-major multi-index (c1,i,c2))."""
    P1 = F(phi1, ry1, rx1, ra1)
    AA = F(A, ra1, n, m, ra2)
    P2 = F(phi2, rx2, ra2, ry2)
    dt = np.result_type(P1, AA, P2)
    B = np.zeros((ry1 * n * ry2, rx1 * m * rx2), dtype=dt)
    for a1 in range(ra1):
        for a2 in range(ra2):
            B += np.kron(P2[:, a2, :].T, np.kron(AA[a1, :, :, a2], P1[:, :, a1]))
    return B


# ---------------------------------------------------------------- truncation
def my_chop3(s, eps):
    """tt_adapt_als.f90:40-61 (== matrix_util.f90:254-275).

    Smallest r such that sum_{k>r} s_k^2 <= (eps*||s||)^2, r >= 1.
    """
    s = np.asarray(s, float)
    n = s.size
    nrm = np.sqrt(np.sum(s * s))
    nrm = (nrm * eps) * (nrm * eps)
    cursum = 0.0
    i = n
    while i > 0:
        cursum += s[i - 1] * s[i - 1]
        if cursum > nrm:
            break
        i -= 1
    return min(i + 1, n)


def residual_rank(err_of, k, res_new, eps2):
    """tt_adapt_als.f90:819-833 (forward) and :937-950 (backward): the
    residual-based truncation rank.  err_of(c) is the relative local residual
    of the core truncated to rank c.  The DO loop runs nn = k-1 down to 1
    and exits at the first nn with err > 2 res_new and err > eps2; after a
    loop that runs out nn is 0; the rank is nn + 1."""
    nn = k - 1
    while nn >= 1:
        err = err_of(nn)
        if err > res_new * 2.0 and err > eps2:
            break
        nn -= 1
    return max(nn, 0) + 1


def amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs, res_new, eps2):
    """One local step's residual-based rank search (tt_adapt_als.f90:809-833
    for direction +1, :928-950 for -1): SVD of the local core unfolded as
    (rx1 n) x rx2 (forward) or rx1 x (n rx2) (backward), candidates
    U(:, :c) diag(s[:c]) V(:, :c)^T, their local residuals through dbfun3,
    and the scan of `residual_rank`.  Returns (rank, errs) with errs[c] the
    residual of candidate c for every c = 1..k-1 (errs[0] = -1)."""
    rows, cols = (rx1 * n, rx2) if direction > 0 else (rx1, n * rx2)
    U, s, VT = np.linalg.svd(F(x, rows, cols), full_matrices=False)
    k = s.size
    shape = (rx1, n, rx2, rx1, n, rx2, ra1, ra2)
    nrhs = np.linalg.norm(rhs)
    errs = np.full(k, -1.0)
    for c in range(1, k):
        cand = (U[:, :c] * s[:c]) @ VT[:c, :]
        errs[c] = np.linalg.norm(dbfun3(*shape, phi1, A, phi2, _flat(cand)) - rhs) / nrhs
    return residual_rank(lambda c: errs[c], k, res_new, eps2), errs


def chop(s, tol=None, rmax=None):
    """svd.f90:86-111: Frobenius chop with optional rank cap.

    Returns (r, err) with err = sqrt of the discarded sum of squares.
    """
    s = np.asarray(s, float)
    r = s.size
    er2 = 0.0
    if rmax is not None and rmax < r:
        er2 = float(np.dot(s[rmax:r], s[rmax:r]))
        r = rmax
    if tol is not None:
        nrm = np.sqrt(np.sum(s * s))
        bound = tol * tol * nrm * nrm
        er = er2 + s[r - 1] * s[r - 1]
        while er < bound:
            er2 = er
            r -= 1
            er = er + s[r - 1] * s[r - 1]
    return r, float(np.sqrt(er2))


def dqr(n, m, A):
    """tt_linalg.f90:30-51 / matrix_util.f90:37-61: thin Householder QR.

    Returns (Q flat n x min(n,m), R flat min(n,m) x m).  LAPACK's sign
    convention (dgeqrf) is kept by numpy.linalg.qr, which calls the same
    routines; comparisons should still be made up to column signs.
    """
    Q, R = np.linalg.qr(F(A, n, m), mode="reduced")
    return _flat(Q), _flat(R)


# ---------------------------------------------------------------- block Jacobi
def djac_gen(ptype, rx1, n, rx2, ra1, ra2, Phi1, A, Phi2):
    """tt_linalg.f90:363-495: block-Jacobi preconditioner blocks.

    'c': rx2*rx1 blocks of n x n, jacs laid out (n, n, rx2, rx1) (:397-429);
    'l': rx2 blocks of (rx1*n)^2 (:434-462); 'r': rx1 blocks of (rx2*n)^2
    (:464-491).  Each block is the inverse of the matching diagonal block of
    the local matrix restricted as described by the reference.
    """
    P1 = F(Phi1, rx1, rx1, ra1)
    AA = F(A, ra1, n, n, ra2)
    P2 = F(Phi2, rx2, ra2, rx2)
    p = ptype.lower()
    if p == "c":
        d1 = np.stack([np.diag(P1[:, :, a]) for a in range(ra1)], axis=1)   # (rx1, ra1)
        d2 = np.stack([np.diag(P2[:, a, :]) for a in range(ra2)], axis=1)   # (rx2, ra2)
        # blk(b1, i, j, b2) = sum d1(b1,a1) A(a1,i,j,a2) d2(b2,a2)   (:412-416)
        blk = np.einsum("pa,aijk,qk->pijq", d1, AA, d2)
        out = np.empty((n, n, rx2, rx1))
        for b1 in range(rx1):
            for b2 in range(rx2):
                out[:, :, b2, b1] = np.linalg.inv(blk[b1, :, :, b2])
        return _flat(out)
    if p == "l":
        d2 = np.stack([np.diag(P2[:, a, :]) for a in range(ra2)], axis=1)
        # blk((c1,i),(b1,j), b2) = sum Phi1(c1,b1,a1) A(a1,i,j,a2) d2(b2,a2)
        blk = np.einsum("cba,aijk,qk->cibjq", P1, AA, d2)
        out = np.empty((rx1 * n, rx1 * n, rx2))
        for b2 in range(rx2):
            out[:, :, b2] = np.linalg.inv(np.reshape(blk[:, :, :, :, b2], (rx1 * n, rx1 * n), order="F"))
        return _flat(out)
    if p == "r":
        d1 = np.stack([np.diag(P1[:, :, a]) for a in range(ra1)], axis=1)
        # blk((i,c2),(j,b2), b1) = sum d1(b1,a1) A(a1,i,j,a2) Phi2(b2,a2,c2)
        blk = np.einsum("pa,aijk,dke->iejdp", d1, AA, P2)
        out = np.empty((n * rx2, n * rx2, rx1))
        for b1 in range(rx1):
            out[:, :, b1] = np.linalg.inv(np.reshape(blk[:, :, :, :, b1], (n * rx2, n * rx2), order="F"))
        return _flat(out)
    raise ValueError(ptype)


def djac_apply(ptype, rx1, n, rx2, jacs, x):
    """tt_linalg.f90:497-533: y = P x with the block-Jacobi blocks."""
    p = ptype.lower()
    X = F(x, rx1, n, rx2)
    if p == "c":
        J = F(jacs, n, n, rx2, rx1)
        Y = np.einsum("ijqp,pjq->piq", J, X)
        return _flat(Y)
    if p == "l":
        J = F(jacs, rx1 * n, rx1 * n, rx2)
        Xm = F(x, rx1 * n, rx2)
        return _flat(np.einsum("abq,bq->aq", J, Xm))
    if p == "r":
        J = F(jacs, n * rx2, n * rx2, rx1)
        Xt = np.reshape(np.transpose(X, (1, 2, 0)), (n * rx2, rx1), order="F")
        Y = np.einsum("abp,bp->ap", J, Xt)
        return _flat(np.transpose(F(_flat(Y), n, rx2, rx1), (2, 0, 1)))
    if p == "n":
        return np.array(x, copy=True)
    raise ValueError(ptype)


# ---------------------------------------------------------------- duchol
def duchol(trans, m, n, dA, numvec):
    """tt_linalg.f90:540-643 (duchol_fort): pivoted incomplete Cholesky of
    the Gram matrix of the n items of dA (columns of an m x n dA for trans
    'N', rows of an n x m dA for 'T'), with the reference's per-step
    re-orthogonalisation (dgeqrf + dorgqr, :613-617) and eigen-rotation of
    the accumulated directions (:619-631).

    Returns (du (n x real), lam (numvec), real).
    """
    A = F(dA, m, n) if trans in "nN" else F(dA, n, m)
    ddiag = (A * A).sum(0) if trans in "nN" else (A * A).sum(1)                # :576-586
    du = np.zeros((n, numvec))
    lam = np.zeros(numvec)
    real = numvec
    for cnt in range(numvec):
        maxi = int(np.argmax(np.abs(ddiag)))                                    # :590 idamax
        du[:, cnt] = A.T @ A[:, maxi] if trans in "nN" else A @ A[maxi, :]      # :600-604
        for j in range(cnt):                                                    # :606-609
            du[:, cnt] -= lam[j] * du[maxi, j] * du[:, j]
        maxa = np.sqrt(abs(du[maxi, cnt]))
        if maxa < 1e-300:                                                       # :611-613
            real = cnt + 1
            break
        du[:, cnt] /= maxa
        ddiag = ddiag - du[:, cnt] ** 2
        Q, Rr = np.linalg.qr(du[:, :cnt + 1])                                   # :620-624
        dx = Rr[:, cnt].copy()
        du[:, :cnt + 1] = Q
        dv = np.outer(dx, dx) + np.diag(lam[:cnt + 1])                          # :627-632
        w, V = np.linalg.eigh(dv)                                               # :634 dsyev
        lam[:cnt + 1] = w
        du[:, :cnt + 1] = du[:, :cnt + 1] @ V                                   # :637-638
    return du[:, :real], lam, real


# ---------------------------------------------------------------- GMRES
def dgmresr_hh(matvec, rhs, nrestart, tol, niters, prec=None):
    """tt_linalg.f90:651-845: right-preconditioned restarted GMRES with
    Householder Arnoldi.  Solves B P z = rhs and returns z (the caller applies
    P, as tt_adapt_als.f90:776-781 does).  `matvec(v)` applies the local
    operator B; `prec(v)` applies P (None for ptype 'n').

    Returns (sol, info) where info = dict(it, j, relres).
    """
    sz = rhs.size
    nrestart = min(nrestart, sz)
    U = np.zeros((sz, nrestart + 1))
    tau = np.zeros(nrestart + 1)
    JJ = np.zeros((2, nrestart))
    R = np.zeros((nrestart, nrestart))
    sol = np.zeros(sz)
    P = prec if prec is not None else (lambda v: v)
    curres = np.inf
    nrmrhs = None
    it = 0
    j = 0
    for it in range(niters):
        w = matvec(P(sol)) - rhs                                         # :686-695
        nrmr = np.linalg.norm(w)
        if it == 0:
            nrmrhs = nrmr
        if nrmr == 0.0:
            break
        dbeta = nrmr if w[0] >= 0 else -nrmr                              # :709-712
        w[0] += dbeta
        tau[:] = 0.0
        tau[0] = -dbeta
        w /= np.linalg.norm(w)
        U[:, 0] = w
        jdone = nrestart
        for j in range(1, nrestart + 1):                                  # :720
            w = -2.0 * U[j - 1, j - 1] * U[:, j - 1]                      # :722-725
            w[j - 1] += 1.0
            for i in range(j - 1, 0, -1):                                 # :726-729
                w += -2.0 * np.dot(U[:, i - 1], w) * U[:, i - 1]
            w /= np.linalg.norm(w)                                        # :730-732
            w = matvec(P(w))                                              # :735-743
            for i in range(1, j + 1):                                     # :746-749
                w += -2.0 * np.dot(U[:, i - 1], w) * U[:, i - 1]
            if j < sz:                                                    # :752-773
                U[:j, j] = 0.0
                U[j:, j] = w[j:]
                dalpha = np.linalg.norm(U[j:, j])
                if dalpha != 0.0:
                    if w[j] < 0.0:
                        dalpha = -dalpha
                    U[j, j] += dalpha
                    U[j:, j] /= np.linalg.norm(U[j:, j])
                    w[j] = -dalpha
                    w[j + 1:] = 0.0
            for i in range(1, j):                                         # :776-780
                db = w[i - 1]
                w[i - 1] = JJ[0, i - 1] * w[i - 1] + JJ[1, i - 1] * w[i]
                w[i] = -JJ[1, i - 1] * db + JJ[0, i - 1] * w[i]
            if j < sz:                                                    # :783-792
                dalpha = np.hypot(w[j - 1], w[j])
                JJ[0, j - 1] = w[j - 1] / dalpha
                JJ[1, j - 1] = w[j] / dalpha
                tau[j] = -JJ[1, j - 1] * tau[j - 1]
                tau[j - 1] = JJ[0, j - 1] * tau[j - 1]
                w[j - 1] = dalpha
                w[j] = 0.0
            R[:j, j - 1] = w[:j]                                          # :794
            curres = abs(tau[j]) / nrmrhs                                 # :797
            if curres < tol:
                jdone = j
                break
        j = min(jdone, nrestart)                                          # :808-814
        t = np.linalg.solve(np.triu(R[:j, :j]), tau[:j])                  # :816 dtrsv
        tau[:j] = t
        w = -2.0 * U[j - 1, j - 1] * tau[j - 1] * U[:, j - 1]             # :819-821
        w[j - 1] += tau[j - 1]
        for i in range(j - 1, 0, -1):                                     # :823-827
            w[i - 1] += tau[i - 1]
            w += -2.0 * np.dot(U[:, i - 1], w) * U[:, i - 1]
        sol -= w                                                          # :828-829
        if curres < tol:
            break
    return sol, dict(it=it, j=j, relres=curres)


# ---------------------------------------------------------------- Krylov exponential
def zgpadm(ideg, H, t):
    """expokit.f:767-935 (ZGPADM): exp(t H) for a small complex matrix by the
    irreducible (ideg, ideg) Pade form +/-(I + 2 q^-1 p) with
    scaling-and-squaring.  Returns (E, ns)."""
    H = np.asarray(H, dtype=np.complex128)
    m = H.shape[0]
    hnorm = abs(t * np.abs(H).sum(axis=1).max())                         # :833-842 (max row sum)
    if hnorm == 0.0:
        raise ValueError("null H in ZGPADM")
    ns = max(0, int(np.log(hnorm) / np.log(2.0)) + 2)                    # :844
    scale = t / 2.0 ** ns
    scale2 = scale * scale
    c = np.zeros(ideg + 1)                                               # :849-854
    c[0] = 1.0
    i, j = ideg + 1, 2 * ideg + 1
    for k in range(1, ideg + 1):
        c[k] = c[k - 1] * (i - k) / (k * (j - k))
    H2 = scale2 * (H @ H)                                                # :858
    p = c[ideg - 1] * np.eye(m, dtype=np.complex128)                     # :862-871
    q = c[ideg] * np.eye(m, dtype=np.complex128)
    odd = 1
    k = ideg - 1
    while k > 0:                                                         # :875-889
        used = q if odd else p
        f = used @ H2
        f = f + c[k - 1] * np.eye(m)
        if odd:
            q = f
        else:
            p = f
        odd = 1 - odd
        k -= 1
    if odd:                                                              # :893-901
        q = scale * (q @ H)
    else:
        p = scale * (p @ H)
    q = q - p                                                            # :902
    p = np.linalg.solve(q, p)                                            # :903 zgesv
    p = 2.0 * p + np.eye(m)                                              # :905-908
    if ns == 0 and odd:                                                  # :910-913
        return -p, ns
    E = p
    for _ in range(ns):                                                  # :917-923
        E = E @ E
    return E, ns


def _nint(x):
    """Fortran NINT: round half away from zero."""
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


def zgexpv(matvec, v, m, t, tol, anorm, mxstep=500):
    """expokit.f:2403-2791 (ZGEXPV): w = exp(t A) v by adaptive Krylov
    steps (Arnoldi basis of size m, Pade on the (m+2)-augmented Hessenberg,
    local error estimate and step-size control).  Returns (w, info)."""
    v = np.asarray(v, dtype=np.complex128)
    n = v.size
    eps = np.finfo(float).eps
    if tol <= eps:                                                       # :2591-2593
        tol = np.sqrt(eps)
    rndoff = eps * anorm
    break_tol = 1.0e-7                                                   # :2596
    delta, gamma = 1.2, 0.9                                              # :2505-2510
    sgn = np.sign(t) if t != 0 else 1.0
    t_out = abs(t)
    w = v.copy()
    beta = np.linalg.norm(w)
    vnorm = beta
    hump = beta
    SQR1 = np.sqrt(0.1)                                                  # :2609
    xm = 1.0 / m
    p2 = tol * (((m + 1) / 2.72) ** (m + 1)) * np.sqrt(2.0 * 3.14 * (m + 1))
    t_new = (1.0 / anorm) * (p2 / (4.0 * beta * anorm)) ** xm
    p1 = 10.0 ** (_nint(np.log10(t_new) - SQR1) - 1)
    t_new = np.trunc(t_new / p1 + 0.55) * p1
    t_now, nstep, nmult = 0.0, 0, 0
    mh = m + 2
    while t_now < t_out:                                                 # :2618
        nstep += 1
        t_step = min(t_out - t_now, t_new)
        V = np.zeros((n, m + 1), dtype=np.complex128)
        H = np.zeros((mh, mh), dtype=np.complex128)
        V[:, 0] = w / beta
        k1, mbrkdwn = 2, m
        avnorm = 0.0
        for j in range(1, m + 1):                                        # :2632-2657 Arnoldi (MGS)
            nmult += 1
            p = matvec(V[:, j - 1])
            for i in range(1, j + 1):
                hij = np.vdot(V[:, i - 1], p)
                p = p - hij * V[:, i - 1]
                H[i - 1, j - 1] = hij
            hj1j = np.linalg.norm(p)
            if hj1j <= break_tol:                                        # :2645-2654 happy breakdown
                k1, mbrkdwn = 0, j
                t_step = t_out - t_now
                break
            H[j, j - 1] = hj1j
            V[:, j] = p / hj1j
        else:
            nmult += 1
            avnorm = np.linalg.norm(matvec(V[:, m]))                     # :2658-2660
        H[m + 1, m] = 1.0                                                # :2666
        while True:                                                      # :2670-2716
            mx = mbrkdwn + k1
            E, _ = zgpadm(6, H[:mx, :mx], sgn * t_step)
            if k1 == 0:
                err_loc = tol
            else:
                p1 = abs(E[m, 0]) * beta
                p2 = abs(E[m + 1, 0]) * beta * avnorm
                if p1 > 10.0 * p2:
                    err_loc, xm = p2, 1.0 / m
                elif p1 > p2:
                    err_loc, xm = (p1 * p2) / (p1 - p2), 1.0 / m
                else:
                    err_loc, xm = p1, 1.0 / (m - 1)
            if k1 != 0 and err_loc > delta * t_step * tol:               # :2702-2716
                t_step = gamma * t_step * (t_step * tol / err_loc) ** xm
                p1 = 10.0 ** (_nint(np.log10(t_step) - SQR1) - 1)
                t_step = np.trunc(t_step / p1 + 0.55) * p1
                continue
            break
        mx = mbrkdwn + max(0, k1 - 1)                                    # :2720-2723
        w = beta * (V[:, :mx] @ E[:mx, 0])
        beta = np.linalg.norm(w)
        hump = max(hump, beta)
        t_new = gamma * t_step * (t_step * tol / err_loc) ** xm          # :2727-2729
        p1 = 10.0 ** (_nint(np.log10(t_new) - SQR1) - 1)
        t_new = np.trunc(t_new / p1 + 0.55) * p1
        t_now += t_step
        if nstep >= mxstep:
            break
    return w, dict(nmult=nmult, nstep=nstep, hump=hump / vnorm)


def zexp_mv(matvec, v, m, tau, tol, anorm):
    """explib.f90:53-125 (zexp_mv): w = exp(tau A) v through ZGEXPV on the
    normalised vector, with the 1 x 1 case done directly."""
    v = np.asarray(v, dtype=np.complex128)
    n = v.size
    m1 = min(m, n)
    if m1 == 1:                                                          # :103-105
        a = matvec(np.ones(1, dtype=np.complex128))[0]
        return v * np.exp(a * tau), dict(nmult=1, nstep=1)
    nrm = np.linalg.norm(v)
    if nrm == 0.0:
        return np.zeros_like(v), dict(nmult=0, nstep=0)
    w, info = zgexpv(matvec, v / nrm, m1, tau, tol, anorm)
    return w * nrm, info


def zsfun(rx1, ry1, ra, rx2, ry2, phi1, phi2, S):
    """tt_ksl.f90:130-136 (zsfun_matvec): S (rx1 x rx2) -> (ry1 x ry2),
    Sy(c1, c2) = sum phi1(c1, b1, a) S(b1, b2) phi2(b2, a, c2), phi1 viewed
    (ry1, rx1 ra) and phi2 (rx2, ra ry2)."""
    res1 = F(S, rx1, rx2) @ F(phi2, rx2, ra * ry2)                       # (rx1, ra*ry2)
    res1 = np.reshape(res1, (rx1 * ra, ry2), order="F")
    return _flat(F(phi1, ry1, rx1 * ra) @ res1)


def onenorm_dense(matvec, n, dtype=np.complex128):
    """Exact 1-norm of the operator through n products with unit vectors
    (the reference's usenrm = 0 mode, normest.f90:112-140)."""
    best = 0.0
    for j in range(n):
        e = np.zeros(n, dtype=dtype)
        e[j] = 1.0
        best = max(best, np.abs(matvec(e)).sum())
    return best

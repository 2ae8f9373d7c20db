"""Oracle restatements of the tt-fort sweep drivers (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.

tt_eigb   -- tt_eigb.f90:110-436, block eigensolver sweep.  The local
             eigensolver the reference calls is PRIMME (DYNAMIC method) with
             the same matvec (dbfun3); PRIMME is vendored C that is not
             available to numpy, so the restatement uses a dense symmetric
             eigensolver for small local problems (as the reference does
             below max_full_size, tt_eigb.f90:234-249) and ARPACK Lanczos
             (scipy.sparse.linalg.eigsh) on the same matvec otherwise, with
             the reference's tolerance eps2/resid_damp.  Eigenvalues are
             pinned by exact diagonalisation; sweep trajectories are not
             reproducible even between two runs of the reference (PRIMME's
             DYNAMIC switching is timing dependent), so parity is judged on
             eigenvalues and residuals.
ztt_ksl   -- tt_ksl.f90:431-668, the projector-splitting time step.
tt_mvk4   -- tt_adapt_als.f90:67-500, the approximate TT-matrix by TT-vector
             product by alternating projections with enrichment; pinned
             against exact products on full vectors.
"""
from __future__ import annotations

import numpy as np

from . import kernels as K
from .kernels import F, _flat


def _svd(M):
    U, s, VT = np.linalg.svd(M, full_matrices=False)
    return U, s, VT


def local_eigs(shape, phi1, A, phi2, x0, B, tol, max_full_size):
    """Lowest B eigenpairs of the local operator (tt_eigb.f90:234-292)."""
    rx1, m, rx2 = shape[:3]
    N = rx1 * m * rx2
    if N < max_full_size or N <= 3 * B + 2:
        Bm = K.d2d_fullmat(rx1, m, rx2, shape[6], shape[7], phi1, A, phi2)
        w, V = np.linalg.eigh(0.5 * (Bm + Bm.T))
        return w[:B], V[:, :B], 0
    import scipy.sparse.linalg as sla
    count = [0]

    def mv(v):
        count[0] += 1
        return K.dbfun3(*shape, phi1, A, phi2, np.asarray(v, float).reshape(-1))

    op = sla.LinearOperator((N, N), matvec=mv, dtype=float)
    w, V = sla.eigsh(op, k=B, which="SA", tol=tol, v0=x0[:, 0] if x0 is not None else None,
                     ncv=min(N, max(2 * B + 1, 20)))
    idx = np.argsort(w)
    return w[idx], V[:, idx], count[0]


def tt_eigb(d, n, ra, crA, crY0, ry, eps, rmax, B, nswp=20, max_full_size=2000, verb=0, start_site=None,
            max_steps=None, time_budget=None):
    """tt_eigb.f90:110-436 restated.  Returns (lambda, cores, ry, stats).

    For timing a bounded sample of a sweep (not part of the reference):
    start_site = k0 (0-based) moves the block index from the last core to
    core k0 through the backward SVD steps without local solves, so that
    the sweep proper begins at core k0; max_steps / time_budget (seconds)
    stop it after that many local steps or once that much time has passed.
    stats['steps'] then lists (local size, seconds, matvecs) per local step
    (eigensolve, truncation and interface update)."""
    n = [int(v) for v in n]
    ra = [int(v) for v in ra]
    ry = [int(v) for v in ry]
    assert ry[d] == B
    eps2 = eps / np.sqrt(d)
    resid_damp = 10.0
    pa = [0]
    for k in range(d):
        pa.append(pa[-1] + ra[k] * n[k] * n[k] * ra[k + 1])
    Acore = [np.asarray(crA[pa[k]:pa[k + 1]], float) for k in range(d)]
    cores, p = [], 0
    for k in range(d):
        sz = ry[k] * n[k] * ry[k + 1]
        cores.append(np.array(crY0[p:p + sz], float))
        p += sz
    phi = [None] * (d + 1)
    phi[0] = np.ones(1)
    phi[d] = np.ones(1)
    # initial QR + phi_left (tt_eigb.f90:201-217)
    for k in range(d):
        mrow, ncol = ry[k] * n[k], ry[k + 1]
        Q, R = np.linalg.qr(K.F(cores[k], mrow, ncol))
        rnew = Q.shape[1]
        cores[k] = Q.reshape(-1, order="F")
        if k < d - 1:
            nxt = R @ K.F(cores[k + 1], ncol, n[k + 1] * ry[k + 2])
            cores[k + 1] = nxt.reshape(-1, order="F")
            ry[k + 1] = rnew
            phi[k + 1] = K.dphi_left(ry[k], n[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1],
                                     phi[k], Acore[k], cores[k], cores[k])
    ry[d] = 1
    k, direction, swp = d - 1, -1, 1
    ermax, fvold, total_mv = 0.0, 0.0, 0
    lam = np.zeros(B)
    last = 0
    steps = []
    dry = start_site is not None and start_site < d - 1
    import time as _time
    t_begin = None
    while swp <= nswp:
        shape = (ry[k], n[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1])
        N = ry[k] * n[k] * ry[k + 1]
        x0 = K.F(cores[k], N, B)
        if dry and k == start_site:
            dry = False
        if dry:
            X, mv = x0, 0   # moving the block index only
        else:
            if t_begin is None:
                t_begin = _time.perf_counter()
            t_step = _time.perf_counter()
            lam, X, mv = local_eigs(shape, phi[k], Acore[k], phi[k + 1], x0, B, eps2 / resid_damp,
                                    max_full_size)
        total_mv += mv
        if not dry:
            fv = lam.sum()
            ermax = max(ermax, (fvold - fv) / abs(fv))
            fvold = fv
        if verb:
            print("swp %d i [%d/%d] N %d ermax %.3e" % (swp, k + 1, d, N, ermax))
        if direction < 0 and k > 0:
            M = np.reshape(X.T.reshape(-1, order="F"), (B * ry[k], n[k] * ry[k + 1]), order="F")
            U, s, VT = _svd(M)
            nn = min(K.my_chop3(s, eps2), rmax)
            cores[k] = VT[:nn, :].reshape(-1, order="F")
            US = U[:, :nn] * s[:nn]                                   # (b, r_k) x j
            US = np.transpose(US.reshape(B, ry[k], nn, order="F"), (1, 2, 0))   # r_k, j, b
            left = K.F(cores[k - 1], ry[k - 1] * n[k - 1], ry[k])
            cores[k - 1] = (left @ US.reshape(ry[k], nn * B, order="F")).reshape(-1, order="F")
            ry[k] = nn
            phi[k] = K.dphi_right(ry[k], n[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1],
                                  phi[k + 1], Acore[k], cores[k], cores[k])
        elif direction > 0 and k < d - 1:
            M = np.reshape(X.reshape(-1, order="F"), (ry[k] * n[k], ry[k + 1] * B), order="F")
            U, s, VT = _svd(M)
            nn = min(K.my_chop3(s, eps2), rmax)
            cores[k] = U[:, :nn].reshape(-1, order="F")
            SVT = (VT[:nn, :] * s[:nn, None]).reshape(nn, ry[k + 1], B, order="F")
            right = K.F(cores[k + 1], ry[k + 1], n[k + 1] * ry[k + 2])
            out = np.stack([SVT[:, :, b] @ right for b in range(B)], axis=2)   # j, (n r), b
            cores[k + 1] = out.reshape(-1, order="F")
            ry[k + 1] = nn
            phi[k + 1] = K.dphi_left(ry[k], n[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1],
                                     phi[k], Acore[k], cores[k], cores[k])
        else:
            cores[k] = X.reshape(-1, order="F")
        if t_begin is not None and not dry:
            steps.append((N, _time.perf_counter() - t_step, mv))
            if (max_steps is not None and len(steps) >= max_steps) or \
                    (time_budget is not None and _time.perf_counter() - t_begin >= time_budget):
                break
        if direction > 0 and k == d - 2:
            direction, k = -1, d - 1
            last = swp
            swp += 1
            if ermax < eps:
                break
            ermax = 0.0
        elif direction < 0 and k == 1:
            direction, k = 1, 0
        else:
            k += direction
    ry[d] = B
    return lam, cores, ry, dict(sweeps=last, total_matvecs=total_mv, steps=steps)


def _zqr(A):
    """zqr (matrix_util.f90:64-90): thin Householder QR, Q (m x k), R (k x n)."""
    Q, R = np.linalg.qr(A, mode="reduced")
    return Q, R


def ztt_ksl(d, n, ra, crA, crY0, ry, tau, order=8, typ=2, tol=1e-8):
    """tt_ksl.f90:431-668 (ztt_ksl): one step of the projector-splitting KSL
    integrator for dY/dt = A Y on fixed TT ranks.  typ 2 (default) is the
    symmetrised scheme (half steps tau/2 right-to-left then left-to-right),
    typ 1 one right-to-left pass.  Local exponentials by zexp_mv with Krylov
    size `order` and tolerance 1e-8; the operator-norm estimate is the exact
    1-norm of the local operator (the reference's usenrm = 0 mode; the
    default Higham estimate only steers the first step size).

    Returns (cores list, ry)."""
    ry = list(int(r) for r in ry)
    pa = np.concatenate([[0], np.cumsum([ra[k] * n[k] * n[k] * ra[k + 1] for k in range(d)])]).astype(int)
    cr, off = [], 0
    for k in range(d):
        sz = ry[k] * n[k] * ry[k + 1]
        cr.append(np.array(crY0[off:off + sz], dtype=np.complex128))
        off += sz
    phi = [None] * (d + 1)
    phi[0] = np.ones(1, dtype=np.complex128)
    phi[d] = np.ones(1, dtype=np.complex128)
    Ak = [np.asarray(crA[pa[k]:pa[k + 1]], dtype=np.complex128) for k in range(d)]
    # initial left-to-right QR and left interfaces (:538-556)
    for i in range(d - 1):
        M = F(cr[i], ry[i] * n[i], ry[i + 1])
        Q, R = _zqr(M)
        rnew = Q.shape[1]
        cr[i] = _flat(Q)
        cr[i + 1] = _flat(R @ F(cr[i + 1], ry[i + 1], n[i + 1] * ry[i + 2]))
        ry[i + 1] = rnew
        phi[i + 1] = K.zphi_left(ry[i], n[i], ry[i + 1], ry[i], n[i], ry[i + 1], ra[i], ra[i + 1],
                                 phi[i], Ak[i], cr[i], cr[i])
    tau0 = tau if typ == 1 else tau / 2.0                                # :566-570

    def local(i, v, t):
        sh = (ry[i], n[i], ry[i + 1], ry[i], n[i], ry[i + 1], ra[i], ra[i + 1])
        mv = lambda x: K.zbfun3(*sh, phi[i], Ak[i], phi[i + 1], x)
        an = K.onenorm_dense(mv, v.size)
        return K.zexp_mv(mv, v, order, t, tol, an)[0]

    def sstep(rx1, ry1, ra_, rx2, ry2, p1, p2, S, t):
        mv = lambda x: K.zsfun(rx1, ry1, ra_, rx2, ry2, p1, p2, x)
        an = K.onenorm_dense(mv, S.size)
        return K.zexp_mv(mv, S, order, t, tol, an)[0]

    i, dir_ = d - 1, -1
    while True:                                                          # :575-651
        if dir_ < 0:
            cur = local(i, cr[i], tau0)
            if i > 0:
                M = F(cur, ry[i], n[i] * ry[i + 1]).T                    # (n r2) x r1
                Q, R = _zqr(M)
                rnew = Q.shape[1]
                cr[i] = _flat(Q.T)
                Rt = R.T                                                 # ry_i x rnew
                phitmp = K.zphi_right(rnew, n[i], ry[i + 1], rnew, n[i], ry[i + 1], ra[i], ra[i + 1],
                                      phi[i + 1], Ak[i], cr[i], cr[i])
                St = sstep(ry[i], ry[i], ra[i], rnew, rnew, phi[i], phitmp, _flat(Rt), -tau0)
                cr[i - 1] = _flat(F(cr[i - 1], ry[i - 1] * n[i - 1], ry[i]) @ F(St, ry[i], rnew))
                ry[i] = rnew
                phi[i] = phitmp
            else:
                cr[i] = cur
        else:
            cur = local(i, cr[i], tau0)
            if i < d - 1:
                M = F(cur, ry[i] * n[i], ry[i + 1])
                Q, R = _zqr(M)
                rnew = Q.shape[1]
                cr[i] = _flat(Q)
                phitmp = K.zphi_left(ry[i], n[i], rnew, ry[i], n[i], rnew, ra[i], ra[i + 1],
                                     phi[i], Ak[i], cr[i], cr[i])
                St = sstep(rnew, rnew, ra[i + 1], ry[i + 1], ry[i + 1], phitmp, phi[i + 1], _flat(R), -tau0)
                cr[i + 1] = _flat(F(St, rnew, ry[i + 1]) @ F(cr[i + 1], ry[i + 1], n[i + 1] * ry[i + 2]))
                ry[i + 1] = rnew
                phi[i + 1] = phitmp
            else:
                cr[i] = cur
        if dir_ > 0 and i == d - 1:                                      # :637-651
            break
        if dir_ < 0 and i == 0:
            if typ == 1:
                break
            dir_ = 1
            continue
        i += dir_
    return cr, ry


def tt_mvk4(d, n, m, rx, ra, crA, crX, crY0, ry, eps, kickrank=5, nswp=20):
    """tt_adapt_als.f90:67-500 (tt_mvk4): y ~ A x for a TT-matrix A (cores
    (ra_k, n_k, m_k, ra_k+1)) and a TT vector x (cores (rx_k, m_k, rx_k+1)),
    by alternating projections started from y0 (cores (ry_k, n_k, ry_k+1)).

    As the reference executes it:
      * y0 is left-orthogonalised by QR (:197-222), building the left
        interfaces with dphi_left;
      * each step sets the core to its projection dbfun3(phi_k, A_k,
        phi_k+1, x_k) and records its relative change (:243-255);
      * left-to-right steps truncate by SVD with my_chop3 at eps/sqrt(d)
        (QR when kickrank = 0) and move the factor right (:269-341);
      * right-to-left steps truncate likewise and enrich the kept basis
        with min(kickrank, n ry) directions from duchol of the partial
        product dbfun3_right(A_k, phi_k+1, x_k), then re-orthogonalise
        (:349-426);
      * a sweep ends at the last core, and the run stops when the largest
        change in the sweep is below eps (:428-447).
    The reference's rmax sizes its work arrays only; no rank is capped
    (ranks beyond rmax overrun those arrays there), and none is here.

    Returns (cores of y as a flat array, ry, sweeps).
    """
    n = [int(v) for v in n[:d]]
    m = [int(v) for v in m[:d]]
    rx = [int(v) for v in rx[:d + 1]]
    ra = [int(v) for v in ra[:d + 1]]
    ry = [int(v) for v in ry[:d + 1]]
    eps2 = eps / np.sqrt(d)
    pa = np.concatenate([[0], np.cumsum([ra[k] * n[k] * m[k] * ra[k + 1] for k in range(d)])])
    px = np.concatenate([[0], np.cumsum([rx[k] * m[k] * rx[k + 1] for k in range(d)])])
    A = [np.asarray(crA[pa[k]:pa[k + 1]], dtype=float) for k in range(d)]
    X = [np.asarray(crX[px[k]:px[k + 1]], dtype=float) for k in range(d)]
    cr, p = [], 0
    for k in range(d):
        sz = ry[k] * n[k] * ry[k + 1]
        cr.append(np.array(crY0[p:p + sz], dtype=float))
        p += sz
    phi = [None] * (d + 1)
    phi[0] = np.ones(1)
    phi[d] = np.ones(1)

    def shape(k):
        return (rx[k], m[k], rx[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1])

    # left-orthogonalisation of y0 (:197-222)
    for k in range(d - 1):
        rnew = min(ry[k] * n[k], ry[k + 1])
        Q, R = np.linalg.qr(F(cr[k], ry[k] * n[k], ry[k + 1]), mode="reduced")
        cr[k + 1] = _flat(R @ F(cr[k + 1], ry[k + 1], n[k + 1] * ry[k + 2]))
        cr[k] = _flat(Q)
        ry[k + 1] = rnew
        phi[k + 1] = K.dphi_left(*shape(k), phi[k], A[k], X[k], cr[k])

    i, direction, swp, err_max = d - 1, -1, 1, 0.0
    while swp <= nswp:
        cur = K.dbfun3(*shape(i), phi[i], A[i], phi[i + 1], X[i])        # :243
        err = np.linalg.norm(cr[i] - cur) / np.linalg.norm(cur)         # :253-255
        err_max = max(err_max, err)
        if direction > 0 and i < d - 1:                                 # :269-341
            M = F(cur, ry[i] * n[i], ry[i + 1])
            if kickrank > 0:
                U, s, VT = _svd(M)
                nn = K.my_chop3(s, eps2)
                Q, R = U[:, :nn], s[:nn, None] * VT[:nn, :]
            else:
                Q, R = np.linalg.qr(M, mode="reduced")
            rnew = Q.shape[1]
            cr[i] = _flat(Q)
            cr[i + 1] = _flat(R @ F(cr[i + 1], ry[i + 1], n[i + 1] * ry[i + 2]))
            ry[i + 1] = rnew
            phi[i + 1] = K.dphi_left(*shape(i), phi[i], A[i], X[i], cr[i])
        if direction < 0 and i > 0:                                     # :349-426
            M = F(cur, ry[i], n[i] * ry[i + 1])
            if kickrank > 0:
                U, s, VT = _svd(M)
                nn = K.my_chop3(s, eps2)
                V = VT[:nn, :].T                                        # (n ry2, nn)
                R = s[:nn, None] * U[:, :nn].T                          # (nn, ry1)
                z = K.dbfun3_right(*shape(i), A[i], phi[i + 1], X[i])   # (rx1 ra1, n ry2)
                du, _, mm = K.duchol("N", rx[i] * ra[i], n[i] * ry[i + 1], z,
                                     min(kickrank, n[i] * ry[i + 1]))
                W = np.hstack([V, du[:, :mm]])
                R = np.vstack([R, np.zeros((mm, ry[i]))])
                Q, R2 = np.linalg.qr(W, mode="reduced")
                R = R2 @ R
            else:
                Q, R = np.linalg.qr(M.T, mode="reduced")
            rnew = Q.shape[1]
            cr[i] = _flat(Q.T)                                          # (rnew, n ry2)
            cr[i - 1] = _flat(F(cr[i - 1], ry[i - 1] * n[i - 1], ry[i]) @ R.T)
            ry[i] = rnew
            phi[i] = K.dphi_right(*shape(i), phi[i + 1], A[i], X[i], cr[i])
        if direction > 0 and i == d - 1:                                # :428-447
            cr[i] = cur
            direction, i, swp = -1, i + 1, swp + 1
            if err_max < eps:
                break
            err_max = 0.0
        if direction < 0 and i == 0:                                    # :448-461
            cr[i] = cur
            direction, i = 1, i - 1
        i += direction
    return np.concatenate(cr), np.array(ry), swp - 1

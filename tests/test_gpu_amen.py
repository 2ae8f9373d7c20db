"""GPU parity of the AMEn path (through the C ABI) against the oracle and
against exact solutions.

* block-Jacobi blocks and their application: 1e-10 relative to the oracle;
* local GMRES: the returned correction meets the requested relative
  residual and agrees with the oracle's Householder GMRES to the tolerance;
* duchol kick directions: eigenvalues to 1e-9 relative and the spanned
  projector to 1e-9 (eigenvectors are defined up to sign);
* the half contractions: 1e-12 relative;
* tt_amen_solve on the Dirichlet Laplacian: the TT solution against the
  tensor-DST exact solution (relative error <= 1e-6 for eps = 1e-8) and the
  full residual against the dense operator.
"""
import numpy as np
import pytest

import mpo
from oracle import kernels as K
from oracle import operators as O

pytestmark = pytest.mark.gpu


def _local_problem(rng, rx1, n, rx2, ra1, ra2, spd=True):
    P1 = rng.standard_normal((rx1, rx1, ra1))
    P2 = rng.standard_normal((rx2, ra2, rx2))
    if spd:   # symmetric interfaces with a dominant identity part
        P1 = 0.5 * (P1 + np.transpose(P1, (1, 0, 2)))
        P2 = 0.5 * (P2 + np.transpose(P2, (2, 1, 0)))
        P1[:, :, 0] = np.eye(rx1) * 2 + 0.1 * P1[:, :, 0]
        P2[:, 0, :] = np.eye(rx2) * 2 + 0.1 * P2[:, 0, :]
    A = 0.2 * rng.standard_normal((ra1, n, n, ra2))
    A[0, :, :, 0] += 4 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    return K._flat(P1), K._flat(A), K._flat(P2)


@pytest.mark.parametrize("ptype,rx1,n,rx2", [("c", 5, 6, 4), ("c", 16, 64, 16), ("l", 4, 5, 3), ("r", 3, 5, 6)])
def test_djac(ttlib, ptype, rx1, n, rx2):
    import ttfort
    rng = np.random.default_rng(rx1 * 100 + n)
    ra1, ra2 = 3, 2
    phi1, A, phi2 = _local_problem(rng, rx1, n, rx2, ra1, ra2)
    J = ttfort.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    Jr = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    assert np.allclose(J, Jr, rtol=1e-10, atol=1e-12 * np.abs(Jr).max())
    x = rng.standard_normal(rx1 * n * rx2)
    y = ttfort.djac_apply(ptype, rx1, n, rx2, J, x)
    assert np.allclose(y, K.djac_apply(ptype, rx1, n, rx2, Jr, x), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("ptype,rx1,n,rx2,shift", [("c", 8, 64, 8, 0.0), ("l", 4, 5, 3, 0.0), ("c", 6, 12, 5, -6.0)])
def test_djac_symmetric_blocks(ttlib, ptype, rx1, n, rx2, shift):
    """Blocks of a symmetric operator are exactly symmetric and take the
    register-tiled inverse without pivoting when positive definite; with a
    negative shift (symmetric, nonsingular, indefinite blocks) the pivoting
    Gauss-Jordan takes them back.  Either way the inverses match the
    oracle's (LU-based) to 1e-10."""
    import ttfort
    rng = np.random.default_rng(rx1 * 31 + n)
    ra1, ra2 = 3, 2
    P1 = rng.standard_normal((rx1, rx1, ra1))
    P2 = rng.standard_normal((rx2, ra2, rx2))
    P1 = 0.5 * (P1 + np.transpose(P1, (1, 0, 2)))
    P2 = 0.5 * (P2 + np.transpose(P2, (2, 1, 0)))
    P1[:, :, 0] = np.eye(rx1) * 2 + 0.1 * P1[:, :, 0]
    P2[:, 0, :] = np.eye(rx2) * 2 + 0.1 * P2[:, 0, :]
    A = 0.2 * rng.standard_normal((ra1, n, n, ra2))
    A = 0.5 * (A + np.transpose(A, (0, 2, 1, 3)))
    A[0, :, :, 0] += 4 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1) + shift * np.eye(n)
    phi1, Af, phi2 = K._flat(P1), K._flat(A), K._flat(P2)
    J = ttfort.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, Af, phi2)
    Jr = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, Af, phi2)
    assert np.allclose(J, Jr, rtol=1e-10, atol=1e-12 * np.abs(Jr).max())


@pytest.mark.parametrize("ptype", ["n", "c"])
def test_gmres(ttlib, ptype):
    import ttfort
    rng = np.random.default_rng(5)
    rx1, n, rx2, ra1, ra2 = 6, 8, 5, 2, 2
    phi1, A, phi2 = _local_problem(rng, rx1, n, rx2, ra1, ra2)
    N = rx1 * n * rx2
    rhs = rng.standard_normal(N)
    jacs = ttfort.djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, A, phi2) if ptype != "n" else None
    tol = 1e-10
    sol, it, relres = ttfort.gmres(phi1, A, phi2, rhs, rx1, n, rx2, ra1, ra2, 40, tol, 4, ptype, jacs)
    B = lambda v: K.dbfun3(rx1, n, rx2, rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    P = (lambda v: K.djac_apply(ptype, rx1, n, rx2, jacs, v)) if ptype != "n" else None
    Pv = P(sol) if P else sol
    true_res = np.linalg.norm(B(Pv) - rhs) / np.linalg.norm(rhs)
    assert relres < tol and true_res < 10 * tol, (relres, true_res)
    ref, info = K.dgmresr_hh(B, rhs, 40, tol, 4, prec=P)
    assert np.linalg.norm(sol - ref) <= 1e-7 * np.linalg.norm(ref)


def test_gmres_long_vectors(ttlib):
    """GMRES on 102 400-entry Krylov vectors: the fused Gram-Schmidt then
    works on row chunks of 512 (200 work-groups) rather than one row per
    thread; same checks as test_gmres."""
    import ttfort
    rng = np.random.default_rng(55)
    rx1, n, rx2, ra1, ra2 = 40, 64, 40, 2, 2
    # _local_problem's interface noise grows like sqrt(rank) and would swamp
    # the identity part at rank 40: scaled down so that the system is well
    # conditioned and GMRES converges
    P1 = rng.standard_normal((rx1, rx1, ra1)) / np.sqrt(rx1)
    P2 = rng.standard_normal((rx2, ra2, rx2)) / np.sqrt(rx2)
    P1 = 0.5 * (P1 + np.transpose(P1, (1, 0, 2)))
    P2 = 0.5 * (P2 + np.transpose(P2, (2, 1, 0)))
    P1[:, :, 0] = np.eye(rx1) * 2 + 0.1 * P1[:, :, 0]
    P2[:, 0, :] = np.eye(rx2) * 2 + 0.1 * P2[:, 0, :]
    Am = 0.2 * rng.standard_normal((ra1, n, n, ra2))
    Am[0, :, :, 0] += 4 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    phi1, A, phi2 = K._flat(P1), K._flat(Am), K._flat(P2)
    N = rx1 * n * rx2
    rhs = rng.standard_normal(N)
    jacs = ttfort.djac_gen("c", rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    tol = 1e-10
    sol, it, relres = ttfort.gmres(phi1, A, phi2, rhs, rx1, n, rx2, ra1, ra2, 40, tol, 4, "c", jacs)
    B = lambda v: K.dbfun3(rx1, n, rx2, rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    P = lambda v: K.djac_apply("c", rx1, n, rx2, jacs, v)
    true_res = np.linalg.norm(B(P(sol)) - rhs) / np.linalg.norm(rhs)
    assert relres < tol and true_res < 10 * tol, (relres, true_res)
    ref, info = K.dgmresr_hh(B, rhs, 40, tol, 4, prec=P)
    assert np.linalg.norm(sol - ref) <= 1e-7 * np.linalg.norm(ref)


@pytest.mark.parametrize("trans,m,n,numvec", [("N", 30, 50, 4), ("T", 40, 64, 4), ("N", 7, 5, 3)])
def test_duchol(ttlib, trans, m, n, numvec):
    import ttfort
    rng = np.random.default_rng(m + n)
    # items graded in size so that the pivot order is well defined
    if trans == "N":
        M = rng.standard_normal((m, n)) * np.logspace(0, -3, n)[None, :]
    else:
        M = rng.standard_normal((n, m)) * np.logspace(0, -3, n)[:, None]
    dA = M.reshape(-1, order="F")
    du, lam, real = ttfort.duchol(trans, m, n, dA, numvec)
    du_r, lam_r, real_r = K.duchol(trans, m, n, dA, numvec)
    assert real == real_r
    assert np.allclose(lam[:real], lam_r[:real], rtol=1e-9)
    assert np.allclose(du @ du.T, du_r @ du_r.T, atol=1e-9)


@pytest.mark.parametrize("side", ["L", "R"])
def test_bfun3_half(ttlib, side):
    import ttfort
    rng = np.random.default_rng(9)
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = 5, 3, 4, 5, 3, 4, 3, 2
    A = rng.standard_normal(ra1 * n * m * ra2)
    x = rng.standard_normal(rx1 * m * rx2)
    if side == "L":
        phi = rng.standard_normal(ry1 * rx1 * ra1)
        y = ttfort.bfun3_half("L", rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi, A, x)
        ref = K.dbfun3_left_intended(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi, A, x)
    else:
        phi = rng.standard_normal(rx2 * ra2 * ry2)
        y = ttfort.bfun3_half("R", rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi, A, x)
        ref = K.dbfun3_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2, A, phi, x)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def _cores(cr, r, n, m=None):
    out, off = [], 0
    for k in range(len(n)):
        sz = int(r[k] * n[k] * (m[k] if m is not None else 1) * r[k + 1])
        out.append(cr[off:off + sz])
        off += sz
    return out


@pytest.mark.parametrize("d,n,prec,kick", [(4, 8, "n", 4), (4, 8, "c", 4), (3, 12, "n", 2), (5, 10, "c", 4)])
def test_amen_laplace(ttlib, d, n, prec, kick):
    import ttfort
    crA, ra, nn = mpo.laplace_dd(d, n)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    eps = 1e-8
    res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, eps, kickrank=kick, prec=prec, nswp=20)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    y = np.ones(n ** d)
    x_ref = O.laplace_dst_solve(d, n, y)
    assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref), (np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref), st)
    Ax = O.mpo_apply_full(_cores(crA, ra, nn, nn), ra, nn, nn, x)
    assert np.linalg.norm(Ax - y) <= 1e-6 * np.linalg.norm(y)
    assert st["sweeps"] >= 1 and st["last_res_max"] < eps


@pytest.mark.parametrize("d,n,prec", [(3, 12, "c"), (4, 8, "n"), (4, 8, "c")])
def test_amen_diffusion(ttlib, d, n, prec):
    """Variable-coefficient 7-point diffusion (rank-3 coefficient, operator
    rank 6) against a dense solve of the same TT-matrix, and the FP64 TT
    residual of the returned solution."""
    import ttfort
    import ttops
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=5, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    eps = 1e-8
    res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, eps, kickrank=4, prec=prec, nswp=20)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    A = O.mpo_to_dense(_cores(crA, ra, nn, nn), ra, nn, nn)
    x_ref = np.linalg.solve(A, np.ones(n ** d))
    assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref), (np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref), st)
    assert ttops.residual(crA, ra, nn, res, rx_out, crY, ry) <= 1e-6
    assert st["flops"] > 0 and st["last_res_max"] < eps


@pytest.mark.parametrize("prec", ["c", "n"])
def test_amen_mixed_precision(ttlib, prec):
    """Local GMRES with fp32 Krylov matvecs and FP64 restart residuals reaches
    the FP64 solver's tolerance and solution."""
    import ttfort
    import ttops
    d, n = 4, 12
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=8, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=3)
    eps = 1e-8
    A = O.mpo_to_dense(_cores(crA, ra, nn, nn), ra, nn, nn)
    x_ref = np.linalg.solve(A, np.ones(n ** d))
    try:
        ttfort.set_local_precision(32)
        res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, eps, kickrank=4, prec=prec, nswp=20)
    finally:
        ttfort.set_local_precision(64)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    assert st["last_res_max"] < eps
    assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    assert ttops.residual(crA, ra, nn, res, rx_out, crY, ry) <= 1e-6


@pytest.mark.parametrize("direction,rx1,n,rx2", [(1, 6, 4, 20), (-1, 20, 4, 6)])
def test_amen_truncation_rank_matches_oracle(ttlib, direction, rx1, n, rx2):
    """The device residual-rank search of one local step (SVD, candidate
    cores, block matvecs, scan) picks the oracle's rank on the same inputs,
    and the candidate residuals it evaluated match the oracle's.  The bound
    2 res_new is put halfway (geometrically) between two candidate residuals
    that lie in the second batch of the device scan, so the decision is
    clear of rounding; 19 candidates span three batches of 8."""
    import ttfort
    rng = np.random.default_rng(31 + direction)
    ra1, ra2 = 3, 2
    phi1 = rng.standard_normal(rx1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * n * ra2)
    phi2 = rng.standard_normal(rx2 * ra2 * rx2)
    rows, cols = (rx1 * n, rx2) if direction > 0 else (rx1, n * rx2)
    k = min(rows, cols)
    Uq, _ = np.linalg.qr(rng.standard_normal((rows, k)))
    Vq, _ = np.linalg.qr(rng.standard_normal((cols, k)))
    s = 10.0 ** (-0.5 * np.arange(k))
    x = ((Uq * s) @ Vq.T).reshape(-1, order="F")
    shape = (rx1, n, rx2, rx1, n, rx2, ra1, ra2)
    rhs = K.dbfun3(*shape, phi1, A, phi2, x)
    _, errs_o = K.amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs, 1.0, 0.0)
    for j in (12, 9, 3):
        bound = np.sqrt(errs_o[j] * errs_o[j + 1])
        res_new, eps2 = bound / 2, 1e-14
        r_o, _ = K.amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs, res_new, eps2)
        r_d, errs_d = ttfort.amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs,
                                                  res_new, eps2)
        assert r_d == r_o == j + 1, (r_d, r_o, j)
        seen = errs_d >= 0
        assert seen[j] and seen[k - 1] and not seen[0]
        assert np.allclose(errs_d[seen], errs_o[seen], rtol=1e-6, atol=1e-13), (errs_d[seen], errs_o[seen])


def test_amen_rank_cap_without_rmax(ttlib):
    """rmax <= 0 caps the ranks at the library's capacity (256 here): with a
    right-hand side of rank 256 on the middle bond, whose solution needs
    more, the truncation keeps 256 columns of a 576 x 256 unfolding and the
    kick must add none beyond the cap."""
    import ttfort
    d, n = 4, 24
    crA, ra, nn = mpo.laplace_dd(d, n)
    ry = [1, n, 256, n, 1]
    rng = np.random.default_rng(41)
    crY = np.concatenate([rng.standard_normal(ry[k] * n * ry[k + 1]) for k in range(d)])
    rx = list(ry)
    crX = np.concatenate([rng.standard_normal(rx[k] * n * rx[k + 1]) for k in range(d)])
    res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, 1e-12, kickrank=4, prec="n", nswp=2,
                                           rmax=0)
    assert max(rx_out) == 256 and st["max_rank"] == 256, (list(rx_out), st["max_rank"])

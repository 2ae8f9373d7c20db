"""GPU dense linear algebra (through the C ABI) against numpy/LAPACK.

SVD: singular values to 1e-12 relative + 1e-14*smax absolute (LAPACK's own
guarantee is absolute, so the comparison is two-sided), orthonormal factors
to 1e-12, reconstruction to 1e-12 relative.  Small symmetric eigensolver:
eigenvalues to 1e-12*||A||, eigen-residual and orthonormality to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_svd(A, U, s, VT, rel=1e-12):
    m, n = A.shape
    k = min(m, n)
    assert U.shape == (m, k) and VT.shape == (k, n) and s.shape == (k,)
    assert np.all(np.diff(s) <= 1e-15 * max(s[0], 1e-300))
    s_ref = np.linalg.svd(A, compute_uv=False)
    smax = max(s_ref[0], 1e-300)
    assert np.all(np.abs(s - s_ref) <= rel * s_ref + 1e-14 * smax), np.abs(s - s_ref).max()
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-12)
    assert np.allclose(VT @ VT.T, np.eye(k), atol=1e-12)
    assert np.linalg.norm(A - (U * s) @ VT) <= 1e-12 * max(np.linalg.norm(A), 1e-300)


@pytest.mark.parametrize("m,n", [(12, 5), (5, 12), (33, 1), (1, 7), (200, 80), (80, 200),
                                 (256, 640), (640, 256), (64, 64), (1000, 17)])
def test_gesvd_random(ttlib, m, n):
    import ttfort
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    U, s, VT, sw = ttfort.gesvd(A)
    _check_svd(A, U, s, VT)
    assert sw <= 15, sw


def test_gesvd_rank_deficient_and_zero(ttlib):
    import ttfort
    rng = np.random.default_rng(5)
    A = rng.standard_normal((100, 10)) @ rng.standard_normal((10, 50))
    U, s, VT, sw = ttfort.gesvd(A)
    _check_svd(A, U, s, VT)
    assert s[10] < 1e-12 * s[0]
    assert sw <= 15, sw
    Z = np.zeros((20, 6))
    U, s, VT, _ = ttfort.gesvd(Z)
    assert np.all(s == 0)
    assert np.allclose(U.T @ U, np.eye(6), atol=1e-12)


def test_gesvd_graded_relative_accuracy(ttlib):
    """Graded singular values 1 .. 1e-14: one-sided Jacobi keeps them to high
    relative accuracy; the truncation tail sums depend on the small ones."""
    import ttfort
    rng = np.random.default_rng(6)
    m, n = 120, 40
    Qa, _ = np.linalg.qr(rng.standard_normal((m, n)))
    Qb, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s_true = np.logspace(0, -14, n)
    A = (Qa * s_true) @ Qb.T
    U, s, VT, _ = ttfort.gesvd(A)
    assert np.all(np.abs(s - s_true) <= 1e-9 * s_true + 2e-16 * n), np.max(np.abs(s - s_true) / s_true)
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 5, 24, 31, 32, 33, 48, 64])
def test_syev_small(ttlib, n):
    """Orders up to 32 take the slot-ordered Jacobi (one 2x2 block per
    thread), 33 and above the double-buffered circle one; odd orders are
    padded with a zero row."""
    import ttfort
    rng = np.random.default_rng(n)
    As = []
    for _ in range(3):
        X = rng.standard_normal((n, n))
        As.append(X + X.T)
    w, Q = ttfort.syev_small(np.array(As))
    for a, wi, qi in zip(As, w, Q):
        nrm = np.linalg.norm(a, 2)
        assert np.allclose(wi, np.linalg.eigvalsh(a), atol=1e-12 * nrm)
        assert np.linalg.norm(a @ qi - qi * wi) <= 1e-12 * nrm * n
        assert np.allclose(qi.T @ qi, np.eye(n), atol=1e-12)
    wd, _ = ttfort.syev_small(np.array(As), desc=True)
    assert np.allclose(wd, w[:, ::-1])


def test_gram(ttlib):
    import ttfort
    rng = np.random.default_rng(7)
    for N, p, q in ((10, 3, 4), (5000, 24, 24), (32768, 8, 16), (100, 64, 64)):
        X = rng.standard_normal((N, p))
        Y = rng.standard_normal((N, q))
        G = ttfort.gram(X, Y)
        assert np.allclose(G, X.T @ Y, rtol=1e-12, atol=1e-10)


def _graded(rng, m, n, lo):
    k = min(m, n)
    Qa, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Qb, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return (Qa * np.logspace(0, lo, k)) @ Qb.T


@pytest.mark.parametrize("m,n,kind", [(640, 256, "rand"), (256, 640, "rand"), (640, 256, "graded"),
                                      (256, 640, "graded"), (200, 200, "graded"), (1000, 17, "graded"),
                                      (17, 1000, "rand"), (33, 1, "rand"), (1, 7, "rand"),
                                      (128, 320, "lowrank"), (4096, 64, "graded"), (64, 4096, "rand"),
                                      (2048, 100, "lowrank"), (4096, 128, "rand")])
@pytest.mark.parametrize("side", ["L", "R"])
def test_svd_side(ttlib, m, n, kind, side):
    """One-sided SVD used by the truncations: singular values to LAPACK's
    absolute accuracy (5e-13 * smax: thousands of rotations per column), orthonormal vectors for the nonzero
    values, and the factorisation A = O (O^T A) (side L) or A = (A O) O^T
    (side R) to 1e-12 relative; sweeps stay few for graded spectra (the
    QR-preconditioned side)."""
    import ttfort
    rng = np.random.default_rng(m * 7 + n * 13 + (side == "R"))
    if kind == "rand":
        A = rng.standard_normal((m, n))
    elif kind == "graded":
        A = _graded(rng, m, n, -13)
    else:
        A = rng.standard_normal((m, 40)) @ rng.standard_normal((40, n))
    s, O, kpos, sw = ttfort.svd_side(A, side)
    k = min(m, n)
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs(s - s_ref)) <= 5e-13 * s_ref[0], np.max(np.abs(s - s_ref)) / s_ref[0]
    if kind == "lowrank":
        assert kpos <= 41 and kpos >= 40
    else:
        assert kpos == k
    Ok = O[:, :kpos]
    assert np.allclose(Ok.T @ Ok, np.eye(kpos), atol=1e-12)
    assert np.all(O[:, kpos:] == 0)
    rec = Ok @ (Ok.T @ A) if side == "L" else (A @ Ok) @ Ok.T
    assert np.linalg.norm(A - rec) <= 1e-12 * np.linalg.norm(A)
    short = (side == "L") == (m < n)
    if short and kind == "graded":
        assert sw <= 15, sw


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("lo", [-6, -10])
def test_svd_side_interleaved_graded_blocks(ttlib, side, lo):
    """An unfolding whose short index runs over two graded blocks, like a
    sweep core's (r_k, n_k) index with r_k fastest after a truncation: rows
    (r, s) weighted by w_r (geometric from 1 to 10^lo) times 1 or 0.7, times
    Gaussian rows.  In its natural order the preconditioning QR meets a large
    column after 128 small ones and the Jacobi iteration on R^T takes 21-32
    sweeps (host emulation of the same schedule); with the columns sorted by
    norm before the QR it takes 4-5.  Accuracy as in test_svd_side."""
    import ttfort
    rng = np.random.default_rng(3 - lo)
    w = np.logspace(0, lo, 128)
    M = np.concatenate([w, 0.7 * w])[:, None] * rng.standard_normal((256, 640))
    A = M if side == "L" else M.T
    s, O, kpos, sw = ttfort.svd_side(A, side)
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.max(np.abs(s - s_ref)) <= 5e-13 * s_ref[0]
    assert kpos == 256
    assert np.allclose(O.T @ O, np.eye(256), atol=1e-12)
    rec = O @ (O.T @ A) if side == "L" else (A @ O) @ O.T
    assert np.linalg.norm(A - rec) <= 1e-12 * np.linalg.norm(A)
    assert sw <= 8, sw

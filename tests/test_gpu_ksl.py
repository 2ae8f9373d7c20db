"""GPU parity of the KSL path (through the C ABI).

* ttf_zexpm (Pade, one work-group) against scipy.linalg.expm: 1e-12;
* ttf_zexp_bfun3 (device Krylov exponential of the local operator) against
  the dense exponential of the explicit local matrix: 1e-8 relative
  (ZGEXPV tolerance 1e-8), and the identity known answer;
* ttf_zqr: orthonormal Q and Q R = A to 1e-12, also for rank-deficient and
  wide blocks;
* ttf_ztt_ksl on the reference fixtures: the identity case exactly
  (e^{i tau} Y), test_ksl.dat against the oracle's restatement (same
  algorithm, 1e-7) and the exact evolution (2e-6, the fixed-rank bound),
  and a full-rank TFIM chain against the exact evolution (1e-7).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import drivers as D
from oracle import kernels as K
from oracle import operators as O

from golden.ksl_fixture import read_ksl_dat
from test_oracle_ksl import _split, tfim_ksl_problem

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("m", [1, 3, 10, 32, 40])
@pytest.mark.parametrize("t", [0.01, 0.7, -2.5])
def test_zexpm(ttlib, m, t):
    import ttfort
    rng = np.random.default_rng(m)
    H = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    E = ttfort.zexpm(H, t)
    ref = sl.expm(t * H)
    assert np.abs(E - ref).max() <= 1e-12 * np.abs(ref).max()


def _zlocal(rng, rx1, n, rx2, ra1, ra2):
    P1 = rng.standard_normal((rx1, rx1, ra1)) + 1j * rng.standard_normal((rx1, rx1, ra1))
    P2 = rng.standard_normal((rx2, ra2, rx2)) + 1j * rng.standard_normal((rx2, ra2, rx2))
    A = 0.3 * (rng.standard_normal((ra1, n, n, ra2)) + 1j * rng.standard_normal((ra1, n, n, ra2)))
    return K._flat(P1), K._flat(A), K._flat(P2)


@pytest.mark.parametrize("shape,order,tau", [((4, 3, 5, 2, 2), 8, 0.3), ((6, 2, 6, 3, 3), 30, 1.0),
                                             ((8, 2, 8, 3, 3), 8, -0.7)])
def test_zexp_local(ttlib, shape, order, tau):
    import ttfort
    rx1, n, rx2, ra1, ra2 = shape
    rng = np.random.default_rng(rx1 * 7 + order)
    phi1, A, phi2 = _zlocal(rng, rx1, n, rx2, ra1, ra2)
    N = rx1 * n * rx2
    v = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    w, anorm, nmult = ttfort.zexp_bfun3(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, tau, order=order)
    B = np.stack([K.zbfun3(rx1, n, rx2, rx1, n, rx2, ra1, ra2, phi1, A, phi2, e) for e in np.eye(N)], axis=1)
    ref = sl.expm(tau * B) @ v
    assert np.linalg.norm(w - ref) <= 1e-7 * np.linalg.norm(ref), (np.linalg.norm(w - ref) / np.linalg.norm(ref), nmult)
    assert 0 < anorm <= np.abs(B).sum(axis=0).max() * (1 + 1e-12)   # Hager's estimate is a lower bound


def test_zexp_identity(ttlib):
    """A = I (phi = 1, A-core = identity): exp(t I) v = e^t v (test_exp.f90)."""
    import ttfort
    n = 4
    v = np.full(n, 1j)
    w, _, _ = ttfort.zexp_bfun3(1, n, 1, 1, 1, np.ones(1), np.eye(n).reshape(-1, order="F"), np.ones(1), v, 1e-2,
                                order=4)
    assert np.allclose(w, np.exp(1e-2) * v, rtol=1e-13)


@pytest.mark.parametrize("rows,cols,rank", [(128, 64, 64), (128, 64, 9), (2, 64, 2), (64, 64, 1), (96, 3, 3)])
def test_zqr(ttlib, rows, cols, rank):
    import ttfort
    rng = np.random.default_rng(rows + cols + rank)
    A = (rng.standard_normal((rows, rank)) + 1j * rng.standard_normal((rows, rank))) @ (
        rng.standard_normal((rank, cols)) + 1j * rng.standard_normal((rank, cols)))
    Q, R = ttfort.zqr(A)
    k = min(rows, cols)
    assert np.allclose(Q.conj().T @ Q, np.eye(k), atol=1e-12)
    assert np.linalg.norm(Q @ R - A) <= 1e-12 * np.linalg.norm(A)


def _run_fixture(name):
    import ttfort
    f = read_ksl_dat(os.path.join(GOLDEN, name))
    res, ry, st = ttfort.ztt_ksl(f["n"], f["ra"], f["A"], f["Y"], f["ry"], f["tau"])
    x = O.tt_full(_split(res, ry, f["n"]), f["n"], ry)
    y0 = O.tt_full(_split(f["Y"], f["ry"], f["n"]), f["n"], f["ry"])
    return f, x, y0, st


def test_ksl_identity_fixture(ttlib):
    f, x, y0, st = _run_fixture("test_eye_ksl.dat")
    assert np.allclose(x, np.exp(1j * f["tau"]) * y0, rtol=1e-10, atol=1e-12)


def test_ksl_fixture(ttlib):
    f, x, y0, st = _run_fixture("test_ksl.dat")
    cr, ry = D.ztt_ksl(f["d"], f["n"], f["ra"], f["A"], f["Y"], f["ry"], f["tau"])
    xo = O.tt_full(cr, f["n"], ry)
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    Ad = O.mpo_to_dense(_split(f["A"], f["ra"], f["n"], f["n"]), f["ra"], f["n"], f["n"])
    ex = sl.expm(f["tau"] * Ad) @ y0
    assert np.linalg.norm(x - ex) <= 2e-6 * np.linalg.norm(ex)
    assert abs(np.linalg.norm(x) - np.linalg.norm(y0)) <= 1e-9 * np.linalg.norm(y0)


@pytest.mark.parametrize("d,typ", [(6, 2), (8, 1), (8, 2)])
def test_ksl_full_rank_tfim(ttlib, d, typ):
    import ttfort
    tau = 0.05
    crA, ra, n, crY, ry = tfim_ksl_problem(d, 4)
    res, ry2, st = ttfort.ztt_ksl(n, ra, crA, crY, ry, tau, typ=typ)
    x = O.tt_full(_split(res, ry2, n), n, ry2)
    y0 = O.tt_full(_split(crY, ry, n), n, ry)
    Ad = O.mpo_to_dense(_split(crA, ra, n, n), ra, n, n)
    ex = sl.expm(tau * Ad) @ y0
    assert np.linalg.norm(x - ex) <= 1e-7 * np.linalg.norm(ex)


def test_ksl_padded_product_start(ttlib):
    """The fixed-rank start of the TFIM benchmark: the all-up product state
    zero-padded to full rank (d = 8, ranks 2^k).  The padded directions are
    completed by the device QR; at full rank the projector-splitting step is
    exact, so several steps must match exp(t A) y0 and keep the energy
    -(d-1) of the all-up state (H commutes with exp(-i t H))."""
    import ttfort
    import mpo
    d, tau, steps = 8, 0.05, 4
    crA, ra, n = mpo.tfim_schroedinger(d)
    cr, ry = mpo.product_state_padded(d, 64)
    y0 = O.tt_full(_split(cr, ry, n), n, ry)
    Ad = O.mpo_to_dense(_split(crA, ra, n, n), ra, n, n)
    for _ in range(steps):
        cr, ry, _ = ttfort.ztt_ksl(n, ra, crA, cr, ry, tau, order=30)
    x = O.tt_full(_split(cr, ry, n), n, ry)
    ex = sl.expm(steps * tau * Ad) @ y0
    assert np.linalg.norm(x - ex) <= 1e-7 * np.linalg.norm(ex)
    H = 1j * Ad
    energy = (x.conj() @ H @ x / (x.conj() @ x)).real
    assert abs(energy + (d - 1)) <= 1e-8 * (d - 1)


def test_ksl_complex64_matvec(ttlib):
    """Complex64 core products in the local exponentials (Krylov basis,
    Hessenberg exponential, S-steps in complex128): the full-rank TFIM step
    and the padded-product start against the exact evolution, with the error
    against the FP64 path stated: within 1e-5 relative (the complex64
    rounding of about 30 Krylov products per local exponential), energy to
    1e-6, and every core product counted as complex64."""
    import ttfort
    import mpo
    d, tau, steps = 8, 0.05, 4
    crA, ra, n = mpo.tfim_schroedinger(d)
    cr0, ry0 = mpo.product_state_padded(d, 64)
    y0 = O.tt_full(_split(cr0, ry0, n), n, ry0)
    Ad = O.mpo_to_dense(_split(crA, ra, n, n), ra, n, n)
    out = {}
    for bits in (64, 32):
        cr, ry = cr0, ry0
        ttfort.set_local_precision(bits)
        try:
            for _ in range(steps):
                cr, ry, st = ttfort.ztt_ksl(n, ra, crA, cr, ry, tau, order=30)
        finally:
            ttfort.set_local_precision(64)
        out[bits] = (O.tt_full(_split(cr, ry, n), n, ry), st)
    x64, x32 = out[64][0], out[32][0]
    ex = sl.expm(steps * tau * Ad) @ y0
    err32 = np.linalg.norm(x32 - ex) / np.linalg.norm(ex)
    err_vs64 = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    print("complex64 matvec: error vs exact %.2e, vs complex128 path %.2e" % (err32, err_vs64))
    assert err32 <= 1e-5 and err_vs64 <= 1e-5
    H = 1j * Ad
    energy = (x32.conj() @ H @ x32 / (x32.conj() @ x32)).real
    assert abs(energy + (d - 1)) <= 1e-6 * (d - 1)
    st32 = out[32][1]
    assert st32["matvecs32"] > 0 and out[64][1]["matvecs32"] == 0

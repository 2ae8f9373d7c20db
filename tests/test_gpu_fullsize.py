"""Full-size GPU runs of the C3 and C4 workloads (the benchmark
configurations), checked through properties that do not depend on size:

* C3, tt_amen_solve on the 6D variable-coefficient diffusion -div(a grad u)
  = 1, n = 64, coefficient rank 3, tol 1e-8, rmax 64 (the benchmark solve):
  the global FP64 residual ||A x - b|| / ||b||, computed in TT format on the
  host, is below 1e-6 (measured 1.95e-7: the tolerance is a local one, and
  tt_amen_solve's stopping rule is the local residual, tt_adapt_als.f90),
  the last sweep's local residuals meet the tolerance, no rank exceeds
  rmax, and the mixed-precision local solver (fp32 Krylov, FP64 restarts)
  ends at the same global residual within a factor of 2.
* C4, ztt_ksl on the transverse-field Ising chain g = 1, d = 40, fixed rank
  64, four steps of 0.05 from the all-up product state: exp(-i t H) is
  unitary and commutes with H, so the norm stays 1 and the energy stays
  -(d - 1) = -39: to 1e-8 with complex128 core products (measured 2e-15
  and 9e-15 relative), and with complex64 ones, whose rounding over about
  30 Krylov products per local exponential makes the state drift, the
  squared norm to 5e-5 and the energy to 1e-5 relative (measured 1.4e-5
  and 5.4e-8 after the four steps).
"""
import time

import numpy as np
import pytest

import mpo


def _say(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)

pytestmark = pytest.mark.gpu


def _tt_expect(d, n, ra, crA, cr, ry):
    """<y|A|y> and <y|y> for a TT vector y and TT-matrix A, by contracting
    the cores left to right on the host."""
    pa = py = 0
    L = np.ones((1, 1, 1), dtype=np.complex128)
    N = np.ones((1, 1), dtype=np.complex128)
    for k in range(d):
        a = crA[pa:pa + ra[k] * n[k] * n[k] * ra[k + 1]].reshape((ra[k], n[k], n[k], ra[k + 1]), order="F")
        y = cr[py:py + ry[k] * n[k] * ry[k + 1]].reshape((ry[k], n[k], ry[k + 1]), order="F")
        pa += a.size
        py += y.size
        # pairwise products (a four-operand einsum is far slower here)
        T = np.tensordot(L, y, axes=([2], [0]))             # (x, a, j, Y)
        T = np.tensordot(T, a, axes=([1, 2], [0, 2]))       # (x, Y, i, A)
        L = np.tensordot(y.conj(), T, axes=([0, 1], [0, 2])).transpose(0, 2, 1)   # (X, A, Y)
        M = np.tensordot(N, y, axes=([1], [0]))             # (x, i, Y)
        N = np.tensordot(y.conj(), M, axes=([0, 1], [0, 1]))   # (X, Y)
    return complex(L.reshape(-1)[0]), complex(N.reshape(-1)[0])


def test_c3_amen_full_size(ttlib):
    import ttfort
    import ttops
    d, n, tol, rmax = 6, 64, 1e-8, 64
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=0, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    out = {}
    for bits in (64, 32):
        _say("C3 solve, local precision", bits)
        ttfort.set_local_precision(bits)
        try:
            res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, tol, kickrank=4, nswp=20, prec="c",
                                                   nrestart=40, niters=2, rmax=rmax)
        finally:
            ttfort.set_local_precision(64)
        _say("solved in %.2f s; TT residual on the host" % st["seconds"])
        g = ttops.residual(crA, ra, nn, res, rx_out, crY, ry)
        _say("bits", bits, "global residual", g, "sweeps", st["sweeps"], "last local", st["last_res_max"],
             "max rank", max(rx_out))
        assert max(rx_out) <= rmax
        assert st["last_res_max"] < tol
        assert g < 1e-6, g
        out[bits] = g
    assert out[32] <= 2.0 * out[64], out


@pytest.mark.parametrize("bits,ntol,etol", [(64, 1e-8, 1e-8), (32, 5e-5, 1e-5)])
def test_c4_ksl_full_size(ttlib, bits, ntol, etol):
    import ttfort
    d, r, tau, steps = 40, 64, 0.05, 4
    crA, ra, n = mpo.tfim_schroedinger(d)
    cr, ry = mpo.product_state_padded(d, r)
    sz0 = int(ra[0] * n[0] * n[0] * ra[1])
    crH = np.concatenate([1j * crA[:sz0], crA[sz0:]])      # H = i A (A carries -i on core 1)
    e0, n0 = _tt_expect(d, n, ra, crH, cr, ry)
    assert abs(n0 - 1.0) < 1e-12 and abs(e0.real + (d - 1)) < 1e-10
    ttfort.set_local_precision(bits)
    try:
        for k in range(steps):
            cr, ry, st = ttfort.ztt_ksl(n, ra, crA, cr, ry, tau, order=30)
            _say("C4 step", k, "bits", bits, "%.3f s" % st["seconds"])
    finally:
        ttfort.set_local_precision(64)
    e1, n1 = _tt_expect(d, n, ra, crH, cr, ry)
    _say("bits", bits, "norm^2", n1.real, "energy", (e1 / n1).real, "max rank", max(ry))
    assert max(ry) <= r
    assert abs(n1.real - 1.0) <= ntol
    assert abs((e1 / n1).real + (d - 1)) <= etol * (d - 1)

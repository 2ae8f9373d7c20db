"""The typed AMEn solver (ttf_dtt_amen_solve, ttamen.f90:29-252) on the GPU.

Pins: the Dirichlet Laplacian's exact solution through the tensor DST
(oracle.operators.laplace_dst_solve, itself pinned against a dense solve),
a dense solve of the variable-coefficient 7-point TT-matrix, and the FP64
TT residual of every result.  Covered: truncation by local residual
(trunc_norm 1, the default) and by the Frobenius chop (0), z-enrichment on
(kickrank 4) and off (0, fixed ranks), local preconditioners 'n' and 'c',
and the dense local solve below max_full_size.  Tolerances: solution within
1e-6 relative of the exact one at tol 1e-8 (the local solves stop at
tol / sqrt(d-1) / 2), residual at most 1e-6.
"""
import numpy as np
import pytest

import mpo
from oracle import operators as O

pytestmark = pytest.mark.gpu


def _cores(cr, r, n, m=None):
    out, off = [], 0
    for k in range(len(n)):
        sz = int(r[k] * n[k] * (m[k] if m is not None else 1) * r[k + 1])
        out.append(cr[off:off + sz])
        off += sz
    return out


@pytest.mark.parametrize("d,n,prec,trunc,kick", [(4, 8, "n", 1, 4), (4, 8, "c", 1, 4), (3, 12, "n", 0, 4),
                                                 (5, 6, "c", 1, 2)])
def test_typed_amen_laplace(ttlib, d, n, prec, trunc, kick):
    import ttfort
    import ttops
    crA, ra, nn = mpo.laplace_dd(d, n)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    res, rx_out, st = ttfort.dtt_amen_solve(nn, ra, crA, ry, crY, 1e-8, rx, crX, kickrank=kick, local_prec=prec,
                                            trunc_norm=trunc, seed=3)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    x_ref = O.laplace_dst_solve(d, n, np.ones(n ** d))
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print("laplace d=%d n=%d prec=%s trunc=%d: error %.2e, sweeps %d, max rank %d"
          % (d, n, prec, trunc, err, st["sweeps"], st["max_rank"]))
    assert err <= 1e-6
    assert ttops.residual(crA, ra, nn, res, rx_out, crY, ry) <= 1e-6
    if trunc == 1:
        assert st["last_res_max"] < 1e-8


@pytest.mark.parametrize("prec", ["n", "c"])
def test_typed_amen_diffusion(ttlib, prec):
    import ttfort
    import ttops
    d, n = 4, 8
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=5, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    res, rx_out, st = ttfort.dtt_amen_solve(nn, ra, crA, ry, crY, 1e-8, rx, crX, local_prec=prec, seed=4)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    A = O.mpo_to_dense(_cores(crA, ra, nn, nn), ra, nn, nn)
    x_ref = np.linalg.solve(A, np.ones(n ** d))
    assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    assert ttops.residual(crA, ra, nn, res, rx_out, crY, ry) <= 1e-6
    assert st["flops"] > 0


def test_typed_amen_fixed_ranks_without_kick(ttlib):
    """kickrank 0: no z, the ranks of the start are kept (QR moves only),
    and from ranks that hold the solution the solve is exact."""
    import ttfort
    d, n = 4, 8
    crA, ra, nn = mpo.laplace_dd(d, n)
    crY, ry = mpo.ones_tt(d, n)
    rx = [1, 8, 8, 8, 1]
    crX = np.concatenate([np.random.default_rng(6).standard_normal(rx[k] * n * rx[k + 1]) for k in range(d)])
    res, rx_out, st = ttfort.dtt_amen_solve(nn, ra, crA, ry, crY, 1e-9, rx, crX, kickrank=0, nswp=10)
    assert list(rx_out) == rx
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    x_ref = O.laplace_dst_solve(d, n, np.ones(n ** d))
    assert np.linalg.norm(x - x_ref) <= 1e-7 * np.linalg.norm(x_ref)


def test_typed_amen_dense_local_solves(ttlib):
    """max_full_size above every local size: all local corrections by the
    dense inverse of the explicit local matrix."""
    import ttfort
    d, n = 3, 4
    crA, ra, nn = mpo.laplace_dd(d, n)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=2)
    res, rx_out, st = ttfort.dtt_amen_solve(nn, ra, crA, ry, crY, 1e-10, rx, crX, max_full_size=129)
    x = O.tt_full(_cores(res, rx_out, nn), nn, rx_out)
    x_ref = O.laplace_dst_solve(d, n, np.ones(n ** d))
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)

"""CPU tests of the host-side operator builders and TT arithmetic: the
variable-coefficient 7-point diffusion TT-matrix against a direct stencil
assembly, the TFIM Schroedinger operator and padded product state, TT cross
interpolation, and the TT residual/norm/rounding helpers against dense
vectors built by the oracle."""
import itertools

import numpy as np
import pytest

import mpo
import ttops
from oracle import operators as O


def _dense_op(crA, ra, n):
    return O.mpo_to_dense(ttops.split(crA, ra, n, n), ra, n, n)


def _stencil(d, n, amp, c):
    """Direct assembly of -div(a grad u), 7-point stencil, half-point
    coefficients averaged from the two neighbouring grid points."""
    h = 1.0 / (n + 1)
    xg = np.arange(n + 2) * h

    def a(p):
        return 1.0 + amp * np.tanh(xg[list(p)] @ c)

    N = n ** d
    B = np.zeros((N, N))
    for p in itertools.product(range(1, n + 1), repeat=d):
        i = sum((p[k] - 1) * n ** k for k in range(d))
        for k in range(d):
            for s in (1, -1):
                q = list(p)
                q[k] += s
                ah = 0.5 * (a(p) + a(q))
                B[i, i] += ah / h ** 2
                if 1 <= q[k] <= n:
                    j = sum((q[l] - 1) * n ** l for l in range(d))
                    B[i, j] -= ah / h ** 2
    return B


@pytest.mark.parametrize("d,n", [(2, 7), (3, 5)])
def test_diffusion_exact_coefficient(d, n):
    # coefficient rank high enough for the cross to be exact on this grid
    crA, ra, nn, info = mpo.diffusion_7pt(d, n, seed=3, coef_rank=n + 2)
    assert info["coef_maxerr"] < 1e-12
    A = _dense_op(crA, ra, nn)
    B = _stencil(d, n, 0.5, info["c"])
    assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max()


def test_diffusion_low_rank_coefficient():
    # the benchmark setting: rank-3 coefficient, operator rank 6, exactly the
    # stencil of the TT coefficient (max deviation from a stated in info)
    d, n = 3, 6
    crA, ra, nn, info = mpo.diffusion_7pt(d, n, seed=1, coef_rank=3)
    assert max(ra) == 6
    A = _dense_op(crA, ra, nn)
    B = _stencil(d, n, 0.5, info["c"])
    h2 = (n + 1) ** 2
    # each row of the stencil sums 4d half-point coefficients
    assert np.abs(A - B).max() <= 4 * d * h2 * info["coef_maxerr"] + 1e-12 * np.abs(B).max()
    assert np.allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > 0


def test_diffusion_constant_is_laplacian():
    d, n = 3, 5
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=0, coef_rank=2, amp=0.0)
    crL, rl, nl = mpo.laplace_dd(d, n)
    assert np.allclose(_dense_op(crA, ra, nn), _dense_op(crL, rl, nl), rtol=0, atol=1e-10)


def test_tt_cross_interpolates():
    rng = np.random.default_rng(0)
    n = [7, 6, 8, 5]
    # a rank-2 function: exact at rank >= 2
    u = [rng.standard_normal(k) for k in n]
    v = [rng.standard_normal(k) for k in n]

    def f(idx):
        return np.prod([u[k][idx[:, k]] for k in range(4)], axis=0) + \
            np.prod([v[k][idx[:, k]] for k in range(4)], axis=0)

    G = mpo.tt_cross(f, n, 3)
    full = O.tt_full([g.reshape(-1, order="F") for g in G], n, [1] + [g.shape[2] for g in G])
    grid = np.array(list(itertools.product(*[range(k) for k in reversed(n)])))[:, ::-1]
    assert np.abs(full - f(grid)).max() < 1e-12 * np.abs(f(grid)).max()


def test_ttops_residual_norm_round():
    rng = np.random.default_rng(2)
    d, n = 4, [3, 4, 2, 3]
    crA, ra, nn = mpo.tfim(d)
    n = list(nn)
    rx = [1, 2, 3, 2, 1]
    ry = [1, 3, 2, 2, 1]
    X = O.random_tt(rng, n, rx)
    Y = O.random_tt(rng, n, ry)
    crX = np.concatenate([np.asarray(c).reshape(-1, order="F") for c in X])
    crY = np.concatenate([np.asarray(c).reshape(-1, order="F") for c in Y])
    A = _dense_op(crA, ra, nn)
    x = O.tt_full(X, n, rx)
    y = O.tt_full(Y, n, ry)
    res = ttops.residual(crA, ra, nn, crX, rx, crY, ry)
    assert res == pytest.approx(np.linalg.norm(A @ x - y) / np.linalg.norm(y), rel=1e-12)
    # a residual far below the operands' norms is not lost to cancellation
    Ax = ttops.matvec(crA, ra, nn, nn, crX, rx)
    crAx, rax = ttops.join(Ax)
    scaled = crAx.copy()
    scaled[:Ax[0].size] *= 1 + 1e-13          # first core only: the vector scales once
    tiny = ttops.norm(ttops.sub(Ax, ttops.split(scaled, rax, n)))
    assert tiny == pytest.approx(1e-13 * np.linalg.norm(A @ x), rel=1e-3)
    # rounding the (redundant) A x representation keeps the vector
    Z = ttops.round_tt(Ax, 1e-13)
    z = O.tt_full([c.reshape(-1, order="F") for c in Z], n, [1] + [c.shape[2] for c in Z])
    assert np.linalg.norm(z - A @ x) <= 1e-12 * np.linalg.norm(A @ x)
    assert max(c.shape[2] for c in Z) <= max(rax)


def test_tfim_schroedinger_and_padded_state():
    d = 5
    crA, ra, n = mpo.tfim_schroedinger(d)
    crH, _, _ = mpo.tfim(d)
    assert np.allclose(_dense_op(crA, ra, n), -1j * _dense_op(crH, ra, n))
    cr, ry = mpo.product_state_padded(d, 8)
    assert list(ry) == [1, 2, 4, 4, 2, 1]
    v = O.tt_full(ttops.split(cr, ry, n), n, ry)
    e = np.zeros(2 ** d)
    e[0] = 1.0
    assert np.allclose(v, e)

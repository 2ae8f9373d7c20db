"""CPU checks of the learned-preconditioner oracle: its gradients against
central finite differences in float64, its block diagonals against dense
blocks built from the block-Jacobi 'c' definition, its targets against a
dense block solve, and bf16 rounding on known values."""
import numpy as np
import pytest

from oracle import kernels as K
from oracle import mlp as ML


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.1415927, -2.5e-3, 65504.0], dtype=np.float32)
    r = ML.bf16_round(x)
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7:
    # ties go to the even mantissa (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] - 3.140625) == 0.0
    for v, rv in zip(x, r):
        assert abs(rv - v) <= abs(v) * 2.0 ** -8


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    theta = rng.uniform(-0.1, 0.1, ML.n_params())
    X = rng.standard_normal((ML.K_IN, 6))
    T = rng.standard_normal((ML.N_OUT, 6))

    def L(th):
        y, _ = ML.forward(th, X, round_weights=False, dtype=np.float64)
        return ML.loss(y, T)

    y, cache = ML.forward(theta, X, round_weights=False, dtype=np.float64)
    g = ML.backward(cache, y, T)
    idx = np.concatenate([rng.integers(0, ML.n_params(), 40),
                          np.arange(ML.n_params() - 64, ML.n_params())])   # includes b3
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        fd = (L(theta + e) - L(theta - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def _local(rng, rx1, n, rx2, ra1, ra2):
    phi1 = rng.standard_normal(rx1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * n * ra2)
    phi2 = rng.standard_normal(rx2 * ra2 * rx2)
    return phi1, A, phi2


def test_block_diagonals_match_dense_blocks():
    rng = np.random.default_rng(1)
    rx1, n, rx2, ra1, ra2 = 3, 7, 4, 2, 3
    phi1, A, phi2 = _local(rng, rx1, n, rx2, ra1, ra2)
    diag, up, lo = ML.block_diagonals(rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    # the dense blocks: inverses of djac_gen 'c' inverted back
    inv = K.F(K.djac_gen("c", rx1, n, rx2, ra1, ra2, phi1, A, phi2), n, n, rx2, rx1)
    for b1 in range(rx1):
        for b2 in range(rx2):
            B = np.linalg.inv(inv[:, :, b2, b1])
            assert np.allclose(diag[:, b1, b2], np.diag(B), rtol=1e-8, atol=1e-8)
            assert np.allclose(up[:-1, b1, b2], np.diag(B, 1), rtol=1e-8, atol=1e-8)
            assert np.allclose(lo[:-1, b1, b2], np.diag(B, -1), rtol=1e-8, atol=1e-8)
            assert up[-1, b1, b2] == 0 and lo[-1, b1, b2] == 0


def test_targets_are_scaled_block_solves():
    rng = np.random.default_rng(2)
    rx1, n, rx2, ra1, ra2 = 2, 5, 3, 2, 2
    phi1, A, phi2 = _local(rng, rx1, n, rx2, ra1, ra2)
    v = rng.standard_normal(rx1 * n * rx2)
    T = ML.block_solve_targets(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    X, scale = ML.features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    jacs = K.djac_gen("c", rx1, n, rx2, ra1, ra2, phi1, A, phi2)
    z = K.F(K.djac_apply("c", rx1, n, rx2, jacs, v), rx1, n, rx2)
    for b2 in range(rx2):
        for b1 in range(rx1):
            b = b1 + rx1 * b2
            # y* scaled back by |r_b| / s_b is the block-Jacobi correction
            assert np.allclose(T[:, b] * scale[b], z[b1, :, b2], rtol=1e-10, atol=1e-12)
            # the residual part of the features is the unit slice
            assert np.linalg.norm(X[3 * n:, b]) == pytest.approx(1.0)

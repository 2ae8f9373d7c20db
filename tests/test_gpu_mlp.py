"""GPU tests of the learned block preconditioner: features, forward pass and
gradients against the oracle (oracle/mlp.py), parameter round trip, training
on pairs collected from plain AMEn runs, and what the trained network gains
in the AMEn solve over GMRES without a preconditioner."""
import numpy as np
import pytest

import mpo
from oracle import mlp as ML

pytestmark = pytest.mark.gpu


def _local(rng, rx1, n, rx2, ra1, ra2):
    phi1 = rng.standard_normal(rx1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * n * ra2)
    phi2 = rng.standard_normal(rx2 * ra2 * rx2)
    return phi1, A, phi2


def test_features_match_oracle(ttlib):
    import ttfort
    rng = np.random.default_rng(0)
    rx1, n, rx2, ra1, ra2 = 3, 64, 5, 2, 3
    phi1, A, phi2 = _local(rng, rx1, n, rx2, ra1, ra2)
    v = rng.standard_normal(rx1 * n * rx2)
    X, sc = ttfort.mlp_features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    Xr, scr = ML.features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    assert np.allclose(X, Xr, rtol=1e-6, atol=1e-6 * np.abs(Xr).max())
    assert np.allclose(sc, scr, rtol=1e-6)


def test_params_round_trip_and_forward(ttlib):
    import ttfort
    ttfort.mlp_init(3)
    theta = ttfort.mlp_get_params()
    assert theta.size == ML.n_params()
    lim = 1.0 / np.sqrt(256)
    assert np.abs(theta[:256 * 256]).max() <= lim and np.abs(theta[256 * 256:256 * 257]).max() == 0
    assert np.abs(theta[-(256 * 64 + 64):]).max() == 0   # W3, b3: the point-Jacobi start
    rng = np.random.default_rng(1)
    theta = rng.uniform(-0.1, 0.1, theta.size).astype(np.float32)
    ttfort.mlp_set_params(theta)
    assert np.array_equal(ttfort.mlp_get_params(), theta)
    X = rng.standard_normal((256, 300)).astype(np.float32)
    # inference runs on the bf16 matrix cores: bf16 weights and layer inputs,
    # fp32 accumulation.  Against the oracle with the same roundings the
    # differences are fp32 summation order plus the rare activation whose
    # bf16 rounding flips on it (one bf16 step, 2^-8, in one input of one
    # output): 2e-3 of the largest output bounds both
    Y = ttfort.mlp_forward(X)
    Yr, _ = ML.forward(theta, X, round_weights=True, dtype=np.float64, round_acts=True)
    assert np.abs(Y - Yr).max() <= 2e-3 * np.abs(Yr).max()
    # and the bf16 activations move the output by a few bf16 steps against
    # the fp32-activation network
    Yf, _ = ML.forward(theta, X, round_weights=True, dtype=np.float64)
    assert np.abs(Y - Yf).max() <= 3e-2 * np.abs(Yf).max()
    # the edge tiles (M and N not multiples of 64) agree with the full batch
    Y7 = ttfort.mlp_forward(np.asfortranarray(X[:, :7]))
    assert np.array_equal(Y7, Y[:, :7])


def test_loss_grad_match_oracle(ttlib):
    import ttfort
    rng = np.random.default_rng(2)
    theta = rng.uniform(-0.08, 0.08, ML.n_params()).astype(np.float32)
    ttfort.mlp_set_params(theta)
    M = 200
    X = rng.standard_normal((256, M)).astype(np.float32)
    T = rng.standard_normal((64, M)).astype(np.float32)
    loss, g = ttfort.mlp_loss_grad(X, T)
    y, cache = ML.forward(theta, X, round_weights=True, dtype=np.float64)
    assert loss == pytest.approx(ML.loss(y, T), rel=1e-5)
    gr = ML.backward(cache, y, T)
    for name, shape in ML.param_shapes():
        off = sum(int(np.prod(s)) for nm, s in ML.param_shapes()[:[p[0] for p in ML.param_shapes()].index(name)])
        sz = int(np.prod(shape))
        a, b = g[off:off + sz], gr[off:off + sz]
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max() + 1e-9, name


def test_precond_is_scaled_network_output(ttlib):
    import ttfort
    rng = np.random.default_rng(3)
    rx1, n, rx2, ra1, ra2 = 4, 64, 3, 2, 2
    phi1, A, phi2 = _local(rng, rx1, n, rx2, ra1, ra2)
    v = rng.standard_normal(rx1 * n * rx2)
    ttfort.mlp_init(5)
    theta = ttfort.mlp_get_params()
    z = ttfort.mlp_precond(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    X, sc = ML.features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v)
    y, _ = ML.forward(theta, X, round_weights=True, dtype=np.float64)
    Z = np.zeros((rx1, n, rx2))
    for b2 in range(rx2):
        for b1 in range(rx1):
            Z[b1, :, b2] = y[:, b1 + rx1 * b2] * sc[b1 + rx1 * b2]
    zr = Z.reshape(-1, order="F")
    assert np.abs(z - zr).max() <= 1e-5 * np.abs(zr).max()


_TRAINED = {}


def _collect_and_train(draws=6, per_draw=400, epochs=30):
    import learned
    import ttfort
    if "theta" not in _TRAINED:
        X, T, info = learned.collect_pairs(draws, per_draw, d=4, n=64, seed0=500, rmax=12, tol=1e-6, nswp=3)
        assert info["pairs"] >= draws * per_draw // 2
        ttfort.mlp_init(1)
        theta0 = ttfort.mlp_get_params()
        theta, hist, _ = learned.train(X, T, epochs=epochs, batch=256, lr=1e-3, seed=1)
        _TRAINED.update(X=X, T=T, theta0=theta0, theta=theta, hist=hist)
    ttfort.mlp_set_params(_TRAINED["theta"])
    return _TRAINED


def test_collected_pairs_and_training(ttlib):
    import ttfort
    t = _collect_and_train()
    X, T = t["X"][:, :2000], t["T"][:, :2000]
    # the residual part of every feature column is a unit vector
    assert np.allclose(np.linalg.norm(X[192:], axis=0), 1.0, atol=1e-5)
    # the trained network matches the oracle's evaluation of its parameters,
    # and halves (at least) the loss of the untrained network, which is the
    # point-Jacobi solve of each block
    yo, _ = ML.forward(t["theta"], X, round_weights=True, dtype=np.float64)
    yb, _ = ML.forward(t["theta"], X, round_weights=True, dtype=np.float64, round_acts=True)
    yd = ttfort.mlp_forward(X)   # bf16 inference
    assert np.abs(yd - yb).max() <= 2e-3 * np.abs(yb).max()
    y0, _ = ML.forward(t["theta0"], X, round_weights=True, dtype=np.float64)
    assert ML.loss(yo, T) < 0.5 * ML.loss(y0, T), (ML.loss(yo, T), ML.loss(y0, T), t["hist"])
    assert t["hist"][-1] < t["hist"][0]


def test_amen_learned_gain_over_plain_gmres(ttlib):
    """On a held-out coefficient draw, AMEn with the trained network in
    flexible GMRES reaches the same local tolerance and residual as AMEn with
    unpreconditioned GMRES, using fewer local matvecs."""
    import ttfort
    import ttops
    _collect_and_train()
    d, n = 4, 64
    crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=999, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=7)
    eps = 1e-6
    out = {}
    for prec in ("n", "m"):
        res, rx_out, st = ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, eps, kickrank=2, nswp=20, prec=prec,
                                               nrestart=40, niters=2, rmax=12)
        out[prec] = (st, ttops.residual(crA, ra, nn, res, rx_out, crY, ry))
    st_n, res_n = out["n"]
    st_m, res_m = out["m"]
    assert st_m["learned_solves"] > 0
    assert st_m["last_res_max"] < eps
    assert res_m <= max(3.0 * res_n, 1e-5)
    assert st_m["total_matvecs"] < st_n["total_matvecs"], (st_m, st_n)

"""Oracle restatement of tt_mvk4 (tt_adapt_als.f90:67-500), the AMEn-style
approximate product y ~ A x of a TT-matrix and a TT vector.

Pins (each against the exact product A x formed on full vectors by
mpo_apply_full, an independent construction):
  * from a rank-2 start the right-to-left enrichment (duchol of the partial
    product) must grow the ranks until y equals A x to rounding, on the
    Heisenberg MPO (d = 8, rx = 3), where the exact ranks are reachable;
  * with truncation active (d = 12, ranks of A x up to 15 < 2^6) the
    relative error stays below eps for eps = 1e-2, 1e-4, 1e-8 and the
    largest rank does not fall as eps does;
  * with kickrank = 0 the reference keeps the ranks of y0 (QR steps only):
    from y0 at the exact ranks the result is A x to rounding, and from a
    rank-2 start the ranks stay 2;
  * the identity operator returns x, at ranks no more than one above x's
    (my_chop3 keeps one singular value past the tolerance).
"""
import numpy as np
import pytest

from oracle import drivers as D
from oracle import operators as O


def _full(flat, r, n):
    cores, p = [], 0
    for k in range(len(n)):
        sz = r[k] * n[k] * r[k + 1]
        cores.append(flat[p:p + sz].reshape((r[k], n[k], r[k + 1]), order="F"))
        p += sz
    return O.tt_full(cores, n, r)


def _graded_x(rng, d, rx):
    # decaying core entries, so that A x has decaying singular values
    return np.concatenate([rng.standard_normal(rx[k] * 2 * rx[k + 1]) * 0.5 ** np.arange(rx[k] * 2 * rx[k + 1])
                           for k in range(d)])


def _random_flat(rng, d, r):
    return np.concatenate([rng.standard_normal(r[k] * 2 * r[k + 1]) for k in range(d)])


def _setup(d, seed, graded):
    cores, ra, nn = O.heisenberg_mpo(d)
    rng = np.random.default_rng(seed)
    rx = [1] + [3] * (d - 1) + [1]
    crX = _graded_x(rng, d, rx) if graded else _random_flat(rng, d, rx)
    x = _full(crX, rx, [2] * d)
    Ax = O.mpo_apply_full(cores, ra, nn, nn, x)
    return O.flatten_cores(cores), ra, rx, crX, Ax, rng


@pytest.mark.parametrize("seed", [0, 1])
def test_enrichment_reaches_exact_product(seed):
    d = 8
    crA, ra, rx, crX, Ax, rng = _setup(d, seed, graded=False)
    ry = [1] + [2] * (d - 1) + [1]
    cy, ryo, sweeps = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, _random_flat(rng, d, ry), ry, 1e-10)
    y = _full(cy, list(ryo), [2] * d)
    assert np.linalg.norm(y - Ax) <= 1e-12 * np.linalg.norm(Ax)
    assert max(ryo) > 2 and sweeps < 20
    # no rank beyond the bound of the exact product (ra rx) or of the mode sizes
    for k in range(d + 1):
        assert ryo[k] <= min(ra[k] * rx[k], 2 ** k, 2 ** (d - k))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_truncation_error_within_eps(seed):
    d = 12
    crA, ra, rx, crX, Ax, rng = _setup(d, seed, graded=True)
    ry = [1] + [2] * (d - 1) + [1]
    y0 = _random_flat(rng, d, ry)
    last_rank = 0
    for eps in (1e-2, 1e-4, 1e-8):
        cy, ryo, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, y0, ry, eps)
        y = _full(cy, list(ryo), [2] * d)
        assert np.linalg.norm(y - Ax) <= eps * np.linalg.norm(Ax), eps
        assert max(ryo) >= last_rank
        last_rank = max(ryo)


def test_kickrank_zero_keeps_ranks():
    d = 12
    crA, ra, rx, crX, Ax, rng = _setup(d, 3, graded=True)
    exact = [1, 2, 4, 8, 15, 15, 15, 15, 15, 8, 4, 2, 1]
    cy, ryo, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, _random_flat(rng, d, exact), exact, 1e-10,
                           kickrank=0)
    assert list(ryo) == exact
    y = _full(cy, exact, [2] * d)
    assert np.linalg.norm(y - Ax) <= 1e-12 * np.linalg.norm(Ax)
    low = [1] + [2] * (d - 1) + [1]
    _, ryo, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, _random_flat(rng, d, low), low, 1e-10,
                          kickrank=0, nswp=3)
    assert list(ryo) == low


def test_identity_operator_returns_x():
    d = 10
    rng = np.random.default_rng(5)
    ra = [1] * (d + 1)
    crA = np.concatenate([np.eye(2).reshape(-1, order="F")] * d)
    rx = [1] + [3] * (d - 1) + [1]
    crX = _random_flat(rng, d, rx)
    x = _full(crX, rx, [2] * d)
    ry = [1] + [2] * (d - 1) + [1]
    cy, ryo, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, _random_flat(rng, d, ry), ry, 1e-10)
    y = _full(cy, list(ryo), [2] * d)
    assert np.linalg.norm(y - x) <= 1e-10 * np.linalg.norm(x)
    # my_chop3 keeps one singular value past the tolerance
    # (tt_adapt_als.f90:52-60), so the ranks end at rx + 1, not rx
    assert max(ryo) <= 3 + 1

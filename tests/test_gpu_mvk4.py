"""tt_mvk4 on the GPU (ttf_tt_mvk4 through the C ABI) against the exact
product A x on full vectors (mpo_apply_full, an independent construction)
and against the oracle's restatement of tt_adapt_als.f90:67-500 on the same
inputs.

  * from a rank-2 start the enrichment grows y to A x to rounding (1e-10),
    on the Heisenberg MPO with a rank-3 x (d = 8) -- the same answer as the
    oracle to 1e-10;
  * with truncation active (d = 12, graded x) the relative error is below
    eps for eps = 1e-2, 1e-4, 1e-8, as the oracle's is;
  * with kickrank = 0 the ranks of y0 are kept (2 stays 2), and from y0 at
    the exact ranks the result is A x;
  * a rank beyond the capacity set by rmax is an error, not an overrun.
"""
import numpy as np
import pytest

from oracle import drivers as D
from oracle import operators as O
from test_oracle_mvk4 import _full, _random_flat, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 1])
def test_mvk4_reaches_exact_product(ttlib, seed):
    import ttfort
    d = 8
    crA, ra, rx, crX, Ax, rng = _setup(d, seed, graded=False)
    ry = [1] + [2] * (d - 1) + [1]
    y0 = _random_flat(rng, d, ry)
    cy, ryo, st = ttfort.tt_mvk4([2] * d, [2] * d, rx, ra, crA, crX, ry, y0, 1e-10, rmax=64)
    y = _full(cy, list(ryo), [2] * d)
    assert np.linalg.norm(y - Ax) <= 1e-10 * np.linalg.norm(Ax), np.linalg.norm(y - Ax) / np.linalg.norm(Ax)
    co, ro, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, y0, ry, 1e-10)
    yo = _full(co, list(ro), [2] * d)
    assert np.linalg.norm(y - yo) <= 1e-10 * np.linalg.norm(yo)
    assert st["converged"] == 1 and st["flops"] > 0


@pytest.mark.parametrize("eps", [1e-2, 1e-4, 1e-8])
def test_mvk4_truncated_accuracy(ttlib, eps):
    import ttfort
    d = 12
    crA, ra, rx, crX, Ax, rng = _setup(d, 3, graded=True)
    ry = [1] + [2] * (d - 1) + [1]
    y0 = _random_flat(rng, d, ry)
    cy, ryo, st = ttfort.tt_mvk4([2] * d, [2] * d, rx, ra, crA, crX, ry, y0, eps, rmax=64)
    y = _full(cy, list(ryo), [2] * d)
    err = np.linalg.norm(y - Ax) / np.linalg.norm(Ax)
    co, ro, _ = D.tt_mvk4(d, [2] * d, [2] * d, rx, ra, crA, crX, y0, ry, eps)
    erro = np.linalg.norm(_full(co, list(ro), [2] * d) - Ax) / np.linalg.norm(Ax)
    print("eps %.0e: device error %.2e (max rank %d), oracle error %.2e" % (eps, err, max(ryo), erro))
    assert err <= eps and erro <= eps


def test_mvk4_without_kick_keeps_ranks(ttlib):
    import ttfort
    d = 8
    crA, ra, rx, crX, Ax, rng = _setup(d, 4, graded=False)
    ry = [1] + [2] * (d - 1) + [1]
    cy, ryo, st = ttfort.tt_mvk4([2] * d, [2] * d, rx, ra, crA, crX, ry, _random_flat(rng, d, ry), 1e-10,
                                 rmax=64, kickrank=0, nswp=4)
    assert list(ryo) == ry
    rex = [1] + [min(2 ** k, 2 ** (d - k), 15) for k in range(1, d)] + [1]
    cy, ryo, st = ttfort.tt_mvk4([2] * d, [2] * d, rx, ra, crA, crX, rex, _random_flat(rng, d, rex), 1e-10,
                                 rmax=64, kickrank=0)
    y = _full(cy, list(ryo), [2] * d)
    assert np.linalg.norm(y - Ax) <= 1e-10 * np.linalg.norm(Ax)


def test_mvk4_rank_capacity_is_an_error(ttlib):
    import ttfort
    d = 8
    crA, ra, rx, crX, Ax, rng = _setup(d, 5, graded=False)
    ry = [1] + [2] * (d - 1) + [1]
    with pytest.raises(ttfort.TTFError, match="capacity"):
        ttfort.tt_mvk4([2] * d, [2] * d, rx, ra, crA, crX, ry, _random_flat(rng, d, ry), 1e-12, rmax=3)

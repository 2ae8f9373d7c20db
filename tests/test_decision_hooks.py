"""The integer decisions of the truncations, through the C ABI, bit for bit
against the oracle on identical inputs (CPU: the hooks are host code).

ttf_my_chop3 is the rule ttf_tt_eigb applies to the singular values it reads
back (tt_adapt_als.f90:40-61); ttf_residual_rank_rule is the scan
ttf_tt_amen_solve runs over candidate residuals (tt_adapt_als.f90:819-833,
:937-950).  Besides random inputs, the singular values and residuals are
placed exactly on the thresholds, where a '>=' for the reference's '>' or a
missing '+1' changes the answer.  Those cases use values whose sums, norms
and products are exact in binary floating point, so every correct
implementation must agree on them to the bit.
"""
import numpy as np
import pytest

from oracle import kernels as K


@pytest.fixture(scope="module")
def hooks():
    import ttfort
    return ttfort


def test_my_chop3_random(hooks):
    rng = np.random.default_rng(21)
    for _ in range(3000):
        n = int(rng.integers(1, 80))
        s = np.sort(10.0 ** rng.uniform(-16, 0, n))[::-1]
        eps = 10.0 ** rng.uniform(-14, 0)
        assert hooks.my_chop3(s, eps) == K.my_chop3(s, eps), (s, eps)


def test_my_chop3_on_the_threshold(hooks):
    # ||s|| = 5 exactly, eps = 0.2: the bound (||s|| eps)^2 is 1.0 exactly
    cases = [
        (np.array([4.0, 2.0, 2.0, 1.0, 0.0]), 0.2, 4),   # tail 1.0 == bound: not dropped by '>'
        (np.array([4.0, 2.0, 2.0, 1.0]), 0.2, 4),
        (np.array([3.0, 4.0]), 0.2, 2),                  # ||s|| = 5, tail 16 > 1
        (np.array([3.0, 2.0, 2.0, 2.0, 2.0, 0.0]), 0.4, 5),   # ||s|| = 5, bound 4.0, tail 4.0 kept
        (np.zeros(6), 1e-3, 1),
        (np.array([1.0]), 0.5, 1),
        (np.array([1.0, 1.0, 1.0, 1.0]), 0.5, 4),        # bound 1, tails 1 (==), 2 (>)
    ]
    for s, eps, want in cases:
        assert K.my_chop3(s, eps) == want, (s, eps)
        assert hooks.my_chop3(s, eps) == want, (s, eps)
    assert hooks.my_chop3(np.zeros(0), 0.1) == K.my_chop3(np.zeros(0), 0.1) == 0


def test_my_chop3_planted_mistakes_differ():
    """The exact cases above separate the reference rule from a '>=' and
    from a rule without the '+1'."""
    def ge_rule(s, eps):
        nrm = (np.sqrt(np.sum(s * s)) * eps) ** 2
        cur, i = 0.0, s.size
        while i > 0:
            cur += s[i - 1] ** 2
            if cur >= nrm:
                break
            i -= 1
        return min(i + 1, s.size)
    s = np.array([4.0, 2.0, 2.0, 1.0, 0.0])
    assert ge_rule(s, 0.2) != K.my_chop3(s, 0.2)
    s = np.array([3.0, 2.0, 2.0, 2.0, 2.0, 0.0])
    assert ge_rule(s, 0.4) != K.my_chop3(s, 0.4)
    # and a rule without the '+1' (the minimal rank) differs on graded values
    s = np.array([1.0, 0.1, 0.01, 0.001])
    assert K.my_chop3(s, 0.05) == K.chop(s, 0.05)[0] + 1


def test_residual_rank_rule_random(hooks):
    rng = np.random.default_rng(22)
    for _ in range(3000):
        k = int(rng.integers(1, 70))
        errs = np.concatenate([[-1.0], np.sort(10.0 ** rng.uniform(-12, 0, k - 1))[::-1]])
        if rng.random() < 0.3:
            rng.shuffle(errs[1:])
        res_new = 10.0 ** rng.uniform(-12, -1)
        eps2 = 10.0 ** rng.uniform(-12, -1)
        want = K.residual_rank(lambda c: errs[c], k, res_new, eps2)
        assert hooks.residual_rank_rule(errs, res_new, eps2) == want


def test_residual_rank_rule_on_the_threshold(hooks):
    errs = np.array([-1.0, 4.0, 2.0, 1.0, 0.5])
    for res_new, eps2, want in ((1.0, 0.1, 2), (0.5, 2.0, 2), (0.49, 0.1, 4), (0.25, 0.5, 4),
                                (0.25, 0.25, 4), (0.2, 0.1, 5), (8.0, 0.1, 1)):
        assert K.residual_rank(lambda c: errs[c], 5, res_new, eps2) == want
        assert hooks.residual_rank_rule(errs, res_new, eps2) == want
    # candidates that straddle a batch boundary of the device scan (8 per
    # block matvec): the first stop is at candidate 9 of 20
    errs = np.concatenate([[-1.0], np.full(9, 1.0), np.full(10, 1e-3)])
    assert hooks.residual_rank_rule(errs, 1e-2, 1e-6) == 10 == K.residual_rank(lambda c: errs[c], 20, 1e-2, 1e-6)
    assert hooks.residual_rank_rule(np.array([-1.0]), 0.0, 0.0) == 1


def test_frobenius_chop_matches_oracle(hooks):
    """ttf_chop (svd.f90:86-111, the typed solver's Frobenius truncation)
    against the oracle's chop, with and without a rank cap, on random
    singular values and on values whose tails sit exactly on the bound (the
    reference keeps a value whose addition reaches the bound: '<')."""
    rng = np.random.default_rng(23)
    for _ in range(3000):
        n = int(rng.integers(1, 60))
        s = np.sort(10.0 ** rng.uniform(-14, 0, n))[::-1]
        tol = 10.0 ** rng.uniform(-12, -0.5)
        rmax = int(rng.integers(1, n + 1)) if rng.random() < 0.3 else None
        assert hooks.chop(s, tol, rmax) == K.chop(s, tol, rmax)[0], (s, tol, rmax)
    # ||s||^2 = 4 and tol 0.5 (both exact): bound 1.0; the tail 1.0 is not < 1.0
    s = np.array([1.0, 1.0, 1.0, 1.0])
    assert hooks.chop(s, 0.5) == K.chop(s, 0.5)[0] == 4
    assert hooks.chop(s, 0.51) == K.chop(s, 0.51)[0] == 3
    # bound 1.5625 (tol 0.25, ||s||^2 = 25): tails 0, 1 dropped, 5 kept
    s = np.array([4.0, 2.0, 2.0, 1.0, 0.0])
    assert hooks.chop(s, 0.25) == K.chop(s, 0.25)[0] == 3
    assert hooks.chop(s, None, 2) == K.chop(s, None, 2)[0] == 2

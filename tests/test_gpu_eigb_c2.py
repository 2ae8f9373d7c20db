"""tt_eigb at the full size of the scaled Heisenberg workload: d = 64 sites,
5 eigenpairs, rank cap 128, tol 1e-10, rank-8 N(0,1) start from seed 0
(local problems up to 128 x 2 x 128 with MPO rank 5).

No exact diagonalisation reaches 2^64 states, so the checks are the
size-independent properties of an eigen-solution, all in FP64:
  * the residual ||H x_b - lambda_b x_b|| of every eigenpair, formed exactly
    in TT format and normed by left-orthogonalisation;
  * orthonormality of the block (X^T X = I up to the truncation);
  * the SU(2) symmetry of the open XXX chain: levels 2-4 are one S = 1
    triplet, so lambda_2 = lambda_3 = lambda_4 up to the truncation;
  * Ritz values of a growing or fixed-rank search space do not increase:
    the eigenvalue sum at the end of each sweep is non-increasing up to the
    truncation's own jitter;
  * honest stopping: converged is 1 exactly when a sweep's ermax is below
    tol, and last_ermax is that last sweep's value.
The bounds are the truncation's: at rank 128 the discarded weight of a
d = 64 chain limits eigenvector accuracy, not the solver.
"""
import numpy as np
import pytest

from oracle import operators as O

pytestmark = pytest.mark.gpu


def test_eigb_heisenberg_d64_rmax128(ttlib):
    import ttfort
    import ttops
    d, B, rmax, tol = 64, 5, 128, 1e-10
    cores, ra, n = O.heisenberg_mpo(d)
    crA = O.flatten_cores(cores)
    rng = np.random.default_rng(0)
    ry = [1] + [8] * (d - 1) + [B]
    crY = np.concatenate([rng.standard_normal(ry[k] * n[k] * ry[k + 1]) for k in range(d)])
    nswp = 4
    lam, res, ry_out, st = ttfort.tt_eigb(n, ra, crA, crY, ry, tol, rmax, B, nswp=nswp)
    print("lambda", lam.tolist())
    print("sweep ermax", st["sweep_ermax"], "lambda sums", st["sweep_lambda_sum"], "converged", st["converged"])
    print("block-preconditioned solves", st["block_prec_solves"], "fallbacks", st["block_prec_fallbacks"])
    assert st["max_rank"] == rmax
    # the middle local problems (order 256 blocks) take the shifted block
    # preconditioner
    assert st["block_prec_solves"] > 0
    assert st["nrec"] == st["sweeps"]
    assert st["last_ermax"] == st["sweep_ermax"][-1]
    assert st["converged"] == (1 if st["last_ermax"] < tol else 0)
    assert np.all(np.diff(lam) >= -1e-12 * np.abs(lam).max())
    r, G = ttops.block_eig_residuals(crA, ra, n, res, ry_out, lam)
    orth = np.abs(G - np.eye(B)).max()
    spread = (lam[3] - lam[1]) / abs(lam[1])
    fv = np.array(st["sweep_lambda_sum"])
    print("residuals", r.tolist(), "orth", orth, "triplet spread", spread, "fv diffs", np.diff(fv).tolist())
    # measured on MI355X: residuals 1.3e-3 .. 1.7e-3 after 4 sweeps (the
    # rank-128 truncation floor of this chain), X^T X - I at 8e-14, triplet
    # spread 2.2e-9, lambda sums decreasing by 3e-5, 2e-8, 4e-9
    assert np.all(r < 5e-3), r
    assert orth < 1e-10, orth
    assert 0.0 <= spread < 1e-7, spread
    assert np.all(np.diff(fv) <= 1e-9 * np.abs(fv[1:])), fv
    # the truncation-limited jitter keeps ermax above tol: not converged
    assert st["converged"] == 0 and st["last_ermax"] > tol

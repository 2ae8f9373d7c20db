"""GPU tests of the concurrent many-chain eigensolver (ttf_tt_eigb_batch):
each chain's eigenvalues against exact diagonalisation of its XXZ
Hamiltonian and against the single-chain driver on the same input, and
chains of the many-chain workload itself (rank cap binding) against the
single-chain driver."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import mpo
from oracle import operators as O

pytestmark = pytest.mark.gpu


def _problems(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        d = int(rng.integers(8, 13))
        delta = float(rng.uniform(0.5, 1.5))
        J = float(rng.uniform(0.8, 1.2))
        crA, ra, n = mpo.xxz(d, J=J, delta=delta)
        crY, ry = mpo.random_block_tt(d, n, 4, 3, seed=seed * 100 + c)
        # rmax 64 is full rank for d <= 12, so the result is exact to the tolerance
        out.append(dict(n=n, ra=ra, crA=crA, crY0=crY, ry=ry, eps=1e-8, rmax=int(rng.integers(56, 65)),
                        d=d, J=J, delta=delta))
    return out


def test_batch_matches_ed_and_single_driver(ttlib):
    import ttfort
    probs = _problems(3, 7)
    out = ttfort.tt_eigb_batch(probs, 3, nswp=10, nworkers=4)
    for p, (lam, cores, ry, st) in zip(probs, out):
        H = O.heisenberg_sparse(p["d"], J=p["J"], delta=p["delta"])
        ed = np.sort(spl.eigsh(H, k=6, which="SA", tol=1e-12)[0])[:3]
        assert np.allclose(np.sort(lam), ed, rtol=1e-7, atol=1e-7), (p["d"], lam, ed)
        lam1, _, _, _ = ttfort.tt_eigb(p["n"], p["ra"], p["crA"], p["crY0"], p["ry"], p["eps"], p["rmax"], 3,
                                       nswp=10)
        assert np.allclose(np.sort(lam), np.sort(lam1), rtol=1e-9, atol=1e-9)
        assert ry[-1] == 3 and cores.size == sum(ry[k] * 2 * ry[k + 1] for k in range(p["d"]))


def test_batch_reports_a_failed_chain(ttlib):
    import ttfort
    probs = _problems(5, 3)
    probs[1]["ry"] = np.array(probs[1]["ry"]).copy()
    probs[1]["ry"][-1] = 2          # inconsistent with B = 3: this chain must fail
    with pytest.raises(ttfort.TTFError, match="1 of 3 chains failed"):
        ttfort.tt_eigb_batch(probs, 3, nswp=4, nworkers=2)


def _c5_chains():
    """The many-chain workload's own distribution (PCG64 seed 5: d ~ U{16..48},
    Delta ~ U(0.5, 1.5), J ~ U(0.8, 1.2), rmax ~ U{32..96}, rank-4 start)."""
    rng = np.random.default_rng(5)
    out = []
    for c in range(256):
        d = int(rng.integers(16, 49))
        delta = float(rng.uniform(0.5, 1.5))
        J = float(rng.uniform(0.8, 1.2))
        r = int(rng.integers(32, 97))
        out.append((c, d, J, delta, r))
    return out


def test_batch_on_workload_chains_matches_single_driver(ttlib):
    """Four of the workload's smallest chains (d 16-17, rmax 33-36: the rank
    cap binds) through the concurrent batch and, one by one, through the
    single-chain driver: the same eigenvalues and ranks."""
    import ttfort
    chains = sorted(_c5_chains(), key=lambda t: t[1] * t[4] ** 3)[:4]
    probs = []
    for c, d, J, delta, r in chains:
        crA, ra, n = mpo.xxz(d, J=J, delta=delta)
        crY, ry = mpo.random_block_tt(d, n, 4, 3, seed=1000 + c)
        probs.append(dict(n=n, ra=ra, crA=crA, crY0=crY, ry=ry, eps=1e-8, rmax=r))
    out = ttfort.tt_eigb_batch(probs, 3, nswp=6, nworkers=4)
    for p, (lam, cores, ry, st) in zip(probs, out):
        lam1, _, ry1, _ = ttfort.tt_eigb(p["n"], p["ra"], p["crA"], p["crY0"], p["ry"], p["eps"], p["rmax"], 3,
                                         nswp=6)
        assert max(ry) == p["rmax"] and max(ry1) == p["rmax"]
        assert np.allclose(np.sort(lam), np.sort(lam1), rtol=1e-7, atol=0), (lam, lam1)

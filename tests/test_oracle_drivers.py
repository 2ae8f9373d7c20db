"""Pins for the oracle's sweep drivers (CPU only)."""
import json
import os

import numpy as np

from oracle import drivers as D
from oracle import operators as O


def test_oracle_eigb_exact_at_full_rank(golden_dir):
    ed = json.load(open(os.path.join(golden_dir, "heisenberg_ed.json")))
    d, B = 8, 5
    cores, ra, n = O.heisenberg_mpo(d)
    crA = O.flatten_cores(cores)
    rng = np.random.default_rng(1)
    ry = [1] + [4] * (d - 1) + [B]
    crY = np.concatenate([rng.standard_normal(ry[k] * n[k] * ry[k + 1]) for k in range(d)])
    lam, cores_out, ryo, st = D.tt_eigb(d, n, ra, crA, crY, ry, 1e-10, 80, B, nswp=6)
    assert np.allclose(lam, ed["8"][:B], atol=1e-10)
    X = O.tt_full(cores_out, n, ryo)
    H = O.mpo_to_dense(cores, ra, n, n)
    assert np.allclose(X.T @ X, np.eye(B), atol=1e-10)
    assert np.linalg.norm(H @ X - X * lam) < 1e-8


def test_oracle_eigb_golden_scenario_consistent(golden_dir):
    """The committed oracle result of the d=20 fixture scenario is above the
    exact levels and within the rank-40 truncation error of them."""
    ed = json.load(open(os.path.join(golden_dir, "heisenberg_ed.json")))
    orc = json.load(open(os.path.join(golden_dir, "eigb_heisen20_oracle.json")))
    ref = np.array(ed["20"][:5])
    lam = np.array(orc["lam"])
    assert np.all(lam >= ref - 1e-12)
    assert np.all(lam - ref <= 5e-6 * np.abs(ref))
    assert max(orc["ry"]) <= orc["rmax"]

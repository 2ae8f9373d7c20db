"""Oracle tt_eigb results on the reference fixture Heisen_20.sdv.

Runs the CPU restatement of tt_eigb (oracle/drivers.py) on the scenario of
the benchmark's small configuration -- B=5 lowest eigenpairs, initial block
TT of rank 4 with N(0,1) cores from numpy's PCG64(seed), rmax 40, tol 1e-6,
max_full_size 200 -- and stores the eigenvalues, sweep count and final ranks
in eigb_heisen20_oracle.json.  The GPU tests compare against these values
and against exact diagonalisation (heisenberg_ed.json).

    python tests/golden/make_eigb_oracle.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import drivers as D  # noqa: E402
from oracle import operators as O  # noqa: E402


def scenario(seed=2, d=20, B=5, r0=4):
    s = O.read_sdv(os.path.join(HERE, "Heisen_20.sdv"))
    n = np.round(np.sqrt(s["n"])).astype(int)
    ra = s["r"]
    crA = O.flatten_cores(s["cores"])
    rng = np.random.default_rng(seed)
    ry = [1] + [r0] * (d - 1) + [B]
    crY = np.concatenate([rng.standard_normal(ry[k] * n[k] * ry[k + 1]) for k in range(d)])
    return n, ra, crA, crY, ry


def main():
    n, ra, crA, crY, ry = scenario()
    t = time.time()
    lam, cores, ryo, st = D.tt_eigb(20, n, ra, crA, crY, ry, 1e-6, 40, 5, nswp=10, max_full_size=200)
    out = dict(seed=2, B=5, rmax=40, tol=1e-6, max_full_size=200, lam=[float(v) for v in lam],
               ry=[int(v) for v in ryo], sweeps=st["sweeps"], seconds=round(time.time() - t, 1))
    print(out)
    with open(os.path.join(HERE, "eigb_heisen20_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

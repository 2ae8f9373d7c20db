"""Exact lowest levels of the open Heisenberg XXX chain (J=1, S=sigma/2).

Writes heisenberg_ed.json with the 8 lowest eigenvalues for d in {8,10,12,20}
computed by sparse Lanczos (scipy.sparse.linalg.eigsh) on the formula
Hamiltonian.  These pin the eigensolver results (tt_eigb) independently of
the reference code; the d=20 case is the Hamiltonian stored in the
reference's fixture Heisen_20.sdv.

    python tests/golden/make_heisenberg_ed.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.operators import heisenberg_sparse  # noqa: E402


def main():
    out = {}
    for d in (8, 10, 12, 20):
        H = heisenberg_sparse(d)
        vals = sla.eigsh(H, k=8, which="SA", tol=1e-14, ncv=40, return_eigenvectors=False)
        out[str(d)] = sorted(float(v) for v in vals)
        print(d, out[str(d)], flush=True)
    with open(os.path.join(HERE, "heisenberg_ed.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

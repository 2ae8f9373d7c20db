"""With TTF_EIG_GRAPHS=1 the FP64 local eigensolver replays its iterations
from captured HIP graphs (one per basis set, recorded after one iteration of
eager launches and kept between local solves of one shape).  The graphs hold
the same kernels with the same arguments as the eager launches, so a whole
tt_eigb run must give the same bits with graphs and without them (the
default), each in a child process: the eigenvalues, the result cores, the
ranks and the matvec counts, over many local solves with their replays and
the host synchronisations between them, and over a second call in the same
process, whose local solves update the graphs kept from the first.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import operators as O

pytestmark = pytest.mark.gpu

_CODE = r'''
import json, os, sys
root, out = sys.argv[1], sys.argv[2]
sys.path[:0] = [root, os.path.join(root, "tt-fort_amd")]
import numpy as np
import ttfort
from oracle import operators as O
d, B, rmax, tol = 16, 5, 32, 1e-10
cores, ra, n = O.heisenberg_mpo(d)
crA = O.flatten_cores(cores)
rng = np.random.default_rng(4)
ry = [1] + [4] * (d - 1) + [B]
crY = np.concatenate([rng.standard_normal(ry[k] * n[k] * ry[k + 1]) for k in range(d)])
res = {}
for rep in range(2):
    lam, cr, ry_out, st = ttfort.tt_eigb(n, ra, crA, crY, ry, tol, rmax, B, nswp=3, max_full_size=50)
    res["lam%d" % rep] = lam.tolist()
    res["ry%d" % rep] = [int(v) for v in ry_out]
    res["mv%d" % rep] = int(st["total_matvecs"])
    np.save(os.path.join(out, "cores%d.npy" % rep), cr)
json.dump(res, open(os.path.join(out, "res.json"), "w"))
'''


def _run(root, out, graphs):
    env = dict(os.environ, TTF_EIG_GRAPHS="1" if graphs else "0")
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([sys.executable, "-c", _CODE, root, out], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.load(open(os.path.join(out, "res.json")))


def test_eigb_graph_replays_match_eager_bits(ttlib, tmp_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    g = _run(root, str(tmp_path / "graphs"), True)
    e = _run(root, str(tmp_path / "eager"), False)
    for rep in range(2):
        assert g["lam%d" % rep] == e["lam%d" % rep]
        assert g["ry%d" % rep] == e["ry%d" % rep]
        assert g["mv%d" % rep] == e["mv%d" % rep]
        cg = np.load(tmp_path / "graphs" / ("cores%d.npy" % rep))
        ce = np.load(tmp_path / "eager" / ("cores%d.npy" % rep))
        assert np.array_equal(cg, ce)
    # and the two calls in one process agree with each other
    assert g["lam0"] == g["lam1"]

"""CPU pins of the oracle's Krylov exponential and KSL step:
* ZGPADM against scipy.linalg.expm on random complex matrices;
* zexp_mv against expm(t A) v, and the identity known answer of the
  reference's test_exp.f90 (exp(t I) v = e^t v);
* ztt_ksl on the reference fixtures: test_eye_ksl.dat has A = i I, so one
  step is exactly e^{i tau} Y; test_ksl.dat against the exact evolution of
  its d = 6 state (the fixed ranks (1,2,4,6,4,2,1) cut the middle bond
  below full rank 8, which bounds the agreement) and conservation of the
  norm for the anti-Hermitian A;
* a full-rank TFIM chain, where the projector-splitting step is exact up
  to the Krylov tolerance.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import mpo
from oracle import drivers as D
from oracle import kernels as K
from oracle import operators as O

from golden.ksl_fixture import read_ksl_dat

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _split(cr, r, n, m=None):
    out, off = [], 0
    for k in range(len(n)):
        sz = int(r[k] * n[k] * (m[k] if m is not None else 1) * r[k + 1])
        out.append(cr[off:off + sz])
        off += sz
    return out


@pytest.mark.parametrize("m", [3, 10, 34])
@pytest.mark.parametrize("t", [0.01, 0.7, -2.5])
def test_zgpadm_vs_expm(m, t):
    rng = np.random.default_rng(m)
    H = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    E, ns = K.zgpadm(6, H, t)
    ref = sl.expm(t * H)
    assert np.abs(E - ref).max() <= 1e-12 * np.abs(ref).max()


def test_zexp_mv_vs_expm_and_identity():
    rng = np.random.default_rng(1)
    A = 0.3 * (rng.standard_normal((50, 50)) + 1j * rng.standard_normal((50, 50)))
    v = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    mv = lambda x: A @ x
    w, _ = K.zexp_mv(mv, v, 30, 0.8, 1e-8, K.onenorm_dense(mv, 50))
    assert np.linalg.norm(w - sl.expm(0.8 * A) @ v) <= 1e-8 * np.linalg.norm(w)
    v = np.full(4, 1j)                                   # test_exp.f90:18-27
    w, _ = K.zexp_mv(lambda x: x, v, 4, 1e-2, 1e-8, 1.0)
    assert np.allclose(w, np.exp(1e-2) * v, rtol=1e-14)


def _ksl_case(name):
    f = read_ksl_dat(os.path.join(GOLDEN, name))
    d, n, ra, ry = f["d"], f["n"], f["ra"], f["ry"]
    y0 = O.tt_full(_split(f["Y"], ry, n), n, ry)
    Ad = O.mpo_to_dense(_split(f["A"], ra, n, n), ra, n, n)
    return f, y0, Ad


def test_ksl_identity_fixture_is_exact():
    f, y0, Ad = _ksl_case("test_eye_ksl.dat")
    assert np.allclose(Ad, 1j * np.eye(Ad.shape[0]))
    cr, ry = D.ztt_ksl(f["d"], f["n"], f["ra"], f["A"], f["Y"], f["ry"], f["tau"])
    x = O.tt_full(cr, f["n"], ry)
    assert np.allclose(x, np.exp(1j * f["tau"]) * y0, rtol=1e-12, atol=1e-13)


def test_ksl_fixture_vs_exact_evolution():
    f, y0, Ad = _ksl_case("test_ksl.dat")
    assert np.abs(Ad + Ad.conj().T).max() == 0.0       # anti-Hermitian: unitary flow
    cr, ry = D.ztt_ksl(f["d"], f["n"], f["ra"], f["A"], f["Y"], f["ry"], f["tau"])
    x = O.tt_full(cr, f["n"], ry)
    ex = sl.expm(f["tau"] * Ad) @ y0
    assert np.linalg.norm(x - ex) <= 2e-6 * np.linalg.norm(ex)
    assert abs(np.linalg.norm(x) - np.linalg.norm(y0)) <= 1e-10 * np.linalg.norm(y0)


def tfim_ksl_problem(d, seed):
    """TFIM chain A = -i H (g = 1) and a full-rank complex start."""
    crH, ra, n = mpo.tfim(d)
    cores = _split(crH, ra, n, n)
    cores[0] = -1j * cores[0]
    crA = np.concatenate([c.astype(np.complex128) for c in cores])
    rng = np.random.default_rng(seed)
    ry = [1] + [int(min(2 ** k, 2 ** (d - k))) for k in range(1, d)] + [1]
    crY = np.concatenate([rng.standard_normal(ry[k] * 2 * ry[k + 1]) + 1j * rng.standard_normal(ry[k] * 2 * ry[k + 1])
                          for k in range(d)])
    return crA, ra, n, crY, np.array(ry)


@pytest.mark.parametrize("typ", [1, 2])
def test_ksl_full_rank_tfim_is_exact(typ):
    d, tau = 6, 0.05
    crA, ra, n, crY, ry = tfim_ksl_problem(d, 4)
    y0 = O.tt_full(_split(crY, ry, n), n, ry)
    Ad = O.mpo_to_dense(_split(crA, ra, n, n), ra, n, n)
    cr, ry2 = D.ztt_ksl(d, n, ra, crA, crY, ry, tau, typ=typ)
    x = O.tt_full(cr, n, ry2)
    ex = sl.expm(tau * Ad) @ y0
    assert np.linalg.norm(x - ex) <= 1e-7 * np.linalg.norm(ex)

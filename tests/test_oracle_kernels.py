"""Pins for the oracle's kernel restatements (CPU only).

Each check uses an independent route -- the Kronecker-built local matrix,
adjoint identities, block inverses of that matrix, closed-form chop cases,
dense solves, the reference's own MPO fixture -- so that a transposed
operand, a dropped term or a wrong index in the oracle fails here.
"""
import json
import os

import numpy as np
import pytest

from oracle import kernels as K
from oracle import operators as O

SHAPES = [
    # rx1, m, rx2, ry1, n, ry2, ra1, ra2
    (1, 2, 1, 1, 2, 1, 1, 1),
    (3, 2, 4, 5, 3, 2, 2, 3),
    (4, 3, 2, 4, 3, 2, 1, 4),
    (6, 2, 5, 6, 2, 5, 5, 5),
    (2, 4, 7, 3, 4, 1, 3, 1),
]


def _rand(rng, shape, cplx=False):
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    g = (lambda k: rng.standard_normal(k) + 1j * rng.standard_normal(k)) if cplx else rng.standard_normal
    return dict(phi1=g(ry1 * rx1 * ra1), A=g(ra1 * n * m * ra2), phi2=g(rx2 * ra2 * ry2),
                x=g(rx1 * m * rx2), y=g(ry1 * n * ry2))


@pytest.mark.parametrize("shape", SHAPES)
def test_bfun3_equals_kron_local_matrix(shape):
    rng = np.random.default_rng(1)
    t = _rand(rng, shape)
    B = K.local_matrix_kron(shape[3], shape[4], shape[5], shape[0], shape[1], shape[2], shape[6],
                            shape[7], t["phi1"], t["A"], t["phi2"])
    y = K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"])
    assert np.allclose(y, B @ t["x"], rtol=1e-13, atol=1e-13 * np.abs(B).sum())


@pytest.mark.parametrize("shape", SHAPES)
def test_zbfun3_and_adjoint(shape):
    rng = np.random.default_rng(2)
    t = _rand(rng, shape, cplx=True)
    B = K.local_matrix_kron(shape[3], shape[4], shape[5], shape[0], shape[1], shape[2], shape[6],
                            shape[7], t["phi1"], t["A"], t["phi2"])
    y = K.zbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"])
    assert np.allclose(y, B @ t["x"], atol=1e-12 * np.abs(B).sum())
    # B^H: <y, B x> == <B^H y, x>
    bt = K.zbfun3_transp(*shape, t["phi1"], t["A"], t["phi2"], t["y"])
    assert np.allclose(np.vdot(t["y"], B @ t["x"]), np.vdot(bt, t["x"]), rtol=1e-11)


@pytest.mark.parametrize("shape", SHAPES)
def test_phi_updates_contract_to_bilinear_form(shape):
    """sum phi_left_new * phi2 == y' B x == sum phi1 * phi_right_new."""
    rng = np.random.default_rng(3)
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    t = _rand(rng, shape)
    yBx = t["y"] @ K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"])
    pl = K.F(K.dphi_left(*shape, t["phi1"], t["A"], t["x"], t["y"]), ry2, rx2, ra2)
    P2 = K.F(t["phi2"], rx2, ra2, ry2)
    assert np.isclose(np.einsum("cba,bac->", pl, P2), yBx, rtol=1e-11)
    pr = K.F(K.dphi_right(*shape, t["phi2"], t["A"], t["x"], t["y"]), rx1, ra1, ry1)
    P1 = K.F(t["phi1"], ry1, rx1, ra1)
    assert np.isclose(np.einsum("cba,bac->", P1, pr), yBx, rtol=1e-11)
    # complex variants conjugate y
    tz = _rand(rng, shape, cplx=True)
    yBxz = np.vdot(tz["y"], K.zbfun3(*shape, tz["phi1"], tz["A"], tz["phi2"], tz["x"]))
    plz = K.F(K.zphi_left(*shape, tz["phi1"], tz["A"], tz["x"], tz["y"]), ry2, rx2, ra2)
    assert np.isclose(np.einsum("cba,bac->", plz, K.F(tz["phi2"], rx2, ra2, ry2)), yBxz, rtol=1e-10)
    prz = K.F(K.zphi_right(*shape, tz["phi2"], tz["A"], tz["x"], tz["y"]), rx1, ra1, ry1)
    assert np.isclose(np.einsum("cba,bac->", K.F(tz["phi1"], ry1, rx1, ra1), prz), yBxz, rtol=1e-10)


def test_phi2_updates():
    rng = np.random.default_rng(4)
    rx1, rx2, ry1, n, ry2 = 3, 4, 2, 3, 5
    p1 = rng.standard_normal(ry1 * rx1)
    p2 = rng.standard_normal(rx2 * ry2)
    x = rng.standard_normal(rx1 * n * rx2)
    y = rng.standard_normal(ry1 * n * ry2)
    # y' (phi1 (x) I (x) phi2) x via the kron local matrix with a rank-1 identity operator
    A = np.eye(n).reshape(-1, order="F")
    B = K.local_matrix_kron(ry1, n, ry2, rx1, n, rx2, 1, 1, p1, A, p2.reshape(-1))
    yBx = y @ B @ x
    l = K.F(K.dphi2_left(rx1, rx2, ry1, n, ry2, p1, x, y), ry2, rx2)
    assert np.isclose(np.sum(l * K.F(p2, rx2, ry2).T), yBx, rtol=1e-12)
    r = K.F(K.dphi2_right(rx1, rx2, ry1, n, ry2, p2, x, y), rx1, ry1)
    assert np.isclose(np.sum(r * K.F(p1, ry1, rx1).T), yBx, rtol=1e-12)


def test_fullmat_equals_kron():
    rng = np.random.default_rng(5)
    r1, n, r2, ra1, ra2 = 3, 2, 4, 3, 2
    p1 = rng.standard_normal(r1 * r1 * ra1)
    A = rng.standard_normal(ra1 * n * n * ra2)
    p2 = rng.standard_normal(r2 * ra2 * r2)
    Bf = K.d2d_fullmat(r1, n, r2, ra1, ra2, p1, A, p2)
    Bk = K.local_matrix_kron(r1, n, r2, r1, n, r2, ra1, ra2, p1, A, p2)
    assert np.allclose(Bf, Bk, atol=1e-13)


def test_chop_rules_hand_cases():
    s = np.array([1.0, 0.1, 0.01, 0.001])
    # my_chop3 keeps one more than the minimal rank (it stops at the first
    # index whose tail sum exceeds the bound and returns that index + 1)
    assert K.my_chop3(s, 0.05) == 3
    assert K.my_chop3(s, 1e-5) == 4
    assert K.my_chop3(s, 0.5) == 2
    assert K.my_chop3(np.zeros(5), 1e-3) == 1
    # svd.f90 chop: minimal rank with tail below tol*||s||
    assert K.chop(s, 0.05)[0] == 2
    assert K.chop(s, 1e-5)[0] == 4
    r, err = K.chop(s, None, rmax=2)
    assert r == 2 and np.isclose(err, np.sqrt(1e-4 + 1e-6))


def test_dqr_invariants():
    rng = np.random.default_rng(6)
    for n, m in ((12, 5), (5, 9), (7, 7)):
        A = rng.standard_normal(n * m)
        Q, R = K.dqr(n, m, A)
        k = min(n, m)
        Qm, Rm = K.F(Q, n, k), K.F(R, k, m)
        assert np.allclose(Qm.T @ Qm, np.eye(k), atol=1e-13)
        assert np.allclose(Qm @ Rm, K.F(A, n, m), atol=1e-12)
        assert np.allclose(np.tril(Rm, -1), 0)


@pytest.mark.parametrize("ptype", ["c", "l", "r"])
def test_block_jacobi_inverts_diagonal_blocks(ptype):
    rng = np.random.default_rng(7)
    rx1, n, rx2, ra1, ra2 = 3, 4, 2, 2, 3
    p1 = rng.standard_normal(rx1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * n * ra2)
    p2 = rng.standard_normal(rx2 * ra2 * rx2)
    B = K.local_matrix_kron(rx1, n, rx2, rx1, n, rx2, ra1, ra2, p1, A, p2)
    N = rx1 * n * rx2
    idx = np.arange(N).reshape((rx1, n, rx2), order="F")
    # block-diagonal part of B selected by the preconditioner type
    if ptype == "c":
        groups = [idx[b1, :, b2] for b2 in range(rx2) for b1 in range(rx1)]
    elif ptype == "l":
        groups = [idx[:, :, b2].reshape(-1, order="F") for b2 in range(rx2)]
    else:
        groups = [idx[b1, :, :].reshape(-1, order="F") for b1 in range(rx1)]
    D = np.zeros_like(B)
    for g in groups:
        D[np.ix_(g, g)] = B[np.ix_(g, g)]
    jacs = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, p1, A, p2)
    x = rng.standard_normal(N)
    y = K.djac_apply(ptype, rx1, n, rx2, jacs, x)
    assert np.allclose(D @ y, x, atol=1e-9 * np.abs(x).max())


@pytest.mark.parametrize("ptype", ["n", "c"])
def test_gmres_solves_local_system(ptype):
    rng = np.random.default_rng(8)
    rx1, n, rx2, ra1, ra2 = 3, 3, 3, 2, 2
    shape = (rx1, n, rx2, rx1, n, rx2, ra1, ra2)
    p1 = rng.standard_normal(rx1 * rx1 * ra1) * 0.2
    A = rng.standard_normal(ra1 * n * n * ra2) * 0.2
    p2 = rng.standard_normal(rx2 * ra2 * rx2) * 0.2
    # make the operator diagonally dominant through the first bond term
    P1 = K.F(p1, rx1, rx1, ra1).copy(); P1[:, :, 0] += 3 * np.eye(rx1)
    AA = K.F(A, ra1, n, n, ra2).copy(); AA[0, :, :, 0] += np.eye(n)
    P2 = K.F(p2, rx2, ra2, rx2).copy(); P2[:, 0, :] += np.eye(rx2)
    p1, A, p2 = [np.reshape(a, -1, order="F") for a in (P1, AA, P2)]
    B = K.local_matrix_kron(rx1, n, rx2, rx1, n, rx2, ra1, ra2, p1, A, p2)
    rhs = rng.standard_normal(B.shape[0])
    mv = lambda v: K.dbfun3(*shape, p1, A, p2, v)
    if ptype == "n":
        sol, info = K.dgmresr_hh(mv, rhs, 40, 1e-12, 3)
        x = sol
    else:
        jacs = K.djac_gen("c", rx1, n, rx2, ra1, ra2, p1, A, p2)
        P = lambda v: K.djac_apply("c", rx1, n, rx2, jacs, v)
        sol, info = K.dgmresr_hh(mv, rhs, 40, 1e-12, 3, prec=P)
        x = P(sol)
    # GMRES solves B P sol = rhs (tt_adapt_als.f90:771-781 then subtracts P sol)
    xs = np.linalg.solve(B, rhs)
    assert np.linalg.norm(x - xs) < 1e-9 * np.linalg.norm(xs)
    assert info["relres"] < 1e-12


def test_heisenberg_mpo_matches_reference_fixture(golden_dir):
    s = O.read_sdv(os.path.join(golden_dir, "Heisen_20.sdv"))
    d = len(s["cores"])
    assert d == 20 and s["dtype"] == np.float64
    nsd = np.round(np.sqrt(s["n"])).astype(int)
    cores, ra, n = O.heisenberg_mpo(d)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(2 ** d)
    y1 = O.mpo_apply_full(s["cores"], s["r"], nsd, nsd, v)
    y2 = O.mpo_apply_full(cores, ra, n, n, v)
    assert np.linalg.norm(y1 - y2) < 1e-13 * np.linalg.norm(y1)
    e = np.zeros(2 ** d); e[0] = 1.0
    assert np.isclose(e @ O.mpo_apply_full(s["cores"], s["r"], nsd, nsd, e), (d - 1) / 4)


def test_heisenberg_ed_golden_consistent(golden_dir):
    ed = json.load(open(os.path.join(golden_dir, "heisenberg_ed.json")))
    cores, ra, n = O.heisenberg_mpo(8)
    H = O.mpo_to_dense(cores, ra, n, n)
    w = np.linalg.eigvalsh(0.5 * (H + H.T))[:8]
    assert np.allclose(w, ed["8"], atol=1e-10)
    assert np.isclose(ed["20"][0], -8.682473334399, atol=1e-11)


# shapes with ra > 1 and enough rows on the contracted side (ry1 >= rx1 ra1,
# ry2 >= rx2 ra2) that contracting with a random interface loses nothing
HALF_SHAPES = [
    # rx1, m, rx2, ry1, n, ry2, ra1, ra2
    (3, 2, 4, 7, 2, 9, 2, 2),
    (2, 3, 2, 9, 3, 7, 4, 3),
    (1, 2, 3, 3, 2, 3, 3, 1),
]


@pytest.mark.parametrize("shape", HALF_SHAPES)
def test_bfun3_right_half_contracts_to_bfun3(shape):
    """sum_{b1,a1} phi1(c1,b1,a1) right(b1,a1,i,c2) == (B x)(c1,i,c2), with B
    the Kronecker-built local matrix; phi1 of full column rank makes the
    check pin every entry of the half product."""
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    rng = np.random.default_rng(11)
    t = _rand(rng, shape)
    half = K.dbfun3_right(*shape, t["A"], t["phi2"], t["x"])
    assert half.size == rx1 * ra1 * n * ry2
    P1 = K.F(t["phi1"], ry1, rx1 * ra1)
    assert np.linalg.matrix_rank(P1) == rx1 * ra1
    y = (P1 @ K.F(half, rx1 * ra1, n * ry2)).reshape(-1, order="F")
    B = K.local_matrix_kron(ry1, n, ry2, rx1, m, rx2, ra1, ra2, t["phi1"], t["A"], t["phi2"])
    assert np.allclose(y, B @ t["x"], rtol=1e-13, atol=1e-13 * np.abs(B).sum())
    # a transposed A (i <-> j) in the half product would fail the same check
    At = np.transpose(K.F(t["A"], ra1, n, m, ra2), (0, 2, 1, 3)).reshape(-1, order="F")
    if n == m:
        bad = K.dbfun3_right(*shape, At, t["phi2"], t["x"])
        ybad = (P1 @ K.F(bad, rx1 * ra1, n * ry2)).reshape(-1, order="F")
        assert not np.allclose(ybad, B @ t["x"], rtol=1e-6)


@pytest.mark.parametrize("shape", HALF_SHAPES)
def test_bfun3_left_half_contracts_to_bfun3(shape):
    """sum_{b2,a2} left(c1,i,b2,a2) phi2(b2,a2,c2) == (B x)(c1,i,c2)."""
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    rng = np.random.default_rng(12)
    t = _rand(rng, shape)
    half = K.dbfun3_left_intended(*shape, t["phi1"], t["A"], t["x"])
    assert half.size == ry1 * n * rx2 * ra2
    P2 = K.F(t["phi2"], rx2 * ra2, ry2)
    y = (K.F(half, ry1 * n, rx2 * ra2) @ P2).reshape(-1, order="F")
    B = K.local_matrix_kron(ry1, n, ry2, rx1, m, rx2, ra1, ra2, t["phi1"], t["A"], t["phi2"])
    assert np.allclose(y, B @ t["x"], rtol=1e-13, atol=1e-13 * np.abs(B).sum())


def _rank_closed_form(errs, res_new, eps2):
    """One more than the largest candidate c in 1..k-1 whose residual exceeds
    both bounds; 1 if there is none."""
    above = [c for c in range(1, len(errs)) if errs[c] > 2 * res_new and errs[c] > eps2]
    return (max(above) if above else 0) + 1


def test_residual_rank_rule_closed_form():
    rng = np.random.default_rng(13)
    for _ in range(2000):
        k = int(rng.integers(1, 40))
        errs = np.concatenate([[-1.0], np.sort(10.0 ** rng.uniform(-12, 0, k - 1))[::-1]])
        if rng.random() < 0.3:
            rng.shuffle(errs[1:])   # non-monotone residual sequences too
        res_new = 10.0 ** rng.uniform(-12, -1)
        eps2 = 10.0 ** rng.uniform(-12, -1)
        got = K.residual_rank(lambda c: errs[c], k, res_new, eps2)
        assert got == _rank_closed_form(errs, res_new, eps2)
    # exact ties: a residual equal to 2 res_new (or to eps2) does not stop
    errs = np.array([-1.0, 4.0, 2.0, 1.0, 0.5])
    assert K.residual_rank(lambda c: errs[c], 5, 1.0, 0.1) == 2      # 2.0 == 2*1.0 does not stop
    assert K.residual_rank(lambda c: errs[c], 5, 0.5, 2.0) == 2      # 2.0 == eps2 does not stop
    assert K.residual_rank(lambda c: errs[c], 5, 0.49, 0.1) == 4      # 1.0 > 0.98 stops at 3
    assert K.residual_rank(lambda c: errs[c], 1, 0.0, 0.0) == 1      # nothing to scan


def test_amen_truncation_rank_properties():
    """The chosen rank's truncation meets the bound, one rank less does not,
    and the full-rank candidate reproduces the core's own residual."""
    rng = np.random.default_rng(14)
    for direction in (1, -1):
        rx1, n, rx2, ra1, ra2 = 5, 3, 6, 2, 3
        phi1 = rng.standard_normal(rx1 * rx1 * ra1)
        A = rng.standard_normal(ra1 * n * n * ra2)
        phi2 = rng.standard_normal(rx2 * ra2 * rx2)
        rows, cols = (rx1 * n, rx2) if direction > 0 else (rx1, n * rx2)
        k = min(rows, cols)
        Uq, _ = np.linalg.qr(rng.standard_normal((rows, k)))
        Vq, _ = np.linalg.qr(rng.standard_normal((cols, k)))
        s = 10.0 ** -np.arange(k)
        x = ((Uq * s) @ Vq.T).reshape(-1, order="F")
        shape = (rx1, n, rx2, rx1, n, rx2, ra1, ra2)
        rhs = K.dbfun3(*shape, phi1, A, phi2, x)
        for res_new in (1e-9, 1e-5, 1e-2):
            r, errs = K.amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs, res_new, 1e-12)
            bound = max(2 * res_new, 1e-12)
            assert 1 <= r <= k
            if r < k:
                assert errs[r] <= bound
            if r > 1:
                assert errs[r - 1] > bound
            # graded singular values: residuals fall with the rank
            assert np.all(np.diff(errs[1:]) < 0)


def test_row_add_and_cut_are_inverse_and_index_exactly():
    rng = np.random.default_rng(15)
    m, n, k = 5, 4, 3
    A = rng.standard_normal(m * n)
    B = rng.standard_normal(k * n)
    C = K.drow_add(m, n, k, A, B)
    Cm = K.F(C, m + k, n)
    for j in range(n):
        for i in range(m + k):
            assert Cm[i, j] == (A[i + m * j] if i < m else B[(i - m) + k * j])
    assert np.array_equal(K.drow_cut(m + k, n, m, C), A)
    assert np.array_equal(K.drow_cut(m, n, m, A), A)


def test_transp_and_perm1324_index_exactly():
    rng = np.random.default_rng(16)
    n1, n2, n3, n4 = 2, 3, 4, 5
    A = rng.standard_normal(n1 * n2 * n3 * n4)
    B = K.dperm1324(n1, n2, n3, n4, A)
    for i4 in range(n4):
        for i3 in range(n3):
            for i2 in range(n2):
                for i1 in range(n1):
                    assert B[i1 + n1 * (i3 + n3 * (i2 + n2 * i4))] == A[i1 + n1 * (i2 + n2 * (i3 + n3 * i4))]
    M = rng.standard_normal(6 * 7)
    T = K.dtransp(6, 7, M)
    assert all(T[j + 7 * i] == M[i + 6 * j] for i in range(6) for j in range(7))

"""The reference-named C symbols (tt_linalg.f90's bind(c) routines and the
PRIMME / Expokit callback shapes) called through the C ABI exactly as a
Fortran caller would: every argument by reference, Fortran default
integers of the library's kind (int32 in libttfort_amd.so, int64 in
libttfort_amd_i8.so), host arrays.  Each is compared with the oracle on the
same inputs: data movement (drow_add, drow_cut, dtransp, dperm1324) bit for
bit, arithmetic to the contraction bar (1e-12 relative), the QR against
LAPACK's dgeqrf + dorgqr (numpy.linalg.qr) with the same signs, GMRES to its
tolerance.
"""
import ctypes

import numpy as np
import pytest

from oracle import kernels as K

pytestmark = pytest.mark.gpu

KINDS = ["i4", "i8"]


def _lib(kind):
    import ttfort
    lib = ttfort.load(kind)
    ttfort.check(lib.ttf_init(0), lib)
    lib.ttf_clear_error()
    return lib


def _int(kind, v):
    return (ctypes.c_int32 if kind == "i4" else ctypes.c_int64)(v)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ok(lib):
    msg = lib.ttf_last_error()
    assert not msg, msg.decode()


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n,m", [(12, 5), (5, 12), (64, 64), (300, 40), (1, 4), (7, 1)])
def test_dqr(kind, n, m):
    lib = _lib(kind)
    rng = np.random.default_rng(n * 100 + m)
    A0 = rng.standard_normal((n, m))
    k = min(n, m)
    A = np.asfortranarray(A0.copy())
    R = np.full(k * max(n, m) + 5, 7.0)   # room for R(k, max(n, m)) plus a guard
    work = np.zeros(4)
    tau = np.zeros(4)
    lib.dqr(ctypes.byref(_int(kind, n)), ctypes.byref(_int(kind, m)), _p(A), _p(R), _p(work),
            ctypes.byref(_int(kind, 4)), _p(tau))
    _ok(lib)
    Qr, Rr = np.linalg.qr(A0, mode="reduced")   # dgeqrf + dorgqr
    Q = A[:, :k]
    assert np.allclose(Q, Qr, atol=1e-12), np.abs(Q - Qr).max()
    Rm = R[:k * m].reshape((k, m), order="F")
    assert np.allclose(Rm, Rr, atol=1e-12 * np.abs(Rr).max()), np.abs(Rm - Rr).max()
    assert np.all(np.tril(Rm, -1) == 0.0)
    # the reference zeroes R(min(n,m), n) and writes columns 1..m: nothing past both
    assert np.all(R[k * max(n, m):] == 7.0)
    if n > m:
        assert np.all(R[k * m:k * n] == 0.0)
    if m > n:   # dgeqrf leaves R's trailing columns in A
        assert np.allclose(A[:, k:], Rr[:, k:], atol=1e-12 * np.abs(Rr).max())


@pytest.mark.parametrize("kind", KINDS)
def test_row_add_cut_transp_perm(kind):
    lib = _lib(kind)
    rng = np.random.default_rng(3)
    m, n, k = 6, 5, 3
    A = rng.standard_normal(m * n)
    B = rng.standard_normal(k * n)
    buf = np.zeros((m + k) * n)
    buf[:m * n] = A
    lib.drow_add(ctypes.byref(_int(kind, m)), ctypes.byref(_int(kind, n)), ctypes.byref(_int(kind, k)), _p(buf),
                 _p(B))
    _ok(lib)
    assert np.array_equal(buf, K.drow_add(m, n, k, A, B))
    cut = buf.copy()
    lib.drow_cut(ctypes.byref(_int(kind, m + k)), ctypes.byref(_int(kind, n)), ctypes.byref(_int(kind, m)),
                 _p(cut))
    _ok(lib)
    assert np.array_equal(cut[:m * n], K.drow_cut(m + k, n, m, buf))
    assert np.array_equal(cut[m * n:], buf[m * n:])   # the rest of A is left as it was
    T = np.zeros(m * n)
    lib.dtransp(ctypes.byref(_int(kind, m)), ctypes.byref(_int(kind, n)), _p(A), _p(T))
    _ok(lib)
    assert np.array_equal(T, K.dtransp(m, n, A))
    dims = (3, 4, 5, 2)
    P = rng.standard_normal(int(np.prod(dims)))
    Q = np.zeros_like(P)
    lib.dperm1324(*[ctypes.byref(_int(kind, v)) for v in dims], _p(P), _p(Q))
    _ok(lib)
    assert np.array_equal(Q, K.dperm1324(*dims, P))


def _local(rng, rx1, n, rx2, ra1, ra2):
    # symmetric positive interfaces so that the Jacobi blocks are invertible
    def spd(r, ra):
        out = np.empty((r, r, ra))
        for a in range(ra):
            M = rng.standard_normal((r, r))
            out[:, :, a] = M @ M.T + r * np.eye(r)
        return out
    P1 = spd(rx1, ra1).reshape(-1, order="F")
    P2 = np.transpose(spd(rx2, ra2), (0, 2, 1)).reshape(-1, order="F")   # (rx2, ra2, rx2)
    A = np.zeros((ra1, n, n, ra2))
    for a1 in range(ra1):
        for a2 in range(ra2):
            M = rng.standard_normal((n, n))
            A[a1, :, :, a2] = M @ M.T + n * np.eye(n)
    return P1, A.reshape(-1, order="F"), P2


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ptype", ["c", "l", "R"])
def test_djac_gen_apply(kind, ptype):
    lib = _lib(kind)
    rng = np.random.default_rng(4)
    rx1, n, rx2, ra1, ra2 = 4, 5, 3, 2, 2
    P1, A, P2 = _local(rng, rx1, n, rx2, ra1, ra2)
    nj = int(lib.ttf_djac_size(ptype.encode(), rx1, n, rx2))
    jacs = np.zeros(nj)
    pt = ctypes.c_char(ptype.encode())
    ints = [ctypes.byref(_int(kind, v)) for v in (rx1, n, rx2, ra1, ra2)]
    lib.djac_gen(ctypes.byref(pt), *ints, _p(P1), _p(A), _p(P2), _p(jacs))
    _ok(lib)
    ref = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, P1, A, P2)
    assert np.allclose(jacs, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    x = rng.standard_normal(rx1 * n * rx2)
    y = np.zeros_like(x)
    work = np.zeros(1)
    lib.djac_apply(ctypes.byref(pt), *ints[:3], _p(jacs), _p(x), _p(y), _p(work))
    _ok(lib)
    yr = K.djac_apply(ptype, rx1, n, rx2, ref, x)
    assert np.allclose(y, yr, rtol=1e-11, atol=1e-12 * np.abs(yr).max())


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("trans,m,n,numvec", [("N", 30, 50, 4), ("T", 40, 64, 4)])
def test_duchol_fort(kind, trans, m, n, numvec):
    lib = _lib(kind)
    rng = np.random.default_rng(m + n)
    if trans == "N":
        M = rng.standard_normal((m, n)) * np.logspace(0, -3, n)[None, :]
    else:
        M = rng.standard_normal((n, m)) * np.logspace(0, -3, n)[:, None]
    dA = M.reshape(-1, order="F")
    du = np.zeros(n * numvec)
    lam = np.zeros(numvec)
    real = _int(kind, 0)
    tr = ctypes.c_char(trans.encode())
    lib.duchol_fort(ctypes.byref(tr), ctypes.byref(_int(kind, m)), ctypes.byref(_int(kind, n)), _p(dA),
                    ctypes.byref(_int(kind, numvec)), _p(du), _p(lam), ctypes.byref(real))
    _ok(lib)
    du_r, lam_r, real_r = K.duchol(trans, m, n, dA, numvec)
    assert real.value == real_r
    U = du.reshape((n, numvec), order="F")[:, :real_r]
    assert np.allclose(lam[:real_r], lam_r[:real_r], rtol=1e-9)
    assert np.allclose(U @ U.T, du_r @ du_r.T, atol=1e-9)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ptype", ["n", "c"])
def test_dgmresr_hh_fort(kind, ptype):
    lib = _lib(kind)
    rng = np.random.default_rng(5)
    rx1, n, rx2, ra1, ra2 = 4, 6, 5, 2, 2
    P1, A, P2 = _local(rng, rx1, n, rx2, ra1, ra2)
    shape = (rx1, n, rx2, rx1, n, rx2, ra1, ra2)
    rhs = rng.standard_normal(rx1 * n * rx2)
    jacs = np.zeros(1)
    if ptype != "n":
        jacs = K.djac_gen(ptype, rx1, n, rx2, ra1, ra2, P1, A, P2)
    sol = np.zeros_like(rhs)
    tol = np.array([1e-10])
    pt = ctypes.c_char(ptype.encode())
    lib.dgmresr_hh_fort(_p(P1), _p(A), _p(P2), _p(rhs), *[ctypes.byref(_int(kind, v)) for v in
                                                         (rx1, n, rx2, ra1, ra2, 40)],
                        _p(tol), ctypes.byref(_int(kind, 4)), ctypes.byref(pt), _p(jacs), _p(sol),
                        ctypes.byref(_int(kind, 0)))
    _ok(lib)
    x = sol if ptype == "n" else K.djac_apply(ptype, rx1, n, rx2, jacs, sol)
    Bx = K.dbfun3(*shape, P1, A, P2, x)
    assert np.linalg.norm(Bx - rhs) <= 1e-9 * np.linalg.norm(rhs)
    B = K.local_matrix_kron(rx1, n, rx2, rx1, n, rx2, ra1, ra2, P1, A, P2)
    assert np.allclose(x, np.linalg.solve(B, rhs), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("kind", KINDS)
def test_callbacks_real_and_complex(kind):
    lib = _lib(kind)
    rng = np.random.default_rng(6)
    shape = (7, 2, 6, 5, 3, 4, 3, 2)
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    P1 = rng.standard_normal(ry1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * m * ra2)
    P2 = rng.standard_normal(rx2 * ra2 * ry2)
    xs, ys = rx1 * m * rx2 + 3, ry1 * n * ry2 + 2   # leading dimensions with padding
    lib.ttf_init_bfun_sizes(*[ctypes.byref(_int(kind, v)) for v in shape + (xs, ys)])
    lib.ttf_init_bfun_main(_p(P1), _p(A), _p(P2))
    _ok(lib)
    k = 4
    X = rng.standard_normal(xs * k)
    Y = np.zeros(ys * k)
    lib.ttf_primme_matvec(_p(X), _p(Y), ctypes.byref(_int(kind, k)), None)
    _ok(lib)
    for b in range(k):
        ref = K.dbfun3(*shape, P1, A, P2, X[b * xs:b * xs + rx1 * m * rx2])
        got = Y[b * ys:b * ys + ry1 * n * ry2]
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # complex: Expokit's matvec and normest's transposed product
    c = lambda sz: rng.standard_normal(sz) + 1j * rng.standard_normal(sz)   # noqa: E731
    Z1, ZA, Z2 = c(ry1 * rx1 * ra1), c(ra1 * n * m * ra2), c(rx2 * ra2 * ry2)
    lib.ttf_zinit_bfun_main(_p(Z1), _p(ZA), _p(Z2))
    _ok(lib)
    zx = c(rx1 * m * rx2)
    zy = np.zeros(ry1 * n * ry2, dtype=np.complex128)
    lib.ttf_zmatvec(_p(zx), _p(zy))
    _ok(lib)
    ref = K.zbfun3(*shape, Z1, ZA, Z2, zx)
    assert np.allclose(zy, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    zt = np.zeros(rx1 * m * rx2, dtype=np.complex128)
    lib.ttf_zmatvec_transp(_p(zy), _p(zt))
    _ok(lib)
    ref = K.zbfun3_transp(*shape, Z1, ZA, Z2, zy)
    assert np.allclose(zt, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())

"""ctypes binding of the C ABI in include/ttfort_amd.h.

This is the Python face of the library used by the tests and the benchmark:
it loads lib/libttfort_amd.so (built by build.py) and raises if the library
is missing -- there is no CPU fallback.  Device arrays are `DeviceArray`
objects wrapping memory from ttf_malloc; helpers upload/download numpy
arrays in column-major (Fortran) order, the layout of the reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

_c_int = C.c_int
_c_i64 = C.c_int64
_vp = C.c_void_p


class TTFError(RuntimeError):
    pass


def lib_path(kind: str = "i4") -> str:
    name = "libttfort_amd.so" if kind == "i4" else "libttfort_amd_i8.so"
    return os.path.join(LIBDIR, name)


_LIB = {}


def load(kind: str = "i4"):
    """Load (once) and return the ctypes CDLL of the HIP library."""
    if kind in _LIB:
        return _LIB[kind]
    path = lib_path(kind)
    if not os.path.exists(path):
        raise TTFError("HIP library not built: %s (run python tt-fort_amd/build.py)" % path)
    lib = C.CDLL(path)
    _declare(lib)
    _LIB[kind] = lib
    return lib


# (name, restype, argtypes)
_SIGS = []


def _sig(name, res, *args):
    _SIGS.append((name, res, list(args)))


_S8 = [_c_int] * 8
_sig("ttf_init", _c_int, _c_int)
_sig("ttf_set_local_precision", _c_int, _c_int)
_sig("ttf_get_local_precision", _c_int)
_sig("ttf_finalize", _c_int)
_sig("ttf_last_error", C.c_char_p)
_sig("ttf_clear_error", None)
_sig("ttf_version", C.c_char_p)
_sig("ttf_malloc", _c_int, C.POINTER(_vp), C.c_size_t)
_sig("ttf_free", _c_int, _vp)
_sig("ttf_memcpy_h2d", _c_int, _vp, _vp, C.c_size_t)
_sig("ttf_memcpy_d2h", _c_int, _vp, _vp, C.c_size_t)
_sig("ttf_memcpy_d2d", _c_int, _vp, _vp, C.c_size_t)
_sig("ttf_memset", _c_int, _vp, _c_int, C.c_size_t)
_sig("ttf_sync", _c_int)
_sig("ttf_timer_start", _c_int)
_sig("ttf_timer_stop", _c_int, C.POINTER(C.c_float))
_sig("ttf_set_autotune", _c_int, _c_int)
_sig("ttf_autotune_count", C.c_long)
for _nm in ("ttf_dbfun3", "ttf_dbfun3_noalloc", "ttf_sbfun3", "ttf_zbfun3", "ttf_zbfun3_transp", "ttf_cbfun3"):
    _sig(_nm, _c_int, *(_S8 + [_c_int, _vp, _vp, _vp, _vp, _c_i64, _vp, _c_i64]))
_sig("ttf_dbfun3_profile", _c_int, *(_S8 + [_c_int, _vp, _vp, _vp, _vp, _c_i64, _vp, _c_i64, _c_int,
                                           C.POINTER(C.c_float)]))
_sig("ttf_dbfun3_chain_profile", _c_int, *(_S8 + [_c_int, _vp, _vp, _vp, _vp, _c_i64, _vp, _c_i64, _c_int,
                                                 C.POINTER(C.c_float), C.POINTER(C.c_int)]))
for _nm in ("ttf_dphi_left", "ttf_dphi_right", "ttf_zphi_left", "ttf_zphi_right"):
    _sig(_nm, _c_int, *(_S8 + [_vp, _vp, _vp, _vp, _vp]))
for _nm in ("ttf_dphi2_left", "ttf_dphi2_right"):
    _sig(_nm, _c_int, *([_c_int] * 5 + [_vp, _vp, _vp, _vp]))
_sig("ttf_d2d_fullmat", _c_int, *([_c_int] * 5 + [_vp, _vp, _vp, _vp]))
_sig("dbfun3", None, *([_vp] * 15))
_sig("dqr", None, *([_vp] * 7))
_sig("drow_add", None, *([_vp] * 5))
_sig("drow_cut", None, *([_vp] * 4))
_sig("dtransp", None, *([_vp] * 4))
_sig("dperm1324", None, *([_vp] * 6))
_sig("djac_gen", None, *([_vp] * 10))
_sig("djac_apply", None, *([_vp] * 8))
_sig("duchol_fort", None, *([_vp] * 8))
_sig("dgmresr_hh_fort", None, *([_vp] * 16))
_sig("ttf_init_bfun_sizes", None, *([_vp] * 10))
_sig("ttf_init_bfun_main", None, _vp, _vp, _vp)
_sig("ttf_zinit_bfun_main", None, _vp, _vp, _vp)
_sig("ttf_primme_matvec", None, _vp, _vp, _vp, _vp)
_sig("ttf_zmatvec", None, _vp, _vp)
_sig("ttf_zmatvec_transp", None, _vp, _vp)
_sig("ttf_dqr", _c_int, _c_int, _c_int, _vp, _vp, _c_i64)
_sig("ttf_dgesvd", _c_int, _c_int, _c_int, _vp, _c_i64, _vp, _c_i64, _vp, _vp, _c_i64,
     C.POINTER(_c_int))
_sig("ttf_dsvd_side", _c_int, _c_int, _c_int, _vp, _c_i64, C.c_char, _vp, _vp, _c_i64,
     C.POINTER(_c_int), C.POINTER(_c_int))
_sig("ttf_dsyev_small", _c_int, _c_int, _c_int, _vp, _vp, _vp, _c_int)
_sig("ttf_dgram", _c_int, _c_int, _c_int, _c_int, _vp, _c_i64, _vp, _c_i64, _vp, _c_int)


class EigbStats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("total_matvecs", C.c_long), ("last_ermax", C.c_double),
                ("max_rank", C.c_int), ("seconds", C.c_double), ("solve_seconds", C.c_double),
                ("trunc_seconds", C.c_double), ("phi_seconds", C.c_double), ("nrec", C.c_int),
                ("sweep_seconds", C.c_double * 64), ("sweep_flops", C.c_double * 64),
                ("sweep_matvecs", C.c_long * 64), ("sweep_ermax", C.c_double * 64),
                ("sweep_lambda_sum", C.c_double * 64), ("converged", C.c_int),
                ("total_matvecs32", C.c_long), ("refreshes", C.c_long), ("sweep_matvecs32", C.c_long * 64),
                ("block_prec_solves", C.c_long), ("block_prec_fallbacks", C.c_long)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v)[:self.nrec] if k.startswith("sweep_") else v
        return out


_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_sig("ttf_tt_eigb", _c_int, _c_int, _IP, _IP, _IP, _DP, _DP, _IP, C.c_double, _c_int, _DP, _c_int,
     _c_int, _c_int, _c_int, _c_int, _DP, _c_i64, C.POINTER(_c_i64), C.POINTER(EigbStats))


class AmenStats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("total_matvecs", C.c_long), ("last_res_max", C.c_double),
                ("last_err_max", C.c_double), ("max_rank", C.c_int), ("seconds", C.c_double),
                ("solve_seconds", C.c_double), ("trunc_seconds", C.c_double), ("phi_seconds", C.c_double),
                ("flops", C.c_double), ("learned_solves", C.c_long), ("learned_fallbacks", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_sig("ttf_my_chop3", _c_int, _c_int, _DP, C.c_double)
_sig("ttf_chop", _c_int, _c_int, _DP, C.c_double, _c_int)
_sig("ttf_residual_rank_rule", _c_int, _c_int, _DP, C.c_double, C.c_double)
_sig("ttf_amen_truncation_rank", _c_int, _c_int, *([_c_int] * 5), _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double,
     _IP, _DP)
_sig("ttf_eig_block_prec", _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, C.c_double, _vp,
     C.POINTER(C.c_int))
_sig("ttf_djac_gen", _c_int, C.c_char, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp)
_sig("ttf_djac_size", _c_i64, C.c_char, _c_int, _c_int, _c_int)
_sig("ttf_djac_apply", _c_int, C.c_char, _c_int, _c_int, _c_int, _vp, _vp, _vp)
_sig("ttf_dgmresr", _c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, C.c_double,
     _c_int, C.c_char, _vp, _vp, _IP, _DP)
_sig("ttf_duchol", _c_int, C.c_char, _c_int, _c_int, _vp, _c_int, _vp, _DP, _IP)
_sig("ttf_dbfun3_left", _c_int, *_S8, _vp, _vp, _vp, _vp)
_sig("ttf_dbfun3_right", _c_int, *_S8, _vp, _vp, _vp, _vp)
_sig("ttf_dtt_amen_solve", _c_int, _c_int, _IP, _IP, _DP, _IP, _DP, C.c_double, _IP, _DP, _c_int, _c_int, C.c_char,
     _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, C.c_uint, _DP, _c_i64, C.POINTER(_c_i64), C.POINTER(AmenStats))
_sig("ttf_tt_amen_solve", _c_int, _c_int, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _IP, C.c_double, _c_int, _c_int,
     _c_int, C.c_char, _c_int, _c_int, _c_int, _DP, _c_i64, C.POINTER(_c_i64), C.POINTER(AmenStats))


class Mvk4Stats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("last_err_max", C.c_double), ("converged", C.c_int), ("max_rank", C.c_int),
                ("seconds", C.c_double), ("flops", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_sig("ttf_tt_mvk4", _c_int, _c_int, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _IP, C.c_double, _c_int, _c_int, _c_int,
     _c_int, _DP, _c_i64, C.POINTER(_c_i64), C.POINTER(Mvk4Stats))


class KslStats(C.Structure):
    _fields_ = [("nsolves", C.c_long), ("matvecs", C.c_long), ("krylov_steps", C.c_long), ("seconds", C.c_double),
                ("flops", C.c_double), ("matvecs32", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_I64P = C.POINTER(_c_i64)
_sig("ttf_tt_eigb_batch", _c_int, _c_int, _IP, _IP, _IP, _DP, _I64P, _DP, _I64P, _IP, _DP, _IP, _c_int, _c_int,
     _c_int, _DP, _DP, _I64P, _I64P, C.POINTER(EigbStats), _IP)
_sig("ttf_zexpm", _c_int, _c_int, C.c_double, _vp, _vp)
_sig("ttf_zexp_bfun3", _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, C.c_double, _c_int,
     C.c_double, _DP, _IP)
_sig("ttf_zqr", _c_int, _c_int, _c_int, _vp, _vp, _vp)
_sig("ttf_ztt_ksl", _c_int, _c_int, _IP, _IP, _IP, _vp, _vp, _IP, C.c_double, _c_int, _c_int, _c_int, _c_int,
     _c_int, _c_int, _c_int, _vp, _c_i64, C.POINTER(_c_i64), C.POINTER(KslStats))

_FP = C.POINTER(C.c_float)
_sig("ttf_mlp_param_count", C.c_long)
_sig("ttf_mlp_init", _c_int, C.c_uint)
_sig("ttf_mlp_set_params", _c_int, _vp)
_sig("ttf_mlp_get_params", _c_int, _vp)
_sig("ttf_mlp_forward", _c_int, _c_int, _vp, _vp)
_sig("ttf_mlp_loss_grad", _c_int, _c_int, _vp, _vp, _DP, _vp)
_sig("ttf_mlp_train", _c_int, C.c_long, _vp, _vp, _c_int, _c_int, C.c_float, C.c_uint, _vp, _c_int, _IP)
_sig("ttf_mlp_features", _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("ttf_mlp_precond", _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp)
_sig("ttf_mlp_collect_begin", _c_int, C.c_long, _c_int)
_sig("ttf_mlp_collect_end", _c_int, C.POINTER(C.c_long))
_sig("ttf_mlp_collected", _c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_long))


def _declare(lib):
    for name, res, args in _SIGS:
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


def exported_symbols():
    """Names declared in this binding (tests compare them with the header)."""
    return [s[0] for s in _SIGS]


def check(status: int, lib=None):
    if status != 0:
        lib = lib or load()
        msg = lib.ttf_last_error().decode(errors="replace")
        lib.ttf_clear_error()
        raise TTFError(msg or "ttf call failed with status %d" % status)


class DeviceArray:
    """A device buffer holding a column-major array."""

    def __init__(self, shape, dtype=np.float64, lib=None):
        self.lib = lib or load()
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        p = _vp()
        check(self.lib.ttf_malloc(C.byref(p), max(self.nbytes, 16)), self.lib)
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, dtype=None, lib=None):
        a = np.asarray(a, dtype=dtype) if dtype is not None else np.asarray(a)
        d = cls(a.shape if a.ndim else (1,), a.dtype, lib)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(np.asarray(a, dtype=self.dtype))
        if a.size != self.size:
            raise ValueError("size mismatch %d != %d" % (a.size, self.size))
        check(self.lib.ttf_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), self.lib)

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(self.lib.ttf_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), self.lib)
        return out

    def free(self):
        if self.ptr:
            check(self.lib.ttf_free(self.ptr), self.lib)
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                self.lib.ttf_free(self.ptr)
        except Exception:
            pass


def sync(lib=None):
    lib = lib or load()
    check(lib.ttf_sync(), lib)


class Timer:
    """Stream-event timer around work on the library stream."""

    def __init__(self, lib=None):
        self.lib = lib or load()

    def __enter__(self):
        check(self.lib.ttf_timer_start(), self.lib)
        return self

    def __exit__(self, *exc):
        ms = C.c_float()
        check(self.lib.ttf_timer_stop(C.byref(ms)), self.lib)
        self.ms = float(ms.value)
        return False


# ---------------------------------------------------------------- helpers
def _dev(a, dtype):
    return DeviceArray.from_numpy(np.ravel(np.asarray(a), order="F"), dtype=dtype)


def bfun3(shape, phi1, A, phi2, x, nvec=1, dtype=np.float64, transp=False):
    """Apply the local operator to nvec column-major vectors (host in/out).

    shape = (rx1, m, rx2, ry1, n, ry2, ra1, ra2).  x has nvec columns of
    rx1*m*rx2 entries (or ry1*n*ry2 when transp=True, complex only)."""
    lib = load()
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    nx, ny = rx1 * m * rx2, ry1 * n * ry2
    if transp:
        nx, ny = ny, nx
    dt = np.dtype(dtype)
    dphi1, dA, dphi2 = _dev(phi1, dt), _dev(A, dt), _dev(phi2, dt)
    dx = _dev(x, dt)
    dy = DeviceArray((ny * nvec,), dt)
    fn = {np.dtype(np.float64): lib.ttf_dbfun3, np.dtype(np.float32): lib.ttf_sbfun3,
          np.dtype(np.complex128): lib.ttf_zbfun3, np.dtype(np.complex64): lib.ttf_cbfun3}[dt]
    if transp:
        fn = lib.ttf_zbfun3_transp
    check(fn(*shape, nvec, dphi1.ptr, dA.ptr, dphi2.ptr, dx.ptr, nx, dy.ptr, ny), lib)
    return dy.numpy()


def phi_left(shape, phi1, A, x, y, complex_=False):
    lib = load()
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    dt = np.complex128 if complex_ else np.float64
    d = [_dev(a, dt) for a in (phi1, A, x, y)]
    out = DeviceArray((ry2 * rx2 * ra2,), dt)
    fn = lib.ttf_zphi_left if complex_ else lib.ttf_dphi_left
    check(fn(*shape, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, out.ptr), lib)
    return out.numpy()


def phi_right(shape, phi2, A, x, y, complex_=False):
    lib = load()
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    dt = np.complex128 if complex_ else np.float64
    d = [_dev(a, dt) for a in (phi2, A, x, y)]
    out = DeviceArray((rx1 * ra1 * ry1,), dt)
    fn = lib.ttf_zphi_right if complex_ else lib.ttf_dphi_right
    check(fn(*shape, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, out.ptr), lib)
    return out.numpy()


def phi2_left(rx1, rx2, ry1, n, ry2, phi1, x, y):
    lib = load()
    d = [_dev(a, np.float64) for a in (phi1, x, y)]
    out = DeviceArray((ry2 * rx2,), np.float64)
    check(lib.ttf_dphi2_left(rx1, rx2, ry1, n, ry2, d[0].ptr, d[1].ptr, d[2].ptr, out.ptr), lib)
    return out.numpy()


def phi2_right(rx1, rx2, ry1, n, ry2, phi2, x, y):
    lib = load()
    d = [_dev(a, np.float64) for a in (phi2, x, y)]
    out = DeviceArray((rx1 * ry1,), np.float64)
    check(lib.ttf_dphi2_right(rx1, rx2, ry1, n, ry2, d[0].ptr, d[1].ptr, d[2].ptr, out.ptr), lib)
    return out.numpy()


def fullmat(rx1, m, rx2, ra1, ra2, phi1, A, phi2):
    lib = load()
    d = [_dev(a, np.float64) for a in (phi1, A, phi2)]
    N = rx1 * m * rx2
    out = DeviceArray((N, N), np.float64)
    check(lib.ttf_d2d_fullmat(rx1, m, rx2, ra1, ra2, d[0].ptr, d[1].ptr, d[2].ptr, out.ptr), lib)
    return out.numpy()


def ref_dbfun3(shape, phi1, A, phi2, x, kind="i4"):
    """Call the reference-compatible `dbfun3` symbol with host arrays."""
    lib = load(kind)
    it = C.c_int32 if kind == "i4" else C.c_int64
    dims = [it(int(v)) for v in shape]
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    y = np.zeros(ry1 * n * ry2)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi1, A, phi2, x)]
    ws = np.zeros(1)
    lib.dbfun3(*[C.cast(C.byref(v), _vp) for v in dims], *[a.ctypes.data for a in arrs],
               y.ctypes.data, ws.ctypes.data, ws.ctypes.data)
    msg = lib.ttf_last_error().decode()
    if msg:
        lib.ttf_clear_error()
        raise TTFError(msg)
    return y


def _ip(a):
    return a.ctypes.data_as(_IP)


def _dp(a):
    return a.ctypes.data_as(_DP)


def gesvd(A):
    """Thin SVD of a host matrix through ttf_dgesvd: returns (U, s, VT, sweeps)."""
    lib = load()
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    k = min(m, n)
    dA = DeviceArray.from_numpy(A)
    dU, dS, dVT = DeviceArray((m, k)), DeviceArray((k,)), DeviceArray((k, n))
    sw = C.c_int()
    check(lib.ttf_dgesvd(m, n, dA.ptr, m, dU.ptr, m, dS.ptr, dVT.ptr, k, C.byref(sw)), lib)
    return dU.numpy(), dS.numpy(), dVT.numpy(), sw.value


def svd_side(A, side):
    """Singular values and one side's singular vectors through ttf_dsvd_side:
    returns (s, O, kpos, sweeps) with O (m x k) for side 'L', (n x k) for 'R'."""
    lib = load()
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    k = min(m, n)
    rows = m if side in ("L", "l") else n
    dA = DeviceArray.from_numpy(A)
    dS, dO = DeviceArray((k,)), DeviceArray((rows, k))
    kp, sw = C.c_int(), C.c_int()
    check(lib.ttf_dsvd_side(m, n, dA.ptr, m, side.encode(), dS.ptr, dO.ptr, rows, C.byref(kp), C.byref(sw)),
          lib)
    return dS.numpy(), dO.numpy(), kp.value, sw.value


def syev_small(As, desc=False):
    """Eigen-decomposition of a stack of symmetric matrices (count, n, n)."""
    lib = load()
    As = np.asarray(As, dtype=np.float64)
    cnt, n, _ = As.shape
    flat = np.concatenate([np.asfortranarray(a).reshape(-1, order="F") for a in As])
    dA = DeviceArray.from_numpy(flat)
    dw, dQ = DeviceArray((cnt * n,)), DeviceArray((cnt * n * n,))
    check(lib.ttf_dsyev_small(n, cnt, dA.ptr, dw.ptr, dQ.ptr, 1 if desc else 0), lib)
    w = dw.numpy().reshape(cnt, n)
    Q = dQ.numpy().reshape(cnt, n * n)
    return w, np.stack([q.reshape(n, n, order="F") for q in Q])


def gram(X, Y):
    lib = load()
    X = np.asfortranarray(X, dtype=np.float64)
    Y = np.asfortranarray(Y, dtype=np.float64)
    N, p = X.shape
    q = Y.shape[1]
    dX, dY, dG = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y), DeviceArray((p, q))
    check(lib.ttf_dgram(N, p, q, dX.ptr, N, dY.ptr, N, dG.ptr, p), lib)
    return dG.numpy()


def tt_eigb(n, ra, crA, crY0, ry, eps, rmax, B, nswp=20, max_full_size=2000, verb=0, kickrank=5):
    """Run ttf_tt_eigb with the reference's argument layout (host arrays).

    Returns (lambda, result cores flat, ry out, stats dict)."""
    lib = load()
    d = len(n)
    n_ = np.ascontiguousarray(n, dtype=np.int32)
    ra_ = np.ascontiguousarray(ra, dtype=np.int32)
    ry_ = np.ascontiguousarray(ry, dtype=np.int32).copy()
    crA_ = np.ascontiguousarray(crA, dtype=np.float64)
    crY_ = np.ascontiguousarray(crY0, dtype=np.float64)
    lam = np.zeros(B)
    cap = int(sum(max(rmax, max(ry)) ** 2 * int(nn) for nn in n) * B + 16)
    res = np.zeros(cap)
    size = C.c_int64()
    st = EigbStats()
    check(lib.ttf_tt_eigb(d, _ip(n_), _ip(n_), _ip(ra_), _dp(crA_), _dp(crY_), _ip(ry_), eps, rmax,
                          _dp(lam), B, kickrank, nswp, max_full_size, verb, _dp(res), cap,
                          C.byref(size), C.byref(st)), lib)
    return lam, res[:size.value].copy(), ry_, st.as_dict()


def my_chop3(s, eps):
    """Rank of the relative chop rule as ttf_tt_eigb applies it (host code)."""
    lib = load()
    s = np.ascontiguousarray(s, dtype=np.float64)
    return int(lib.ttf_my_chop3(s.size, _dp(s), eps))


def chop(s, tol=None, rmax=None):
    """The Frobenius chop (svd.f90:86-111) as the typed solver applies it."""
    lib = load()
    s = np.ascontiguousarray(s, dtype=np.float64)
    return int(lib.ttf_chop(s.size, _dp(s), -1.0 if tol is None else tol, 0 if rmax is None else rmax))


def dtt_amen_solve(n, ra, crA, ry, crY, tol, rx, crX0, kickrank=4, nswp=20, local_prec="n", local_iters=2,
                   local_restart=40, trunc_norm=1, max_full_size=50, verb=0, rmax=0, seed=0):
    """Run ttf_dtt_amen_solve (the typed dtt_amen_solve on flat arrays).
    Returns (result cores flat, rx out, stats dict)."""
    lib = load()
    d = len(n)
    n_ = np.ascontiguousarray(n, dtype=np.int32)
    ra_ = np.ascontiguousarray(ra, dtype=np.int32)
    ry_ = np.ascontiguousarray(ry, dtype=np.int32)
    rx_ = np.ascontiguousarray(rx, dtype=np.int32).copy()
    crA_ = np.ascontiguousarray(crA, dtype=np.float64)
    crY_ = np.ascontiguousarray(crY, dtype=np.float64)
    crX_ = np.ascontiguousarray(crX0, dtype=np.float64)
    R = (rmax if rmax > 0 else 256) + kickrank
    cap = int(sum(R * int(v) * R for v in n) + 16)
    res = np.zeros(cap)
    size = C.c_int64()
    st = AmenStats()
    check(lib.ttf_dtt_amen_solve(d, _ip(n_), _ip(ra_), _dp(crA_), _ip(ry_), _dp(crY_), tol, _ip(rx_), _dp(crX_),
                                 kickrank, nswp, local_prec.encode(), local_iters, local_restart, trunc_norm,
                                 max_full_size, verb, rmax, seed, _dp(res), cap, C.byref(size), C.byref(st)), lib)
    return res[:size.value].copy(), rx_, st.as_dict()


def residual_rank_rule(errs, res_new, eps2):
    """The AMEn residual-rank scan on given candidate residuals errs[c],
    c = 1..k-1 (errs[0] unused; k = len(errs))."""
    lib = load()
    e = np.ascontiguousarray(errs, dtype=np.float64)
    return int(lib.ttf_residual_rank_rule(e.size, _dp(e), res_new, eps2))


def amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, phi1, A, phi2, x, rhs, res_new, eps2):
    """The device residual-rank search of one AMEn local step.  Returns
    (rank, errs) with errs[c] the candidate residuals evaluated (-1 if not)."""
    lib = load()
    arrs = [DeviceArray.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for v in (phi1, A, phi2, x, rhs)]
    k = min(rx1 * n, rx2) if direction > 0 else min(rx1, n * rx2)
    errs = np.zeros(k)
    rank = C.c_int()
    check(lib.ttf_amen_truncation_rank(direction, rx1, n, rx2, ra1, ra2, *[a.ptr for a in arrs], res_new, eps2,
                                       C.byref(rank), _dp(errs)), lib)
    return rank.value, errs


def djac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, A, phi2):
    """Block-Jacobi blocks through ttf_djac_gen (host arrays in and out)."""
    lib = load()
    sz = int(lib.ttf_djac_size(ptype.encode(), rx1, n, rx2))
    d1, dA, d2 = (DeviceArray.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for v in (phi1, A, phi2))
    dj = DeviceArray((max(sz, 1),))
    check(lib.ttf_djac_gen(ptype.encode(), rx1, n, rx2, ra1, ra2, d1.ptr, dA.ptr, d2.ptr, dj.ptr), lib)
    return dj.numpy()[:sz]


def eig_block_prec(ry1, n, ry2, ra1, ra2, phi1, A, phi2, sigma):
    """Inverses of the shifted 'l' blocks through ttf_eig_block_prec: returns
    (inverses flattened as ttf_djac_gen 'l', positive_definite)."""
    lib = load()
    nb = ry1 * n
    d1, dA, d2 = (DeviceArray.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for v in (phi1, A, phi2))
    dj = DeviceArray((nb * nb * ry2,))
    pd = C.c_int(0)
    check(lib.ttf_eig_block_prec(ry1, n, ry2, ra1, ra2, d1.ptr, dA.ptr, d2.ptr, float(sigma), dj.ptr,
                                 C.byref(pd)), lib)
    return dj.numpy(), bool(pd.value)


def djac_apply(ptype, rx1, n, rx2, jacs, x):
    lib = load()
    dj = DeviceArray.from_numpy(np.ascontiguousarray(jacs, dtype=np.float64))
    dx = DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=np.float64))
    dy = DeviceArray((len(x),))
    check(lib.ttf_djac_apply(ptype.encode(), rx1, n, rx2, dj.ptr, dx.ptr, dy.ptr), lib)
    return dy.numpy()


def gmres(phi1, A, phi2, rhs, rx1, n, rx2, ra1, ra2, nrestart, tol, niters, ptype="n", jacs=None):
    """Local GMRES through ttf_dgmresr: returns (sol, (cycle, j), relres)."""
    lib = load()
    arrs = [DeviceArray.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for v in (phi1, A, phi2, rhs)]
    dj = DeviceArray.from_numpy(np.ascontiguousarray(jacs, dtype=np.float64)) if jacs is not None else None
    ds = DeviceArray((len(rhs),))
    it = np.zeros(2, dtype=np.int32)
    rr = np.zeros(1)
    check(lib.ttf_dgmresr(arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, arrs[3].ptr, rx1, n, rx2, ra1, ra2, nrestart, tol,
                          niters, ptype.encode(), dj.ptr if dj else None, ds.ptr, _ip(it), _dp(rr)), lib)
    return ds.numpy(), tuple(it), float(rr[0])


def duchol(trans, m, n, dA, numvec):
    """Kick directions through ttf_duchol: returns (du (n x real), lam, real)."""
    lib = load()
    dd = DeviceArray.from_numpy(np.ascontiguousarray(dA, dtype=np.float64))
    du = DeviceArray((n * numvec,))
    lam = np.zeros(numvec)
    real = np.zeros(1, dtype=np.int32)
    check(lib.ttf_duchol(trans.encode(), m, n, dd.ptr, numvec, du.ptr, _dp(lam), _ip(real)), lib)
    r = int(real[0])
    return du.numpy()[: n * r].reshape((n, r), order="F"), lam, r


def bfun3_half(side, rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi, A, x):
    """ttf_dbfun3_left (side 'L': phi = Phi1) or ttf_dbfun3_right ('R': phi = Phi2)."""
    lib = load()
    dp, dA, dx = (DeviceArray.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for v in (phi, A, x))
    if side == "L":
        dy = DeviceArray((ry1 * n * rx2 * ra2,))
        check(lib.ttf_dbfun3_left(rx1, m, rx2, ry1, n, ry2, ra1, ra2, dp.ptr, dA.ptr, dx.ptr, dy.ptr), lib)
    else:
        dy = DeviceArray((rx1 * ra1 * n * ry2,))
        check(lib.ttf_dbfun3_right(rx1, m, rx2, ry1, n, ry2, ra1, ra2, dA.ptr, dp.ptr, dx.ptr, dy.ptr), lib)
    return dy.numpy()


def tt_amen_solve(n, ra, crA, ry, crY, rx, crX0, eps, kickrank=4, nswp=20, verb=0, prec="n", nrestart=40,
                  niters=2, rmax=0):
    """Run ttf_tt_amen_solve with the reference's argument layout (host arrays).

    Returns (result cores flat, rx out, stats dict)."""
    lib = load()
    d = len(n)
    n_ = np.ascontiguousarray(n, dtype=np.int32)
    ra_ = np.ascontiguousarray(ra, dtype=np.int32)
    ry_ = np.ascontiguousarray(ry, dtype=np.int32)
    rx_ = np.ascontiguousarray(rx, dtype=np.int32).copy()
    crA_ = np.ascontiguousarray(crA, dtype=np.float64)
    crY_ = np.ascontiguousarray(crY, dtype=np.float64)
    crX_ = np.ascontiguousarray(crX0, dtype=np.float64)
    R = rmax if rmax > 0 else 256
    cap = int(sum(R * int(nn) * R for nn in n) + 16)
    res = np.zeros(cap)
    size = C.c_int64()
    st = AmenStats()
    check(lib.ttf_tt_amen_solve(d, _ip(n_), _ip(n_), _ip(ry_), _ip(ra_), _dp(crA_), _dp(crY_), _dp(crX_), _ip(rx_),
                                eps, kickrank, nswp, verb, prec.encode(), nrestart, niters, rmax, _dp(res), cap,
                                C.byref(size), C.byref(st)), lib)
    return res[:size.value].copy(), rx_, st.as_dict()


def tt_mvk4(n, m, rx, ra, crA, crX, ry, crY0, eps, rmax=0, kickrank=5, nswp=20, verb=0):
    """Run ttf_tt_mvk4 with the reference's argument layout (host arrays).

    Returns (result cores flat, ry out, stats dict)."""
    lib = load()
    d = len(n)
    n_ = np.ascontiguousarray(n, dtype=np.int32)
    m_ = np.ascontiguousarray(m, dtype=np.int32)
    rx_ = np.ascontiguousarray(rx, dtype=np.int32)
    ra_ = np.ascontiguousarray(ra, dtype=np.int32)
    ry_ = np.ascontiguousarray(ry, dtype=np.int32).copy()
    crA_ = np.ascontiguousarray(crA, dtype=np.float64)
    crX_ = np.ascontiguousarray(crX, dtype=np.float64)
    crY_ = np.ascontiguousarray(crY0, dtype=np.float64)
    R = max(rmax, int(max(ry)))
    cap = int(sum(R * int(v) * (R + kickrank) for v in n) + 16)
    res = np.zeros(cap)
    size = C.c_int64()
    st = Mvk4Stats()
    check(lib.ttf_tt_mvk4(d, _ip(n_), _ip(m_), _ip(rx_), _ip(ra_), _dp(crA_), _dp(crX_), _dp(crY_), _ip(ry_), eps,
                          rmax, kickrank, nswp, verb, _dp(res), cap, C.byref(size), C.byref(st)), lib)
    return res[:size.value].copy(), ry_, st.as_dict()


def zexpm(H, t):
    """exp(t H) of a small complex matrix through ttf_zexpm."""
    lib = load()
    H = np.asfortranarray(H, dtype=np.complex128)
    m = H.shape[0]
    dH = DeviceArray.from_numpy(H)
    dE = DeviceArray((m, m), np.complex128)
    check(lib.ttf_zexpm(m, t, dH.ptr, dE.ptr), lib)
    return dE.numpy()


def zexp_bfun3(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, tau, order=8, tol=1e-8):
    """exp(tau B) v for the zbfun3 local operator: returns (w, anorm, nmult)."""
    lib = load()
    arrs = [DeviceArray.from_numpy(np.asarray(x, dtype=np.complex128)) for x in (phi1, A, phi2, v)]
    dw = DeviceArray((len(v),), np.complex128)
    an = np.zeros(1)
    nm = np.zeros(1, dtype=np.int32)
    check(lib.ttf_zexp_bfun3(rx1, n, rx2, ra1, ra2, arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, arrs[3].ptr, dw.ptr, tau,
                             order, tol, _dp(an), _ip(nm)), lib)
    return dw.numpy(), float(an[0]), int(nm[0])


def zqr(A):
    """Complex QR through ttf_zqr: returns (Q, R)."""
    lib = load()
    A = np.asfortranarray(A, dtype=np.complex128)
    rows, cols = A.shape
    k = min(rows, cols)
    dA = DeviceArray.from_numpy(A)
    dQ, dR = DeviceArray((rows, k), np.complex128), DeviceArray((k, cols), np.complex128)
    check(lib.ttf_zqr(rows, cols, dA.ptr, dQ.ptr, dR.ptr), lib)
    return dQ.numpy(), dR.numpy()


def ztt_ksl(n, ra, crA, crY0, ry, tau, typ=2, order=8, usenrm=1, verb=0):
    """Run ttf_ztt_ksl with the reference's argument layout (host arrays).

    Returns (result cores flat (complex), ry out, stats dict)."""
    lib = load()
    d = len(n)
    n_ = np.ascontiguousarray(n, dtype=np.int32)
    ra_ = np.ascontiguousarray(ra, dtype=np.int32)
    ry_ = np.ascontiguousarray(ry, dtype=np.int32).copy()
    crA_ = np.ascontiguousarray(crA, dtype=np.complex128)
    crY_ = np.ascontiguousarray(crY0, dtype=np.complex128)
    cap = int(sum(max(ry_) ** 2 * int(nn) for nn in n) + 16)
    res = np.zeros(cap, dtype=np.complex128)
    size = C.c_int64()
    st = KslStats()
    check(lib.ttf_ztt_ksl(d, _ip(n_), _ip(n_), _ip(ra_), crA_.ctypes.data, crY_.ctypes.data, _ip(ry_), tau, 0, 0, 0,
                          verb, typ, order, usenrm, res.ctypes.data, cap, C.byref(size), C.byref(st)), lib)
    return res[:size.value].copy(), ry_, st.as_dict()


# ---------------------------------------------------- learned preconditioner
def mlp_param_count():
    return int(load().ttf_mlp_param_count())


def mlp_init(seed=0):
    lib = load()
    check(lib.ttf_mlp_init(seed), lib)


def mlp_set_params(theta):
    lib = load()
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    if theta.size != mlp_param_count():
        raise ValueError("parameter vector has %d entries, expected %d" % (theta.size, mlp_param_count()))
    check(lib.ttf_mlp_set_params(theta.ctypes.data), lib)


def mlp_get_params():
    lib = load()
    out = np.empty(mlp_param_count(), dtype=np.float32)
    check(lib.ttf_mlp_get_params(out.ctypes.data), lib)
    return out


def mlp_forward(X):
    """Network output (64 x M) for host features X (256 x M)."""
    lib = load()
    X = np.asfortranarray(X, dtype=np.float32)
    M = X.shape[1]
    dX = DeviceArray.from_numpy(X)
    dY = DeviceArray((64, M), np.float32)
    check(lib.ttf_mlp_forward(M, dX.ptr, dY.ptr), lib)
    return dY.numpy()


def mlp_loss_grad(X, T):
    lib = load()
    X = np.asfortranarray(X, dtype=np.float32)
    T = np.asfortranarray(T, dtype=np.float32)
    dX, dT = DeviceArray.from_numpy(X), DeviceArray.from_numpy(T)
    loss = C.c_double()
    g = np.empty(mlp_param_count(), dtype=np.float32)
    check(lib.ttf_mlp_loss_grad(X.shape[1], dX.ptr, dT.ptr, C.byref(loss), g.ctypes.data), lib)
    return loss.value, g


def mlp_train_device(X_ptr, T_ptr, nsamples, epochs, batch=1024, lr=1e-3, seed=0):
    """Train on device data (256 x n, 64 x n floats); returns the per-epoch
    mean batch loss."""
    lib = load()
    hist = np.zeros(max(epochs, 1))
    nh = C.c_int()
    check(lib.ttf_mlp_train(nsamples, X_ptr, T_ptr, epochs, batch, lr, seed, hist.ctypes.data, hist.size,
                            C.byref(nh)), lib)
    return hist[:nh.value].copy()


def mlp_train(X, T, epochs, batch=1024, lr=1e-3, seed=0):
    X = np.asfortranarray(X, dtype=np.float32)
    T = np.asfortranarray(T, dtype=np.float32)
    dX, dT = DeviceArray.from_numpy(X), DeviceArray.from_numpy(T)
    return mlp_train_device(dX.ptr, dT.ptr, X.shape[1], epochs, batch, lr, seed)


def mlp_features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v):
    lib = load()
    d = [_dev(a, np.float64) for a in (phi1, A, phi2, v)]
    X = DeviceArray((256, rx1 * rx2), np.float32)
    sc = DeviceArray((rx1 * rx2,), np.float32)
    check(lib.ttf_mlp_features(rx1, n, rx2, ra1, ra2, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, X.ptr, sc.ptr), lib)
    return X.numpy(), sc.numpy()


def mlp_precond(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v):
    lib = load()
    d = [_dev(a, np.float64) for a in (phi1, A, phi2, v)]
    z = DeviceArray((rx1 * n * rx2,), np.float64)
    check(lib.ttf_mlp_precond(rx1, n, rx2, ra1, ra2, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, z.ptr), lib)
    return z.numpy()


def mlp_collect_begin(capacity, per_apply):
    lib = load()
    check(lib.ttf_mlp_collect_begin(capacity, per_apply), lib)


def mlp_collect_end():
    lib = load()
    c = C.c_long()
    check(lib.ttf_mlp_collect_end(C.byref(c)), lib)
    return c.value


def mlp_collected(host=False):
    """(X_ptr, T_ptr, count) of the collected pairs, or host copies."""
    lib = load()
    xp, tp, c = _vp(), _vp(), C.c_long()
    check(lib.ttf_mlp_collected(C.byref(xp), C.byref(tp), C.byref(c)), lib)
    if not host:
        return xp.value, tp.value, c.value
    X = np.empty((256, c.value), dtype=np.float32, order="F")
    T = np.empty((64, c.value), dtype=np.float32, order="F")
    if c.value:
        check(lib.ttf_memcpy_d2h(X.ctypes.data, xp.value, X.nbytes), lib)
        check(lib.ttf_memcpy_d2h(T.ctypes.data, tp.value, T.nbytes), lib)
    return X, T, c.value


def set_local_precision(bits):
    """Precision of the matvecs inside the local iterative solvers (64 or 32)."""
    lib = load()
    check(lib.ttf_set_local_precision(bits), lib)


def tt_eigb_batch(problems, B, nswp=20, nworkers=8):
    """Solve many tt_eigb problems concurrently.  problems: list of dicts with
    n, ra, crA, crY0, ry, eps, rmax.  Returns a list of (lambda, cores, ry,
    stats) in the same order."""
    lib = load()
    nc = len(problems)
    d = np.array([len(p["n"]) for p in problems], dtype=np.int32)
    n_all = np.concatenate([np.asarray(p["n"], dtype=np.int32) for p in problems])
    ra_all = np.concatenate([np.asarray(p["ra"], dtype=np.int32) for p in problems])
    ry_all = np.concatenate([np.asarray(p["ry"], dtype=np.int32) for p in problems]).copy()
    crA = [np.asarray(p["crA"], dtype=np.float64) for p in problems]
    crY = [np.asarray(p["crY0"], dtype=np.float64) for p in problems]
    crA_off = np.concatenate([[0], np.cumsum([a.size for a in crA])]).astype(np.int64)
    crY_off = np.concatenate([[0], np.cumsum([a.size for a in crY])]).astype(np.int64)
    caps = [int(sum(max(p["rmax"], max(p["ry"])) ** 2 * int(nn) for nn in p["n"]) * B + 16) for p in problems]
    res_off = np.concatenate([[0], np.cumsum(caps)]).astype(np.int64)
    eps = np.array([p["eps"] for p in problems], dtype=np.float64)
    rmax = np.array([p["rmax"] for p in problems], dtype=np.int32)
    lam = np.zeros(nc * B)
    res = np.zeros(int(res_off[-1]))
    sizes = np.zeros(nc, dtype=np.int64)
    stats = (EigbStats * nc)()
    status = np.zeros(nc, dtype=np.int32)
    crA_all = np.concatenate(crA)
    crY_all = np.concatenate(crY)
    check(lib.ttf_tt_eigb_batch(nc, _ip(d), _ip(n_all), _ip(ra_all), _dp(crA_all), crA_off.ctypes.data_as(_I64P),
                                _dp(crY_all), crY_off.ctypes.data_as(_I64P), _ip(ry_all), _dp(eps), _ip(rmax), B,
                                nswp, nworkers, _dp(lam), _dp(res), res_off.ctypes.data_as(_I64P),
                                sizes.ctypes.data_as(_I64P), stats, _ip(status)), lib)
    out, orr = [], 0
    for c in range(nc):
        ry = ry_all[orr:orr + d[c] + 1].copy()
        orr += d[c] + 1
        out.append((lam[c * B:(c + 1) * B].copy(), res[res_off[c]:res_off[c] + sizes[c]].copy(), ry,
                    stats[c].as_dict() if hasattr(stats[c], "as_dict") else None))
    return out


# ------------------------------------------------------------------ SDV files
_sig("ttf_sdv_dims", _c_int, C.c_char_p, _c_int, _IP, _IP, _IP, _IP, _I64P, _IP)
_sig("ttf_dtt_read", _c_int, C.c_char_p, _c_i64, _vp)
_sig("ttf_ztt_read", _c_int, C.c_char_p, _c_i64, _vp)
_sig("ttf_dtt_write", _c_int, C.c_char_p, _c_int, _c_int, _IP, _IP, _vp)
_sig("ttf_ztt_write", _c_int, C.c_char_p, _c_int, _c_int, _IP, _IP, _vp)


class SDVError(TTFError):
    """A failed SDV read or write; `status` is the library's code (the
    reference readers' info: -1 no such file, -2 not a TT file, -3 other
    version, > 0 errno; -4 truncated, -5 invalid sizes or buffer, -6 real /
    complex mismatch)."""

    def __init__(self, msg, status):
        super().__init__(msg)
        self.status = status


def _sdv_check(status, lib):
    if status != 0:
        msg = lib.ttf_last_error().decode(errors="replace")
        lib.ttf_clear_error()
        raise SDVError(msg or "SDV call failed with status %d" % status, status)


def _sdv_flat(path):
    """(l, m, n, r, x): sizes and the cores back to back, as the file holds
    them."""
    lib = load()
    fn = os.fsencode(path)
    l, m, cplx = C.c_int(), C.c_int(), C.c_int()
    count = _c_i64()
    _sdv_check(lib.ttf_sdv_dims(fn, 0, C.byref(l), C.byref(m), None, None, C.byref(count), C.byref(cplx)), lib)
    d = m.value - l.value + 1
    n = np.zeros(d, dtype=np.int32)
    r = np.zeros(d + 1, dtype=np.int32)
    _sdv_check(lib.ttf_sdv_dims(fn, d, None, None, _ip(n), _ip(r), C.byref(count), None), lib)
    x = np.zeros(max(count.value, 1), dtype=np.complex128 if cplx.value else np.float64)
    rd = lib.ttf_ztt_read if cplx.value else lib.ttf_dtt_read
    _sdv_check(rd(fn, count.value, x.ctypes.data), lib)
    return l.value, m.value, n, r, x[:count.value]


def sdv_read(path):
    """Read a tensor train from an SDV file (ttio.f90:196-399, dtt_read and
    ztt_read, chosen by the file's header).  Returns dict(l, m, n, r, cores):
    n has m - l + 1 entries, r has m - l + 2, core k is a column-major array
    of shape (r[k], n[k], r[k + 1]), float64 or complex128.  Host code only:
    no device is needed."""
    l, m, n, r, x = _sdv_flat(path)
    cores, p = [], 0
    for k in range(len(n)):
        sz = int(r[k]) * int(n[k]) * int(r[k + 1])
        cores.append(x[p:p + sz].reshape((int(r[k]), int(n[k]), int(r[k + 1])), order="F"))
        p += sz
    return dict(l=l, m=m, n=n, r=r, cores=cores)


def sdv_to_arrays(path):
    """The tensor train of an SDV file as the flat arrays of the ttpy
    converters (dtt_read, then python_conv.f90:12-56, dsdv_to_arrays /
    zsdv_to_arrays): returns (n, r, d, ps, cr) with ps the 1-based core
    offsets, ps[0] = 1 and ps[k + 1] = ps[k] + r[k] n[k] r[k + 1], and cr
    the cores back to back.  n is the file's n(l:m); the reference copies
    n(1:d) instead, which differs when l != 1."""
    _, _, n, r, x = _sdv_flat(path)
    d = len(n)
    ps = np.ones(d + 1, dtype=np.int32)
    ps[1:] = 1 + np.cumsum(r[:-1].astype(np.int64) * n * r[1:])
    return n, r, d, ps, x


def arrays_to_sdv(n, r, d, ps, cr, path):
    """Write flat arrays (n, r, d, ps, cr) as python_conv.f90:59-97
    (darrays_to_sdv / zarrays_to_sdv, cores l = 1 .. d) followed by
    dtt_write / ztt_write do; cr complex selects the complex format."""
    n = np.asarray(n, dtype=np.int32)[:d]
    r = np.asarray(r, dtype=np.int32)[:d + 1]
    ps = np.asarray(ps, dtype=np.int64)[:d + 1]
    cr = np.asarray(cr)
    sizes = r[:-1].astype(np.int64) * n * r[1:]
    if np.any(ps[:-1] < 1) or np.any(ps[:-1] - 1 + sizes > cr.size):
        raise SDVError("arrays_to_sdv: core offsets ps run outside cr", -5)
    cores = [cr[ps[k] - 1:ps[k] - 1 + sizes[k]].reshape((int(r[k]), int(n[k]), int(r[k + 1])), order="F")
             for k in range(d)]
    sdv_write(path, cores, l=1)


def sdv_write(path, cores, l=1):
    """Write cores (arrays of shape (r[k], n[k], r[k + 1]), consecutive
    ranks matching) to an SDV file as ttio.f90:29-192 (dtt_write, or
    ztt_write when any core is complex) does.  Replaces an existing file."""
    lib = load()
    if not cores:
        raise SDVError("sdv_write: no cores", -5)
    cplx = any(np.iscomplexobj(c) for c in cores)
    dt = np.complex128 if cplx else np.float64
    shapes = [np.shape(c) for c in cores]
    for k, s in enumerate(shapes):
        if len(s) != 3 or (k > 0 and s[0] != shapes[k - 1][2]):
            raise SDVError("sdv_write: core %d has shape %r; cores must be (r[k], n[k], r[k+1]) with "
                           "matching ranks" % (k, s), -5)
    n = np.array([s[1] for s in shapes], dtype=np.int32)
    r = np.array([shapes[0][0]] + [s[2] for s in shapes], dtype=np.int32)
    x = np.concatenate([np.asarray(c, dtype=dt).reshape(-1, order="F") for c in cores])
    wr = lib.ttf_ztt_write if cplx else lib.ttf_dtt_write
    _sdv_check(wr(os.fsencode(path), l, l + len(cores) - 1, _ip(n), _ip(r), x.ctypes.data), lib)

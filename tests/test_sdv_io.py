"""SDV reader and writer of the library (ttf_sdv_dims, ttf_dtt_read,
ttf_ztt_read, ttf_dtt_write, ttf_ztt_write; reference ttio.f90:10-399).

These entry points are host code, so the tests run on the CPU through the
C-ABI.  Pins:
  * the reference's own fixtures Heisen_20.sdv (real; its MPO is pinned to
    the Heisenberg formula in test_oracle_kernels.py) and cme20_A.sdv
    (complex, from test_ksl_cme.f90) read to the same bits as the oracle's
    independent reader, and rewrite to the same bytes after the header;
  * the header the writers produce, built field by field from ttio.f90:10-17
    and the writers' assignments (ttio.f90:73-74, 154-156);
  * round trips over ragged ranks, a single core, l != 1 and complex cores;
  * every status code, including the reference readers' -1 / -2 / -3.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from oracle import operators as O
import ttfort


def _payload(path):
    return open(path, "rb").read()[128:]


def _expected_header(l, m, complex_):
    # txt, ver(2), inf(4) = (tt_size, complex flag, 0, 0), comment(64), i(8)
    return (b"TT      " + struct.pack("<2i", 1, 0) + struct.pack("<4i", 1023, 1 if complex_ else 0, 0, 0)
            + b"\0" * 64 + struct.pack("<8i", l, m, 0, 0, 0, 0, 0, 0))


@pytest.mark.parametrize("name,complex_", [("Heisen_20.sdv", False), ("cme20_A.sdv", True)])
def test_reference_fixture_matches_oracle_reader(golden_dir, name, complex_):
    path = os.path.join(golden_dir, name)
    got = ttfort.sdv_read(path)
    ref = O.read_sdv(path)
    assert (got["l"], got["m"]) == (ref["l"], ref["m"]) == (1, len(ref["cores"]))
    assert np.array_equal(got["n"], ref["n"]) and np.array_equal(got["r"], ref["r"])
    assert (ref["dtype"] == np.complex128) == complex_
    for k, (a, b) in enumerate(zip(got["cores"], ref["cores"])):
        assert a.dtype == ref["dtype"]
        assert a.shape == (ref["r"][k], ref["n"][k], ref["r"][k + 1])
        assert np.array_equal(a.reshape(-1, order="F"), b), "core %d" % k


@pytest.mark.parametrize("name", ["Heisen_20.sdv", "cme20_A.sdv"])
def test_reference_fixture_rewrites_to_same_bytes(golden_dir, tmp_path, name):
    src = os.path.join(golden_dir, name)
    t = ttfort.sdv_read(src)
    out = str(tmp_path / name)
    ttfort.sdv_write(out, t["cores"], l=t["l"])
    assert _payload(out) == _payload(src)
    complex_ = np.iscomplexobj(t["cores"][0])
    assert open(out, "rb").read()[:128] == _expected_header(t["l"], t["m"], complex_)


def _random_tt(rng, n, r, complex_):
    cores = []
    for k in range(len(n)):
        c = rng.standard_normal((r[k], n[k], r[k + 1]))
        if complex_:
            c = c + 1j * rng.standard_normal(c.shape)
        cores.append(np.asfortranarray(c))
    return cores


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("n,r,l", [
    ([2, 3, 5, 2], [1, 2, 7, 3, 1], 1),
    ([7], [1, 1], 1),                      # a single core
    ([4, 4, 4], [3, 1, 6, 2], 3),          # boundary ranks above 1, l != 1
    ([1, 9, 1], [1, 1, 1, 1], 1),          # unit modes and ranks
])
def test_round_trip(tmp_path, n, r, l, complex_):
    rng = np.random.default_rng(len(n) * 10 + l)
    cores = _random_tt(rng, n, r, complex_)
    path = str(tmp_path / "t.sdv")
    ttfort.sdv_write(path, cores, l=l)
    got = ttfort.sdv_read(path)
    assert got["l"] == l and got["m"] == l + len(n) - 1
    assert list(got["n"]) == n and list(got["r"]) == r
    for a, b in zip(got["cores"], cores):
        assert a.dtype == (np.complex128 if complex_ else np.float64)
        assert np.array_equal(a, b)
    # the oracle's reader, which tells real from complex by the payload size,
    # agrees with the header flag the writer set
    ref = O.read_sdv(path)
    assert ref["dtype"] == (np.complex128 if complex_ else np.float64)
    for a, b in zip(got["cores"], ref["cores"]):
        assert np.array_equal(a.reshape(-1, order="F"), b)
    assert open(path, "rb").read()[:128] == _expected_header(l, l + len(n) - 1, complex_)
    size = 128 + 8 + 4 * (2 * len(n) + 1) + sum(r[k] * n[k] * r[k + 1] for k in range(len(n))) * (
        16 if complex_ else 8)
    assert os.path.getsize(path) == size


def test_dims_query(golden_dir):
    lib = ttfort.load()
    path = os.fsencode(os.path.join(golden_dir, "Heisen_20.sdv"))
    l, m, cplx = C.c_int(), C.c_int(), C.c_int(-1)
    count = C.c_int64()
    assert lib.ttf_sdv_dims(path, 0, C.byref(l), C.byref(m), None, None, C.byref(count), C.byref(cplx)) == 0
    assert (l.value, m.value, cplx.value) == (1, 20, 0)
    # ranks 1, 4, 5, ..., 5, 4, 1 with 4 = 2 x 2 entries per mode: two cores
    # 1 x 4 x 4, two 4 x 4 x 5 and sixteen 5 x 4 x 5
    assert count.value == 2 * 16 + 2 * 80 + 16 * 100
    n = np.zeros(20, dtype=np.int32)
    r = np.zeros(21, dtype=np.int32)
    assert lib.ttf_sdv_dims(path, 19, None, None, ttfort._ip(n), ttfort._ip(r), None, None) == -5
    lib.ttf_clear_error()
    assert lib.ttf_sdv_dims(path, 20, None, None, ttfort._ip(n), ttfort._ip(r), None, None) == 0
    assert list(n) == [4] * 20 and list(r) == [1, 4] + [5] * 17 + [4, 1]


def _status(fn):
    with pytest.raises(ttfort.SDVError) as e:
        fn()
    return e.value.status


def _write_bytes(path, b):
    with open(path, "wb") as f:
        f.write(b)
    return path


def test_status_codes(tmp_path, golden_dir):
    real = open(os.path.join(golden_dir, "Heisen_20.sdv"), "rb").read()
    # the reference readers' info codes
    assert _status(lambda: ttfort.sdv_read(str(tmp_path / "absent.sdv"))) == -1
    assert _status(lambda: ttfort.sdv_read(_write_bytes(str(tmp_path / "x1"), b"XX" + real[2:]))) == -2
    bad_ver = real[:8] + struct.pack("<i", 2) + real[12:]
    assert _status(lambda: ttfort.sdv_read(_write_bytes(str(tmp_path / "x2"), bad_ver))) == -3
    # truncated in the header, in the sizes and in the cores
    for cut in (100, 140, len(real) - 8):
        p = _write_bytes(str(tmp_path / ("cut%d" % cut)), real[:cut])
        assert _status(lambda: ttfort.sdv_read(p)) == -4, cut
    # a negative rank, an empty core range
    neg = bytearray(real)
    off_r = 128 + 8 + 4 * 20 + 4 * 3
    neg[off_r:off_r + 4] = struct.pack("<i", -5)
    assert _status(lambda: ttfort.sdv_read(_write_bytes(str(tmp_path / "x3"), bytes(neg)))) == -5
    empty = real[:128] + struct.pack("<2i", 5, 4) + real[136:]
    assert _status(lambda: ttfort.sdv_read(_write_bytes(str(tmp_path / "x4"), empty))) == -5
    # a corrupt rank whose cores cannot be in the file is reported by the
    # size query itself, before a caller allocates for it
    huge = bytearray(real)
    huge[off_r:off_r + 4] = struct.pack("<i", 0x7fffffff)
    p = _write_bytes(str(tmp_path / "x6"), bytes(huge))
    lib = ttfort.load()
    cnt = C.c_int64(-7)
    assert lib.ttf_sdv_dims(os.fsencode(p), 0, None, None, None, None, C.byref(cnt), None) == -4
    lib.ttf_clear_error()
    assert _status(lambda: ttfort.sdv_read(p)) == -4
    # a directory as the file to read: errno (EISDIR)
    assert _status(lambda: ttfort.sdv_read(str(tmp_path))) > 0
    # a directory as the file to write: errno (EISDIR)
    t = ttfort.sdv_read(os.path.join(golden_dir, "Heisen_20.sdv"))
    assert _status(lambda: ttfort.sdv_write(str(tmp_path), t["cores"])) > 0
    # mismatched consecutive ranks are refused before anything is written
    bad = [np.zeros((1, 2, 3)), np.zeros((4, 2, 1))]
    assert _status(lambda: ttfort.sdv_write(str(tmp_path / "x5"), bad)) == -5
    assert not os.path.exists(str(tmp_path / "x5"))


def test_kind_and_capacity_checks(tmp_path, golden_dir):
    lib = ttfort.load()
    real = os.fsencode(os.path.join(golden_dir, "Heisen_20.sdv"))
    cplx = os.fsencode(os.path.join(golden_dir, "cme20_A.sdv"))
    buf = np.zeros(1 << 16, dtype=np.complex128)
    assert lib.ttf_ztt_read(real, buf.size, buf.ctypes.data) == -6
    lib.ttf_clear_error()
    assert lib.ttf_dtt_read(cplx, 2 * buf.size, buf.ctypes.data) == -6
    lib.ttf_clear_error()
    small = np.zeros(10)
    assert lib.ttf_dtt_read(real, small.size, small.ctypes.data) == -5
    assert "room for 10" in lib.ttf_last_error().decode()
    lib.ttf_clear_error()
    # the error path leaves the caller's buffer untouched
    assert not small.any()
    # writer argument checks: null sizes, negative mode size
    n = np.array([2, 2], dtype=np.int32)
    r = np.array([1, 2, 1], dtype=np.int32)
    x = np.zeros(8)
    p = os.fsencode(str(tmp_path / "w.sdv"))
    assert lib.ttf_dtt_write(p, 1, 2, None, ttfort._ip(r), x.ctypes.data) == -5
    n_bad = np.array([2, -2], dtype=np.int32)
    assert lib.ttf_dtt_write(p, 1, 2, ttfort._ip(n_bad), ttfort._ip(r), x.ctypes.data) == -5
    assert lib.ttf_dtt_write(p, 1, 0, ttfort._ip(n), ttfort._ip(r), x.ctypes.data) == -5
    lib.ttf_clear_error()
    assert lib.ttf_dtt_write(p, 1, 2, ttfort._ip(n), ttfort._ip(r), x.ctypes.data) == 0


@pytest.mark.parametrize("name", ["Heisen_20.sdv", "cme20_A.sdv"])
def test_flat_arrays_round_trip(golden_dir, tmp_path, name):
    """sdv_to_arrays / arrays_to_sdv (python_conv.f90:12-97): 1-based core
    offsets, the cores back to back, and a file rewritten from the arrays
    identical to the reference's after the header."""
    src = os.path.join(golden_dir, name)
    n, r, d, ps, cr = ttfort.sdv_to_arrays(src)
    ref = O.read_sdv(src)
    assert d == len(ref["cores"]) and ps[0] == 1
    assert all(ps[k + 1] - ps[k] == r[k] * n[k] * r[k + 1] for k in range(d))
    assert ps[d] - 1 == cr.size
    assert np.array_equal(cr, np.concatenate(ref["cores"]))
    out = str(tmp_path / name)
    ttfort.arrays_to_sdv(n, r, d, ps, cr, out)
    assert _payload(out) == _payload(src)


def test_flat_arrays_with_gaps_and_bad_offsets(tmp_path):
    rng = np.random.default_rng(4)
    n, r = [2, 3, 2], [1, 2, 2, 1]
    cores = _random_tt(rng, n, r, False)
    sizes = [c.size for c in cores]
    # cores placed with gaps between them in a larger array
    ps = [1, 1 + sizes[0] + 5, 1 + sizes[0] + 5 + sizes[1] + 3, 0]
    ps[3] = ps[2] + sizes[2]
    cr = np.full(ps[3] - 1 + 4, np.nan)
    for k in range(3):
        cr[ps[k] - 1:ps[k] - 1 + sizes[k]] = cores[k].reshape(-1, order="F")
    path = str(tmp_path / "g.sdv")
    ttfort.arrays_to_sdv(n, r, 3, ps, cr, path)
    got = ttfort.sdv_read(path)
    for a, b in zip(got["cores"], cores):
        assert np.array_equal(a, b)
    bad = list(ps)
    bad[2] = cr.size   # the last core would run past the end of cr
    assert _status(lambda: ttfort.arrays_to_sdv(n, r, 3, bad, cr, str(tmp_path / "b.sdv"))) == -5

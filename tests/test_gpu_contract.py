"""GPU parity of the contraction kernels against the oracle (through the C ABI).

Tolerance: FP64 kernels agree with the oracle to 1e-12 relative to the
problem scale (sum of |terms|); FP32 kernels to 2e-5 relative (single
precision with FP32 accumulation over K <= 1280).
"""
import numpy as np
import pytest

from oracle import kernels as K

pytestmark = pytest.mark.gpu

SHAPES = [
    # rx1, m, rx2, ry1, n, ry2, ra1, ra2
    (1, 2, 1, 1, 2, 1, 1, 1),
    (1, 2, 4, 1, 2, 4, 1, 5),
    (3, 2, 4, 5, 3, 2, 2, 3),
    (17, 2, 33, 17, 2, 33, 5, 5),
    (40, 2, 40, 40, 2, 40, 5, 5),
    (64, 3, 65, 63, 3, 65, 4, 6),
    (128, 2, 128, 128, 2, 128, 5, 5),
    # the fused kernel's edges: an odd b1 range split in two (the second half
    # one row), c2 tiles cut short, three-column c2 tiles
    (5, 2, 7, 3, 2, 9, 2, 3),
    (96, 2, 96, 96, 2, 96, 5, 5),
    (21, 3, 11, 22, 3, 13, 3, 2),
]


def _rand(rng, shape, nvec=1, cplx=False):
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    g = (lambda k: rng.standard_normal(k) + 1j * rng.standard_normal(k)) if cplx else rng.standard_normal
    return dict(phi1=g(ry1 * rx1 * ra1), A=g(ra1 * n * m * ra2), phi2=g(rx2 * ra2 * ry2),
                x=g(rx1 * m * rx2 * nvec), y=g(ry1 * n * ry2))


def _scale(shape, t):
    # bound on |y| entries: products of absolute values
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    return (np.abs(t["phi1"]).max() * np.abs(t["A"]).max() * np.abs(t["phi2"]).max()
            * np.abs(t["x"]).max() * rx1 * m * rx2 * ra1 * ra2)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("nvec", [1, 5, 8])
def test_dbfun3(ttlib, shape, nvec):
    import ttfort
    rng = np.random.default_rng(10)
    t = _rand(rng, shape, nvec)
    nx = shape[0] * shape[1] * shape[2]
    y = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], nvec=nvec)
    for v in range(nvec):
        ref = K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"][v * nx:(v + 1) * nx])
        got = y[v * ref.size:(v + 1) * ref.size]
        assert np.abs(got - ref).max() <= 1e-13 * _scale(shape, t), (shape, v)


@pytest.mark.parametrize("shape", SHAPES[2:])
def test_sbfun3_fp32(ttlib, shape):
    import ttfort
    rng = np.random.default_rng(11)
    t = _rand(rng, shape, 2)
    y = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], nvec=2, dtype=np.float32)
    nx = shape[0] * shape[1] * shape[2]
    for v in range(2):
        ref = K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"][v * nx:(v + 1) * nx])
        got = y[v * ref.size:(v + 1) * ref.size]
        assert np.linalg.norm(got - ref) <= 2e-5 * np.linalg.norm(ref) + 1e-6 * _scale(shape, t)


@pytest.mark.parametrize("shape", SHAPES[2:])
def test_cbfun3_complex64(ttlib, shape):
    """The complex64 chain (fp32 matrix cores, fp32 accumulation) against the
    oracle's complex128 zbfun3: 2e-5 relative, the FP32 bar."""
    import ttfort
    rng = np.random.default_rng(13)
    t = _rand(rng, shape, 2, cplx=True)
    nx = shape[0] * shape[1] * shape[2]
    ny = shape[3] * shape[4] * shape[5]
    y = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], nvec=2, dtype=np.complex64)
    for v in range(2):
        ref = K.zbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"][v * nx:(v + 1) * nx])
        got = y[v * ny:(v + 1) * ny]
        assert np.linalg.norm(got - ref) <= 2e-5 * np.linalg.norm(ref) + 4e-6 * _scale(shape, t)


@pytest.mark.parametrize("shape", SHAPES[:6])
def test_zbfun3_and_transp(ttlib, shape):
    import ttfort
    rng = np.random.default_rng(12)
    t = _rand(rng, shape, 2, cplx=True)
    nx = shape[0] * shape[1] * shape[2]
    ny = shape[3] * shape[4] * shape[5]
    y = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], nvec=2, dtype=np.complex128)
    for v in range(2):
        ref = K.zbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"][v * nx:(v + 1) * nx])
        assert np.abs(y[v * ny:(v + 1) * ny] - ref).max() <= 1e-13 * 4 * _scale(shape, t)
    xt = np.concatenate([t["y"], t["y"][::-1]])
    yt = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], xt, nvec=2, dtype=np.complex128,
                      transp=True)
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    sc_t = (np.abs(t["phi1"]).max() * np.abs(t["A"]).max() * np.abs(t["phi2"]).max()
            * np.abs(xt).max() * ry1 * n * ry2 * ra1 * ra2 * 4)
    for v in range(2):
        ref = K.zbfun3_transp(*shape, t["phi1"], t["A"], t["phi2"], xt[v * ny:(v + 1) * ny])
        assert np.abs(yt[v * nx:(v + 1) * nx] - ref).max() <= 1e-13 * sc_t


@pytest.mark.parametrize("shape", SHAPES)
def test_phi_left_right(ttlib, shape):
    import ttfort
    rng = np.random.default_rng(13)
    t = _rand(rng, shape)
    sc = _scale(shape, t) * shape[3] * shape[4] * shape[5] * np.abs(t["y"]).max()
    pl = ttfort.phi_left(shape, t["phi1"], t["A"], t["x"], t["y"])
    assert np.abs(pl - K.dphi_left(*shape, t["phi1"], t["A"], t["x"], t["y"])).max() <= 1e-13 * sc
    pr = ttfort.phi_right(shape, t["phi2"], t["A"], t["x"], t["y"])
    assert np.abs(pr - K.dphi_right(*shape, t["phi2"], t["A"], t["x"], t["y"])).max() <= 1e-13 * sc


@pytest.mark.parametrize("shape", SHAPES[:6])
def test_zphi_left_right(ttlib, shape):
    import ttfort
    rng = np.random.default_rng(14)
    t = _rand(rng, shape, cplx=True)
    sc = _scale(shape, t) * shape[3] * shape[4] * shape[5] * np.abs(t["y"]).max() * 4
    pl = ttfort.phi_left(shape, t["phi1"], t["A"], t["x"], t["y"], complex_=True)
    assert np.abs(pl - K.zphi_left(*shape, t["phi1"], t["A"], t["x"], t["y"])).max() <= 1e-13 * sc
    pr = ttfort.phi_right(shape, t["phi2"], t["A"], t["x"], t["y"], complex_=True)
    assert np.abs(pr - K.zphi_right(*shape, t["phi2"], t["A"], t["x"], t["y"])).max() <= 1e-13 * sc


def test_phi2(ttlib):
    import ttfort
    rng = np.random.default_rng(15)
    for rx1, rx2, ry1, n, ry2 in ((3, 4, 2, 3, 5), (40, 33, 41, 2, 30), (1, 1, 1, 2, 1)):
        p1 = rng.standard_normal(ry1 * rx1)
        p2 = rng.standard_normal(rx2 * ry2)
        x = rng.standard_normal(rx1 * n * rx2)
        y = rng.standard_normal(ry1 * n * ry2)
        l = ttfort.phi2_left(rx1, rx2, ry1, n, ry2, p1, x, y)
        assert np.allclose(l, K.dphi2_left(rx1, rx2, ry1, n, ry2, p1, x, y), rtol=1e-12, atol=1e-11)
        r = ttfort.phi2_right(rx1, rx2, ry1, n, ry2, p2, x, y)
        assert np.allclose(r, K.dphi2_right(rx1, rx2, ry1, n, ry2, p2, x, y), rtol=1e-12, atol=1e-11)


def test_fullmat(ttlib):
    import ttfort
    rng = np.random.default_rng(16)
    for r1, n, r2, ra1, ra2 in ((3, 2, 4, 3, 2), (10, 2, 9, 5, 5), (1, 2, 1, 1, 1)):
        p1 = rng.standard_normal(r1 * r1 * ra1)
        A = rng.standard_normal(ra1 * n * n * ra2)
        p2 = rng.standard_normal(r2 * ra2 * r2)
        B = ttfort.fullmat(r1, n, r2, ra1, ra2, p1, A, p2)
        assert np.allclose(B, K.d2d_fullmat(r1, n, r2, ra1, ra2, p1, A, p2), atol=1e-12)


def test_y_aliases_x(ttlib):
    import ttfort
    rng = np.random.default_rng(17)
    shape = (20, 2, 20, 20, 2, 20, 5, 5)
    t = _rand(rng, shape, 3)
    lib = ttfort.load()
    d = [ttfort.DeviceArray.from_numpy(t[k]) for k in ("phi1", "A", "phi2", "x")]
    N = 20 * 2 * 20
    ttfort.check(lib.ttf_dbfun3(*shape, 3, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, N, d[3].ptr, N))
    got = d[3].numpy()
    for v in range(3):
        ref = K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"][v * N:(v + 1) * N])
        assert np.allclose(got[v * N:(v + 1) * N], ref, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("kind", ["i4", "i8"])
def test_reference_dbfun3_symbol(ttlib, kind):
    import ttfort
    rng = np.random.default_rng(18)
    shape = (7, 2, 9, 8, 2, 6, 3, 4)
    t = _rand(rng, shape)
    y = ttfort.ref_dbfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], kind=kind)
    ref = K.dbfun3(*shape, t["phi1"], t["A"], t["phi2"], t["x"])
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-11)


def test_invalid_sizes_report_error(ttlib):
    import ttfort
    lib = ttfort.load()
    st = lib.ttf_dbfun3(0, 2, 1, 1, 2, 1, 1, 1, 1, None, None, None, None, 1, None, 1)
    assert st != 0
    assert b"invalid sizes" in lib.ttf_last_error()
    lib.ttf_clear_error()


def test_dbfun3_rect_tiles_pinned(ttlib, tmp_path):
    """The 64 x 80 GEMM tiles (four waves in rows), which the autotuner may
    pick for stage 1 when N is a multiple of 80, pinned through a tuning
    table in a child process (the table is read at the first GEMM of a
    process) at the (128, 2, 128, R = 5) shape, against the oracle; both
    K chunks (32 and 16) are covered."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    tune = tmp_path / "tune.txt"
    # type M N K nbatch flags bm bk ks bn: stage 1 is M=256 N=640 K=128 with
    # A and C row-fast (flags 5), one batch per vector
    pinned = ["0 256 640 128 8 5 64 32 1 80", "0 256 640 128 4 5 64 16 1 80"]
    tune.write_text("\n".join(pinned) + "\n")
    code = r'''
import json, os, sys
root = sys.argv[1]
sys.path[:0] = [root, os.path.join(root, "tt-fort_amd")]
import numpy as np
import ttfort
from oracle import kernels as K
lib = ttfort.load()
ttfort.check(lib.ttf_init(0), lib)
shape = (128, 2, 128, 128, 2, 128, 5, 5)
rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
rng = np.random.default_rng(21)
worst = 0.0
for nvec in (8, 4):
    phi1 = rng.standard_normal(ry1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * m * ra2)
    phi2 = rng.standard_normal(rx2 * ra2 * ry2)
    x = rng.standard_normal(rx1 * m * rx2 * nvec)
    y = ttfort.bfun3(shape, phi1, A, phi2, x, nvec=nvec)
    nx, ny = rx1 * m * rx2, ry1 * n * ry2
    scale = (np.abs(phi1).max() * np.abs(A).max() * np.abs(phi2).max() * np.abs(x).max()
             * rx1 * m * rx2 * ra1 * ra2)
    for v in range(nvec):
        ref = K.dbfun3(*shape, phi1, A, phi2, x[v * nx:(v + 1) * nx])
        worst = max(worst, float(np.abs(y[v * ny:(v + 1) * ny] - ref).max() / scale))
print(json.dumps({"rel": worst}))
'''
    # the GEMM chain, not the fused kernel, at this shape
    env = dict(os.environ, TTF_TUNE_FILE=str(tune), TTF_BFUN3_FUSED="0")
    r = subprocess.run([sys.executable, "-c", code, root], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["rel"] <= 1e-13, res
    # the pinned stage-1 entries were used: no new stage-1 line was tuned and appended
    lines = [ln.strip() for ln in tune.read_text().splitlines() if ln.strip()]
    assert [ln for ln in lines if ln.startswith("0 256 640 128 ")] == pinned, lines


@pytest.mark.parametrize("nvec", [1, 8, 16])
def test_dbfun3_fused_matches_gemm_chain(ttlib, tmp_path, nvec):
    """At the (128, 2, 128, R = 5) local shape the library runs the fused
    one-launch kernel; the GEMM chain (TTF_BFUN3_FUSED=0, in a child
    process) gives the same products to 1e-13 of the problem scale, and two
    fused calls give identical bits (the two halves of the b1 range meet by
    commutative additions onto zero)."""
    import json
    import os
    import subprocess
    import sys
    import ttfort
    shape = (128, 2, 128, 128, 2, 128, 5, 5)
    rng = np.random.default_rng(30 + nvec)
    t = _rand(rng, shape, nvec)
    for k in ("phi1", "A", "phi2", "x"):
        np.save(tmp_path / (k + ".npy"), t[k])
    lib = ttfort.load()
    d = {k: ttfort.DeviceArray.from_numpy(t[k]) for k in ("phi1", "A", "phi2", "x")}
    nx = ny = 128 * 2 * 128
    dy = ttfort.DeviceArray((ny * nvec,))
    ms = ttfort.C.c_float()
    fused = ttfort.C.c_int()
    ttfort.check(lib.ttf_dbfun3_chain_profile(*shape, nvec, d["phi1"].ptr, d["A"].ptr, d["phi2"].ptr,
                                              d["x"].ptr, nx, dy.ptr, ny, 3, ttfort.C.byref(ms),
                                              ttfort.C.byref(fused)), lib)
    assert fused.value == 1
    y1 = dy.numpy().copy()
    y2 = ttfort.bfun3(shape, t["phi1"], t["A"], t["phi2"], t["x"], nvec=nvec)
    assert np.array_equal(y1, y2)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = r'''
import os, sys
root, d, nvec = sys.argv[1], sys.argv[2], int(sys.argv[3])
sys.path[:0] = [root, os.path.join(root, "tt-fort_amd")]
import numpy as np
import ttfort
t = {k: np.load(os.path.join(d, k + ".npy")) for k in ("phi1", "A", "phi2", "x")}
y = ttfort.bfun3((128, 2, 128, 128, 2, 128, 5, 5), t["phi1"], t["A"], t["phi2"], t["x"], nvec=nvec)
np.save(os.path.join(d, "y_chain.npy"), y)
'''
    env = dict(os.environ, TTF_BFUN3_FUSED="0")
    r = subprocess.run([sys.executable, "-c", code, root, str(tmp_path), str(nvec)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    yc = np.load(tmp_path / "y_chain.npy")
    assert np.abs(y1 - yc).max() <= 1e-13 * _scale(shape, t)

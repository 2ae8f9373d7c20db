"""ttf_dbfun3_noalloc, the block matvec that allocates nothing (its stage
intermediates and split-K slices live in a fixed 64 MiB device array), is
checked against ttf_dbfun3 on the same inputs: two calls in a row, the
second with a smaller problem, through the C ABI.  The two entry points run
the same kernels with the same split-K factors, so the products agree to
rounding (1e-13 relative, the contraction bar is 1e-12); the free device
memory does not drop across no-allocation calls; a product too large for
the array is refused with an error before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mem_free():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _inputs(rng, shape, nvec):
    import ttfort
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    host = dict(phi1=rng.standard_normal(ry1 * rx1 * ra1), A=rng.standard_normal(ra1 * n * m * ra2),
                phi2=rng.standard_normal(rx2 * ra2 * ry2), x=rng.standard_normal(rx1 * m * rx2 * nvec))
    dev = {k: ttfort.DeviceArray.from_numpy(v) for k, v in host.items()}
    return host, dev


def _call(lib, name, shape, nvec, dev, y):
    import ttfort
    rx1, m, rx2, ry1, n, ry2 = shape[:6]
    f = getattr(lib, name)
    ttfort.check(f(*shape, nvec, dev["phi1"].ptr, dev["A"].ptr, dev["phi2"].ptr, dev["x"].ptr, rx1 * m * rx2,
                   y.ptr, ry1 * n * ry2), lib)
    ttfort.sync(lib)
    return y.numpy()


@pytest.mark.parametrize("calls", [[((128, 2, 128, 128, 2, 128, 5, 5), 8), ((40, 2, 36, 40, 2, 36, 5, 5), 3)]])
def test_noalloc_matches_dbfun3(ttlib, calls):
    import ttfort
    lib = ttlib
    rng = np.random.default_rng(77)
    # one small product first, so that loading the kernels' code objects
    # does not count against the no-allocation calls below
    small = (4, 2, 4, 4, 2, 4, 2, 2)
    _, dev = _inputs(rng, small, 1)
    _call(lib, "ttf_dbfun3_noalloc", small, 1, dev, ttfort.DeviceArray((32,)))
    for shape, nvec in calls:
        print("no-alloc bfun3 shape", shape, "nvec", nvec, flush=True)
        host, dev = _inputs(rng, shape, nvec)
        ry1, n, ry2 = shape[3:6]
        y0 = ttfort.DeviceArray((ry1 * n * ry2 * nvec,))
        y1 = ttfort.DeviceArray((ry1 * n * ry2 * nvec,))
        free0 = _mem_free()
        got = _call(lib, "ttf_dbfun3_noalloc", shape, nvec, dev, y1)
        free1 = _mem_free()
        assert free0 - free1 < (1 << 20), (free0, free1)
        ref = _call(lib, "ttf_dbfun3", shape, nvec, dev, y0)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("max relative difference", err, "identical bits", bool(np.array_equal(got, ref)), flush=True)
        assert err <= 1e-13, err


def test_noalloc_rejects_oversized_products(ttlib):
    import ttfort
    lib = ttlib
    shape = (128, 2, 128, 128, 2, 128, 5, 5)
    # 32 vectors need about 19 M doubles of scratch, over the 8 M of the array;
    # the refusal comes before any launch, so the pointers are never read
    rc = lib.ttf_dbfun3_noalloc(*shape, 32, None, None, None, None, 128 * 2 * 128, None, 128 * 2 * 128)
    assert rc != 0
    msg = lib.ttf_last_error().decode()
    assert "ttf_dbfun3_noalloc" in msg and "scratch" in msg, msg
    lib.ttf_clear_error()

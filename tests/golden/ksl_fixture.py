"""Reader of the reference's KSL fixtures (test_ksl.dat, test_eye_ksl.dat;
layout of test_ksl.f90:9-33: Fortran stream records of int32 d, n(d), m(d),
ra(d+1), ry(d+1), Asize, complex128 crA(Asize), Ysize, complex128
crY(Ysize), float64 tau, int32 rmax, kickrank, nswp, verb)."""
import numpy as np


def read_ksl_dat(path):
    b = open(path, "rb").read()
    o = 0

    def ints(k):
        nonlocal o
        v = np.frombuffer(b, dtype="<i4", count=k, offset=o)
        o += 4 * k
        return v.astype(np.int64)

    d = int(ints(1)[0])
    n, m, ra, ry = ints(d), ints(d), ints(d + 1), ints(d + 1)
    asz = int(ints(1)[0])
    A = np.frombuffer(b, dtype="<c16", count=asz, offset=o).copy()
    o += 16 * asz
    ysz = int(ints(1)[0])
    Y = np.frombuffer(b, dtype="<c16", count=ysz, offset=o).copy()
    o += 16 * ysz
    tau = float(np.frombuffer(b, dtype="<f8", count=1, offset=o)[0])
    o += 8
    rmax, kickrank, nswp, verb = (int(x) for x in ints(4))
    assert o == len(b)
    return dict(d=d, n=n, m=m, ra=ra, ry=ry, A=A, Y=Y, tau=tau, rmax=rmax, kickrank=kickrank, nswp=nswp, verb=verb)

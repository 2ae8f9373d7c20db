// Reference-compatible entry points: the tt-fort `bind(c)` symbols with
// their by-reference Fortran argument lists (tt_linalg.f90), routed to the
// device.  Host arrays are staged through the library's host-I/O arena on
// every call; the native ttf_* entry points keep data resident instead.
//
// Built twice: with ttf_fint = int32 (libttfort_amd.so) and with
// -DTTF_FINT64, ttf_fint = int64 (libttfort_amd_i8.so), matching the
// reference's i4 and i8 build presets.
#include "contract.h"

#include <cstring>

using namespace ttf;

extern "C" {

// tt_linalg.f90:127-159.  res1/res2 are the caller's workspaces of the
// reference implementation; the device path keeps its own.
void dbfun3(const ttf_fint* rx1, const ttf_fint* m, const ttf_fint* rx2, const ttf_fint* ry1,
            const ttf_fint* n, const ttf_fint* ry2, const ttf_fint* ra1, const ttf_fint* ra2,
            const double* phi1, const double* A, const double* phi2, const double* x, double* y,
            double* /*res1*/, double* /*res2*/) {
  guarded([&] {
    Shape s{(int)*rx1, (int)*m, (int)*rx2, (int)*ry1, (int)*n, (int)*ry2, (int)*ra1, (int)*ra2};
    TTF_REQUIRE(s.rx1 > 0 && s.m > 0 && s.rx2 > 0 && s.ry1 > 0 && s.n > 0 && s.ry2 > 0 &&
                    s.ra1 > 0 && s.ra2 > 0,
                "dbfun3: invalid sizes");
    const size_t n1 = (size_t)s.ry1 * s.rx1 * s.ra1;
    const size_t na = (size_t)s.ra1 * s.n * s.m * s.ra2;
    const size_t n2 = (size_t)s.rx2 * s.ra2 * s.ry2;
    const size_t nx = (size_t)s.rx1 * s.m * s.rx2;
    const size_t ny = (size_t)s.ry1 * s.n * s.ry2;
    hipStream_t st = stream();
    double* d = (double*)arena(kArenaHostIO, sizeof(double) * (n1 + na + n2 + nx + ny));
    double *dphi1 = d, *dA = d + n1, *dphi2 = dA + na, *dx = dphi2 + n2, *dy = dx + nx;
    TTF_HIP(hipMemcpyAsync(dphi1, phi1, n1 * sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dA, A, na * sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dphi2, phi2, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dx, x, nx * sizeof(double), hipMemcpyHostToDevice, st));
    bfun3_chain<double>(s, 1, dphi1, dA, dphi2, dx, (int64_t)nx, dy, (int64_t)ny, st);
    read_back(y, dy, ny * sizeof(double), st);
  });
}

}  // extern "C"

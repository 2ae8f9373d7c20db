// Reference-compatible entry points: the tt-fort `bind(c)` symbols with
// their by-reference Fortran argument lists (tt_linalg.f90:30-845), and C
// versions of the local-operator callbacks the reference hands PRIMME and
// Expokit (tt_eigb.f90:15-49, tt_ksl.f90:20-90), all routed to the device.
// Host arrays are staged through the library's host-I/O arenas on every
// call; the native ttf_* entry points keep data resident instead.
//
// Built twice: with ttf_fint = int32 (libttfort_amd.so) and with
// -DTTF_FINT64, ttf_fint = int64 (libttfort_amd_i8.so), matching the
// reference's i4 and i8 build presets.  The reference's Fortran `character`
// arguments arrive as pointers to one byte.
//
// These routines have no status return (the reference's have none either):
// a failure leaves the message in ttf_last_error(), which a caller checks
// when it needs to.
#include "amen.h"
#include "linalg.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace ttf;

namespace {

typedef hipDoubleComplex zc;

// Device copies of host arrays for one call: inputs first, outputs after.
struct Staging {
  std::vector<size_t> off;
  size_t total = 0;
  double* base = nullptr;
  size_t add(size_t n) {
    off.push_back(total);
    total += (n + 31) & ~size_t(31);   // 256-byte aligned slices
    return off.size() - 1;
  }
  void alloc() { base = (double*)arena(kArenaHostIO, sizeof(double) * std::max<size_t>(total, 1)); }
  double* at(size_t i) const { return base + off[i]; }
};

void upload(double* d, const double* h, size_t n, hipStream_t st) {
  if (n) TTF_HIP(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, st));
}

// C (m + k, n) = [A (m, n); B (k, n)]: the rows of B appended under A
__global__ void row_add_kernel(int m, int n, int k, const double* __restrict__ A, const double* __restrict__ B,
                               double* __restrict__ C) {
  const int64_t mk = (int64_t)m + k, total = mk * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % mk, j = e / mk;
    C[e] = i < m ? A[i + (int64_t)m * j] : B[(i - m) + (int64_t)k * j];
  }
}

// C (k, n) = the first k rows of A (m, n)
__global__ void row_cut_kernel(int m, int n, int k, const double* __restrict__ A, double* __restrict__ C) {
  const int64_t total = (int64_t)k * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % k, j = e / k;
    C[e] = A[i + (int64_t)m * j];
  }
}

// B (n1, n3, n2, n4) = A (n1, n2, n3, n4), one output element per thread
__global__ void perm1324_kernel(int n1, int n2, int n3, int64_t total, const double* __restrict__ A,
                                double* __restrict__ B) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t t = e;
    const int i1 = (int)(t % n1); t /= n1;
    const int i3 = (int)(t % n3); t /= n3;
    const int i2 = (int)(t % n2);
    const int64_t i4 = t / n2;
    B[e] = A[i1 + (int64_t)n1 * (i2 + (int64_t)n2 * (i3 + (int64_t)n3 * i4))];
  }
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 16384); }

// ---- callback context: the sizes and interfaces of one local step
struct Callback {
  Shape s{};
  int64_t xsize = 0, ysize = 0;
  bool real_ready = false, complex_ready = false;
  double* dphi1 = nullptr;   // one device block: phi1 | A | phi2 (real or complex)
  size_t cap = 0;
  size_t n1 = 0, na = 0, n2 = 0;
};
Callback g_cb;

double* callback_block(size_t doubles) {
  if (g_cb.cap < doubles) {
    if (g_cb.dphi1) {
      TTF_HIP(hipStreamSynchronize(stream()));
      TTF_HIP(hipFree(g_cb.dphi1));
      g_cb.dphi1 = nullptr;
      g_cb.cap = 0;
    }
    TTF_HIP(hipMalloc((void**)&g_cb.dphi1, sizeof(double) * doubles));
    g_cb.cap = doubles;
  }
  return g_cb.dphi1;
}

}  // namespace

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
    upload(dphi1, phi1, n1, st);
    upload(dA, A, na, st);
    upload(dphi2, phi2, n2, st);
    upload(dx, x, nx, st);
    bfun3_chain<double>(s, 1, dphi1, dA, dphi2, dx, (int64_t)nx, dy, (int64_t)ny, st);
    read_back(y, dy, ny * sizeof(double), st);
  });
}

// tt_linalg.f90:30-51: Householder QR in place.  A (n, m) returns Q in its
// first min(n, m) columns (dorgqr) and, when m > n, R's trailing columns in
// the rest (what dgeqrf leaves there); R is the reference's R(min(n, m), n):
// zeroed, then its upper triangle R(1:min(j, n), j), j = 1..m, written with
// leading dimension min(n, m).  work, lwork and tau are not used.
void dqr(const ttf_fint* n_, const ttf_fint* m_, double* A, double* R, double* /*work*/,
         const ttf_fint* /*lwork*/, double* /*tau*/) {
  guarded([&] {
    const int n = (int)*n_, m = (int)*m_;
    TTF_REQUIRE(n >= 1 && m >= 1, "dqr: invalid sizes n=%d m=%d", n, m);
    const int k = std::min(n, m);
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)n * m), iq = sg.add((size_t)n * k), ir = sg.add((size_t)k * m);
    sg.alloc();
    upload(sg.at(ia), A, (size_t)n * m, st);
    householder_qr(n, m, sg.at(ia), sg.at(iq), sg.at(ir), k, st);
    std::vector<double> hr((size_t)k * m);
    read_back(hr.data(), sg.at(ir), sizeof(double) * hr.size(), st);
    read_back(A, sg.at(iq), sizeof(double) * (size_t)n * k, st);
    std::memset(R, 0, sizeof(double) * (size_t)k * n);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i <= std::min(j, k - 1); ++i) R[i + (size_t)k * j] = hr[i + (size_t)k * j];
    for (int j = k; j < m; ++j)
      for (int i = 0; i < n; ++i) A[i + (size_t)n * j] = hr[i + (size_t)k * j];
  });
}

// tt_linalg.f90:54-69: the k rows of B (k, n) appended under A (m, n);
// A becomes (m + k, n) in place and must have room for it.
void drow_add(const ttf_fint* m_, const ttf_fint* n_, const ttf_fint* k_, double* A, const double* B) {
  guarded([&] {
    const int m = (int)*m_, n = (int)*n_, k = (int)*k_;
    TTF_REQUIRE(m >= 0 && n >= 0 && k >= 0, "drow_add: invalid sizes");
    if ((int64_t)(m + k) * n == 0) return;
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)m * n), ib = sg.add((size_t)k * n), ic = sg.add((size_t)(m + k) * n);
    sg.alloc();
    upload(sg.at(ia), A, (size_t)m * n, st);
    upload(sg.at(ib), B, (size_t)k * n, st);
    hipLaunchKernelGGL(row_add_kernel, dim3(grid_for((int64_t)(m + k) * n)), dim3(256), 0, st, m, n, k, sg.at(ia),
                       sg.at(ib), sg.at(ic));
    TTF_HIP(hipGetLastError());
    read_back(A, sg.at(ic), sizeof(double) * (size_t)(m + k) * n, st);
  });
}

// tt_linalg.f90:71-79: keep the first k rows of A (m, n), packed as (k, n).
void drow_cut(const ttf_fint* m_, const ttf_fint* n_, const ttf_fint* k_, double* A) {
  guarded([&] {
    const int m = (int)*m_, n = (int)*n_, k = (int)*k_;
    TTF_REQUIRE(m >= 0 && n >= 0 && k >= 0 && k <= m, "drow_cut: invalid sizes m=%d n=%d k=%d", m, n, k);
    if ((int64_t)k * n == 0) return;
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)m * n), ic = sg.add((size_t)k * n);
    sg.alloc();
    upload(sg.at(ia), A, (size_t)m * n, st);
    hipLaunchKernelGGL(row_cut_kernel, dim3(grid_for((int64_t)k * n)), dim3(256), 0, st, m, n, k, sg.at(ia),
                       sg.at(ic));
    TTF_HIP(hipGetLastError());
    read_back(A, sg.at(ic), sizeof(double) * (size_t)k * n, st);
  });
}

// tt_linalg.f90:82-98: B (m, n) = A (n, m)^T.
void dtransp(const ttf_fint* n_, const ttf_fint* m_, const double* A, double* B) {
  guarded([&] {
    const int n = (int)*n_, m = (int)*m_;
    TTF_REQUIRE(n >= 0 && m >= 0, "dtransp: invalid sizes");
    if ((int64_t)n * m == 0) return;
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)n * m), ib = sg.add((size_t)n * m);
    sg.alloc();
    upload(sg.at(ia), A, (size_t)n * m, st);
    transpose(n, m, sg.at(ia), n, sg.at(ib), m, st);
    read_back(B, sg.at(ib), sizeof(double) * (size_t)n * m, st);
  });
}

// tt_linalg.f90:101-121: B (n1, n3, n2, n4) = A (n1, n2, n3, n4).
void dperm1324(const ttf_fint* n1_, const ttf_fint* n2_, const ttf_fint* n3_, const ttf_fint* n4_, const double* A,
               double* B) {
  guarded([&] {
    const int n1 = (int)*n1_, n2 = (int)*n2_, n3 = (int)*n3_, n4 = (int)*n4_;
    TTF_REQUIRE(n1 >= 0 && n2 >= 0 && n3 >= 0 && n4 >= 0, "dperm1324: invalid sizes");
    const int64_t total = (int64_t)n1 * n2 * n3 * n4;
    if (total == 0) return;
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)total), ib = sg.add((size_t)total);
    sg.alloc();
    upload(sg.at(ia), A, (size_t)total, st);
    hipLaunchKernelGGL(perm1324_kernel, dim3(grid_for(total)), dim3(256), 0, st, n1, n2, n3, total, sg.at(ia),
                       sg.at(ib));
    TTF_HIP(hipGetLastError());
    read_back(B, sg.at(ib), sizeof(double) * (size_t)total, st);
  });
}

// tt_linalg.f90:363-495: block-Jacobi blocks of a local operator whose x
// and y interfaces coincide: Phi1 (rx1, rx1, ra1), A (ra1, n, n, ra2),
// Phi2 (rx2, ra2, rx2); jacs in the reference's layout per ptype.
void djac_gen(const char* ptype, const ttf_fint* rx1_, const ttf_fint* n_, const ttf_fint* rx2_,
              const ttf_fint* ra1_, const ttf_fint* ra2_, const double* Phi1, const double* A, const double* Phi2,
              double* jacs) {
  guarded([&] {
    const int rx1 = (int)*rx1_, n = (int)*n_, rx2 = (int)*rx2_, ra1 = (int)*ra1_, ra2 = (int)*ra2_;
    const char p = norm_ptype(*ptype);
    TTF_REQUIRE(p == 'c' || p == 'l' || p == 'r', "djac_gen: ptype '%c'", *ptype);
    const size_t nj = jac_size(p, rx1, n, rx2);
    hipStream_t st = stream();
    Staging sg;
    const size_t i1 = sg.add((size_t)rx1 * rx1 * ra1), ia = sg.add((size_t)ra1 * n * n * ra2),
                 i2 = sg.add((size_t)rx2 * ra2 * rx2), ij = sg.add(nj);
    sg.alloc();
    upload(sg.at(i1), Phi1, (size_t)rx1 * rx1 * ra1, st);
    upload(sg.at(ia), A, (size_t)ra1 * n * n * ra2, st);
    upload(sg.at(i2), Phi2, (size_t)rx2 * ra2 * rx2, st);
    jac_gen(p, rx1, n, rx2, ra1, ra2, sg.at(i1), sg.at(ia), sg.at(i2), sg.at(ij), st);
    read_back(jacs, sg.at(ij), sizeof(double) * nj, st);
  });
}

// tt_linalg.f90:497-533: y = P x with the blocks of djac_gen.  work1 is
// not used.
void djac_apply(const char* ptype, const ttf_fint* rx1_, const ttf_fint* n_, const ttf_fint* rx2_,
                const double* jacs, const double* x, double* y, double* /*work1*/) {
  guarded([&] {
    const int rx1 = (int)*rx1_, n = (int)*n_, rx2 = (int)*rx2_;
    const char p = norm_ptype(*ptype);
    TTF_REQUIRE(p == 'c' || p == 'l' || p == 'r', "djac_apply: ptype '%c'", *ptype);
    const size_t nj = jac_size(p, rx1, n, rx2), nx = (size_t)rx1 * n * rx2;
    hipStream_t st = stream();
    Staging sg;
    const size_t ij = sg.add(nj), ix = sg.add(nx), iy = sg.add(nx);
    sg.alloc();
    upload(sg.at(ij), jacs, nj, st);
    upload(sg.at(ix), x, nx, st);
    jac_apply(p, rx1, n, rx2, sg.at(ij), sg.at(ix), sg.at(iy), st);
    read_back(y, sg.at(iy), sizeof(double) * nx, st);
  });
}

// tt_linalg.f90:540-643: pivoted incomplete Cholesky of the Gram matrix of
// the n items of dA (m x n for trans 'N', n x m for 'T'): du (n, numvec),
// lam (numvec), real_numvec.
void duchol_fort(const char* trans, const ttf_fint* m_, const ttf_fint* n_, const double* dA,
                 const ttf_fint* numvec_, double* du, double* lam, ttf_fint* real_numvec) {
  guarded([&] {
    const int m = (int)*m_, n = (int)*n_, numvec = (int)*numvec_;
    TTF_REQUIRE(m >= 1 && n >= 1 && numvec >= 1, "duchol_fort: invalid sizes m=%d n=%d numvec=%d", m, n, numvec);
    hipStream_t st = stream();
    Staging sg;
    const size_t ia = sg.add((size_t)m * n), iu = sg.add((size_t)n * numvec);
    sg.alloc();
    upload(sg.at(ia), dA, (size_t)m * n, st);
    std::vector<double> hl(numvec, 0.0);
    const int r = ttf::duchol(*trans, m, n, sg.at(ia), numvec, sg.at(iu), hl.data(), st);
    read_back(du, sg.at(iu), sizeof(double) * (size_t)n * numvec, st);
    for (int i = 0; i < std::min(r, numvec); ++i) lam[i] = hl[i];
    *real_numvec = (ttf_fint)r;
  });
}

// tt_linalg.f90:651-845: right-preconditioned restarted GMRES for the local
// system with Phi1 (rx1, rx1, ra1), A (ra1, n, n, ra2), Phi2 (rx2, ra2,
// rx2); sol (rx1 n rx2) is the solution before the preconditioner, as the
// reference returns it.  Arnoldi by classical Gram-Schmidt applied twice;
// the Krylov dimension is at most 63.  verb is not used.
void dgmresr_hh_fort(const double* Phi1, const double* A, const double* Phi2, const double* rhs,
                     const ttf_fint* rx1_, const ttf_fint* n_, const ttf_fint* rx2_, const ttf_fint* ra1_,
                     const ttf_fint* ra2_, const ttf_fint* nrestart, const double* tol, const ttf_fint* niters,
                     const char* ptype, const double* jacs, double* sol, const ttf_fint* /*verb*/) {
  guarded([&] {
    const int rx1 = (int)*rx1_, n = (int)*n_, rx2 = (int)*rx2_, ra1 = (int)*ra1_, ra2 = (int)*ra2_;
    const char p = norm_ptype(*ptype);
    TTF_REQUIRE(p == 'n' || p == 'c' || p == 'l' || p == 'r', "dgmresr_hh_fort: ptype '%c'", *ptype);
    const size_t nj = p == 'n' ? 0 : jac_size(p, rx1, n, rx2), nx = (size_t)rx1 * n * rx2;
    hipStream_t st = stream();
    Staging sg;
    const size_t i1 = sg.add((size_t)rx1 * rx1 * ra1), ia = sg.add((size_t)ra1 * n * n * ra2),
                 i2 = sg.add((size_t)rx2 * ra2 * rx2), ir = sg.add(nx), ij = sg.add(nj), is = sg.add(nx);
    sg.alloc();
    upload(sg.at(i1), Phi1, (size_t)rx1 * rx1 * ra1, st);
    upload(sg.at(ia), A, (size_t)ra1 * n * n * ra2, st);
    upload(sg.at(i2), Phi2, (size_t)rx2 * ra2 * rx2, st);
    upload(sg.at(ir), rhs, nx, st);
    upload(sg.at(ij), jacs, nj, st);
    LocalOp op{Shape{rx1, n, rx2, rx1, n, rx2, ra1, ra2}, sg.at(i1), sg.at(ia), sg.at(i2)};
    JacPrec P;
    if (p != 'n') P = JacPrec{p, rx1, n, rx2, sg.at(ij)};
    gmres_solve(op, P, sg.at(ir), sg.at(is), (int)*nrestart, *tol, (int)*niters, st);
    read_back(sol, sg.at(is), sizeof(double) * nx, st);
  });
}

// ---- the callbacks of the local eigen and exponential solvers.
// init_bfun_sizes / init_bfun_main (tt_eigb.f90:29-49, tt_ksl.f90:55-90)
// bind one local step's sizes and interfaces; here the interfaces and the
// operator core are uploaded once, so each callback moves only x and y.
void ttf_init_bfun_sizes(const ttf_fint* rx1, const ttf_fint* m, const ttf_fint* rx2, const ttf_fint* ry1,
                         const ttf_fint* n, const ttf_fint* ry2, const ttf_fint* ra1, const ttf_fint* ra2,
                         const ttf_fint* xsize, const ttf_fint* ysize) {
  guarded([&] {
    Shape s{(int)*rx1, (int)*m, (int)*rx2, (int)*ry1, (int)*n, (int)*ry2, (int)*ra1, (int)*ra2};
    TTF_REQUIRE(s.rx1 > 0 && s.m > 0 && s.rx2 > 0 && s.ry1 > 0 && s.n > 0 && s.ry2 > 0 && s.ra1 > 0 &&
                    s.ra2 > 0 && *xsize >= (ttf_fint)s.rx1 * s.m * s.rx2 && *ysize >= (ttf_fint)s.ry1 * s.n * s.ry2,
                "init_bfun_sizes: invalid sizes");
    g_cb.s = s;
    g_cb.xsize = (int64_t)*xsize;
    g_cb.ysize = (int64_t)*ysize;
    g_cb.n1 = (size_t)s.ry1 * s.rx1 * s.ra1;
    g_cb.na = (size_t)s.ra1 * s.n * s.m * s.ra2;
    g_cb.n2 = (size_t)s.rx2 * s.ra2 * s.ry2;
    g_cb.real_ready = g_cb.complex_ready = false;
  });
}

void ttf_init_bfun_main(const double* phi1, const double* A, const double* phi2) {
  guarded([&] {
    TTF_REQUIRE(g_cb.n1 > 0, "init_bfun_main: init_bfun_sizes first");
    hipStream_t st = stream();
    double* d = callback_block(g_cb.n1 + g_cb.na + g_cb.n2);
    upload(d, phi1, g_cb.n1, st);
    upload(d + g_cb.n1, A, g_cb.na, st);
    upload(d + g_cb.n1 + g_cb.na, phi2, g_cb.n2, st);
    g_cb.real_ready = true;
    g_cb.complex_ready = false;
  });
}

void ttf_zinit_bfun_main(const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* phi2) {
  guarded([&] {
    TTF_REQUIRE(g_cb.n1 > 0, "init_bfun_main: init_bfun_sizes first");
    hipStream_t st = stream();
    double* d = callback_block(2 * (g_cb.n1 + g_cb.na + g_cb.n2));
    upload(d, (const double*)phi1, 2 * g_cb.n1, st);
    upload(d + 2 * g_cb.n1, (const double*)A, 2 * g_cb.na, st);
    upload(d + 2 * (g_cb.n1 + g_cb.na), (const double*)phi2, 2 * g_cb.n2, st);
    g_cb.complex_ready = true;
    g_cb.real_ready = false;
  });
}

// PRIMME's matrixMatvec (primme/primme.h:140-141), as tt_eigb.f90:15-27
// binds it: y(:, b) = B x(:, b) for the blockSize columns of the host
// arrays x (leading dimension xsize) and y (ysize), one block product on
// the device.  The PRIMME parameter block is not read.
void ttf_primme_matvec(void* x, void* y, ttf_fint* blockSize, void* /*primme*/) {
  guarded([&] {
    TTF_REQUIRE(g_cb.real_ready, "primme_matvec: init_bfun_main first");
    const int k = (int)*blockSize;
    if (k <= 0) return;
    const Shape& s = g_cb.s;
    hipStream_t st = stream();
    const size_t nx = (size_t)g_cb.xsize * k, ny = (size_t)g_cb.ysize * k;
    double* d = (double*)arena(kArenaHostIO, sizeof(double) * (nx + ny));
    upload(d, (const double*)x, nx, st);
    const double* dp = g_cb.dphi1;
    bfun3_chain<double>(s, k, dp, dp + g_cb.n1, dp + g_cb.n1 + g_cb.na, d, g_cb.xsize, d + nx, g_cb.ysize, st);
    read_back(y, d + nx, sizeof(double) * ny, st);
  });
}

// Expokit's matvec(x, y) and normest's matvec_transp(x, y) as
// tt_ksl.f90:25-50 binds them: y = B x and y = B^H x, complex(8).
static void zcallback(const ttf_zdouble* x, ttf_zdouble* y, bool transp) {
  TTF_REQUIRE(g_cb.complex_ready, "zmatvec: init_bfun_main (complex) first");
  const Shape& s = g_cb.s;
  hipStream_t st = stream();
  const size_t nx = transp ? (size_t)s.ry1 * s.n * s.ry2 : (size_t)s.rx1 * s.m * s.rx2;
  const size_t ny = transp ? (size_t)s.rx1 * s.m * s.rx2 : (size_t)s.ry1 * s.n * s.ry2;
  zc* d = (zc*)arena(kArenaHostIO, sizeof(zc) * (nx + ny));
  upload((double*)d, (const double*)x, 2 * nx, st);
  const zc* dp = (const zc*)g_cb.dphi1;
  if (transp)
    zbfun3_transp_chain(s, 1, dp, dp + g_cb.n1, dp + g_cb.n1 + g_cb.na, d, (int64_t)nx, d + nx, (int64_t)ny, st);
  else
    zbfun3_chain(s, 1, dp, dp + g_cb.n1, dp + g_cb.n1 + g_cb.na, d, (int64_t)nx, d + nx, (int64_t)ny, st);
  read_back(y, d + nx, sizeof(zc) * ny, st);
}

void ttf_zmatvec(const ttf_zdouble* x, ttf_zdouble* y) {
  guarded([&] { zcallback(x, y, false); });
}

void ttf_zmatvec_transp(const ttf_zdouble* x, ttf_zdouble* y) {
  guarded([&] { zcallback(x, y, true); });
}

}  // extern "C"

// Local-operator contractions of the tt-fort hot path on the gfx950 matrix
// cores: the projected matvec (bfun3 family) and the interface (phi)
// updates.  Each is a chain of strided batched GEMMs (gemm_mfma.h) whose
// stride choices replace the reference's dgemm + dtransp sequences; the
// intermediates live in the library's contraction arena.
//
// Index names follow the reference comments (ttals.f90:18-29):
//   x (b1,j,b2) = (rx1,m,rx2)      y (c1,i,c2) = (ry1,n,ry2)
//   A (a1,i,j,a2) = (ra1,n,m,ra2)  phi1 (c1,b1,a1)  phi2 (b2,a2,c2)
#include "contract.h"

#include <type_traits>

namespace ttf {

// ------------------------------------------------------------------ bfun3
// Stage 1 (ttals.f90:20):  T1(b1,j,a2,c2) = sum_b2 x(b1,j,b2) phi2(b2,a2,c2)
// Stage 2 (ttals.f90:24):  T2(b1,a1,i,c2) = sum_{j,a2} A(a1,i,j,a2) T1(b1,j,a2,c2)
//                          (batched over (c2, vector); written directly in the
//                           layout the reference reaches after dtransp :26)
// Stage 3 (ttals.f90:29):  y(c1,i,c2) = sum_{b1,a1} phi1(c1,b1,a1) T2(b1,a1,i,c2)
template <typename T>
static GemmOp<T> bfun3_op1(const Shape& s, int nvec, const T* phi2, const T* x, int64_t ldx, T* T1) {
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  GemmOp<T> g{};
  g.M = s.rx1 * s.m; g.N = s.ra2 * s.ry2; g.K = s.rx2;
  g.A = x; g.sAi = 1; g.sAk = (int64_t)s.rx1 * s.m; g.sAb0 = ldx; g.sAb1 = 0;
  g.B = phi2; g.sBk = 1; g.sBj = s.rx2; g.sBb0 = 0; g.sBb1 = 0;
  g.C = T1; g.sCi = 1; g.sCj = (int64_t)s.rx1 * s.m; g.sCb0 = sz1; g.sCb1 = 0;
  g.nb0 = nvec; g.nbatch = nvec; g.alpha = T(1); g.beta = T(0); g.ksplit = 1;
  return g;
}

template <typename T>
static GemmOp<T> bfun3_op2(const Shape& s, int nvec, const T* A, const T* T1, T* T2) {
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  GemmOp<T> g{};
  g.M = s.rx1; g.N = s.ra1 * s.n; g.K = s.m * s.ra2;
  g.A = T1; g.sAi = 1; g.sAk = s.rx1; g.sAb0 = (int64_t)s.rx1 * s.m * s.ra2; g.sAb1 = sz1;
  g.B = A; g.sBk = (int64_t)s.ra1 * s.n; g.sBj = 1; g.sBb0 = 0; g.sBb1 = 0;
  g.C = T2; g.sCi = 1; g.sCj = s.rx1; g.sCb0 = (int64_t)s.rx1 * s.ra1 * s.n; g.sCb1 = sz2;
  g.nb0 = s.ry2; g.nbatch = s.ry2 * nvec; g.alpha = T(1); g.beta = T(0); g.ksplit = 1;
  return g;
}

template <typename T>
static GemmOp<T> bfun3_op3(const Shape& s, int nvec, const T* phi1, const T* T2, T* y, int64_t ldy) {
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  GemmOp<T> g{};
  g.M = s.ry1; g.N = s.n * s.ry2; g.K = s.rx1 * s.ra1;
  g.A = phi1; g.sAi = 1; g.sAk = s.ry1; g.sAb0 = 0; g.sAb1 = 0;
  g.B = T2; g.sBk = 1; g.sBj = (int64_t)s.rx1 * s.ra1; g.sBb0 = sz2; g.sBb1 = 0;
  g.C = y; g.sCi = 1; g.sCj = s.ry1; g.sCb0 = ldy; g.sCb1 = 0;
  g.nb0 = nvec; g.nbatch = nvec; g.alpha = T(1); g.beta = T(0); g.ksplit = 1;
  return g;
}

template <typename T>
void bfun3_stage12(const Shape& s, int nvec, const T* A, const T* phi2, const T* x, int64_t ldx,
                   T* T1, T* T2, hipStream_t st) {
  gemm_launch(bfun3_op1(s, nvec, phi2, x, ldx, T1), st);
  gemm_launch(bfun3_op2(s, nvec, A, (const T*)T1, T2), st);
}

// true when the column blocks [x, x + ldx nvec) and [y, y + ldy nvec) overlap
template <typename T>
static bool overlaps(const T* x, int64_t ldx, const T* y, int64_t ldy, int nvec) {
  const T* xe = x + ldx * nvec;
  const T* ye = y + ldy * nvec;
  return x < ye && y < xe;
}

template <typename T>
void bfun3_chain(const Shape& s, int nvec, const T* phi1, const T* A, const T* phi2, const T* x,
                 int64_t ldx, T* y, int64_t ldy, hipStream_t st) {
  if (nvec <= 0) return;
  if constexpr (std::is_same<T, double>::value) {
    if (!overlaps(x, ldx, y, ldy, nvec) && bfun3_fused(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st)) return;
  }
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  T* T1 = (T*)arena(kArenaBfun, sizeof(T) * (size_t)(sz1 + sz2) * nvec);
  T* T2 = T1 + sz1 * nvec;
  bfun3_stage12(s, nvec, A, phi2, x, ldx, T1, T2, st);
  gemm_launch(bfun3_op3(s, nvec, phi1, (const T*)T2, y, ldy), st);
}

size_t bfun3_workspace(const Shape& s, int nvec) {
  if (nvec <= 0) return 0;
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  const size_t inter = (size_t)(sz1 + sz2) * nvec;
  size_t split = 0;
  split = std::max(split, gemm_workspace(bfun3_op1<double>(s, nvec, nullptr, nullptr, 1, nullptr)));
  split = std::max(split, gemm_workspace(bfun3_op2<double>(s, nvec, nullptr, nullptr, nullptr)));
  split = std::max(split, gemm_workspace(bfun3_op3<double>(s, nvec, nullptr, nullptr, nullptr, 1)));
  return inter + split;
}

template <typename T>
void bfun3_chain_ws(const Shape& s, int nvec, const T* phi1, const T* A, const T* phi2, const T* x,
                    int64_t ldx, T* y, int64_t ldy, T* ws, size_t ws_cap, hipStream_t st) {
  if (nvec <= 0) return;
  // the fused kernel keeps its intermediates in LDS and allocates nothing
  if constexpr (std::is_same<T, double>::value) {
    if (!overlaps(x, ldx, y, ldy, nvec) && bfun3_fused(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st)) return;
  }
  const size_t need = bfun3_workspace(s, nvec);
  TTF_REQUIRE(ws != nullptr && need <= ws_cap, "bfun3: workspace of %zu elements exceeds the %zu given", need,
              ws_cap);
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  T* T1 = ws;
  T* T2 = T1 + sz1 * nvec;
  T* split = T2 + sz2 * nvec;
  const size_t split_cap = ws_cap - (size_t)(sz1 + sz2) * nvec;
  gemm_launch_ws(bfun3_op1(s, nvec, phi2, x, ldx, T1), split, split_cap, st);
  gemm_launch_ws(bfun3_op2(s, nvec, A, (const T*)T1, T2), split, split_cap, st);
  gemm_launch_ws(bfun3_op3(s, nvec, phi1, (const T*)T2, y, ldy), split, split_cap, st);
}

template void bfun3_chain_ws<double>(const Shape&, int, const double*, const double*, const double*,
                                     const double*, int64_t, double*, int64_t, double*, size_t, hipStream_t);
template void bfun3_stage12<double>(const Shape&, int, const double*, const double*, const double*,
                                    int64_t, double*, double*, hipStream_t);
template void bfun3_chain<double>(const Shape&, int, const double*, const double*, const double*,
                                  const double*, int64_t, double*, int64_t, hipStream_t);
template void bfun3_chain<float>(const Shape&, int, const float*, const float*, const float*,
                                 const float*, int64_t, float*, int64_t, hipStream_t);

// ------------------------------------------------------------------ complex
typedef hipDoubleComplex zc;
static const zc kZ1 = {1.0, 0.0};
static const zc kZ0 = {0.0, 0.0};

template <typename R>
static ZGemmOpT<R> zop() {
  ZGemmOpT<R> o{};
  o.nb0 = 1; o.nbatch = 1;
  o.alpha = CplxT<R>::make(R(1), R(0));
  o.beta = CplxT<R>::make(R(0), R(0));
  return o;
}

static void zlaunch(const ZGemmOp& g, hipStream_t st) { zgemm_launch(g, st); }
static void zlaunch(const CGemmOp& g, hipStream_t st) { cgemm_launch(g, st); }

// The complex chain in either precision (complex128 or complex64 operands,
// intermediates and accumulation).
template <typename R>
static void zbfun3_chain_t(const Shape& s, int nvec, const typename CplxT<R>::type* phi1,
                           const typename CplxT<R>::type* A, const typename CplxT<R>::type* phi2,
                           const typename CplxT<R>::type* x, int64_t ldx, typename CplxT<R>::type* y, int64_t ldy,
                           hipStream_t st) {
  typedef typename CplxT<R>::type C_t;
  if (nvec <= 0) return;
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  C_t* T1 = (C_t*)arena(kArenaBfun, sizeof(C_t) * (size_t)(sz1 + sz2) * nvec);
  C_t* T2 = T1 + sz1 * nvec;
  ZGemmOpT<R> g1 = zop<R>();
  g1.M = s.rx1 * s.m; g1.N = s.ra2 * s.ry2; g1.K = s.rx2;
  g1.A = x; g1.sAi = 1; g1.sAk = (int64_t)s.rx1 * s.m; g1.sAb0 = ldx;
  g1.B = phi2; g1.sBk = 1; g1.sBj = s.rx2;
  g1.C = T1; g1.sCi = 1; g1.sCj = (int64_t)s.rx1 * s.m; g1.sCb0 = sz1;
  g1.nb0 = nvec; g1.nbatch = nvec;
  zlaunch(g1, st);
  ZGemmOpT<R> g2 = zop<R>();
  g2.M = s.rx1; g2.N = s.ra1 * s.n; g2.K = s.m * s.ra2;
  g2.A = T1; g2.sAi = 1; g2.sAk = s.rx1; g2.sAb0 = (int64_t)s.rx1 * s.m * s.ra2; g2.sAb1 = sz1;
  g2.B = A; g2.sBk = (int64_t)s.ra1 * s.n; g2.sBj = 1;
  g2.C = T2; g2.sCi = 1; g2.sCj = s.rx1; g2.sCb0 = (int64_t)s.rx1 * s.ra1 * s.n; g2.sCb1 = sz2;
  g2.nb0 = s.ry2; g2.nbatch = s.ry2 * nvec;
  zlaunch(g2, st);
  ZGemmOpT<R> g3 = zop<R>();
  g3.M = s.ry1; g3.N = s.n * s.ry2; g3.K = s.rx1 * s.ra1;
  g3.A = phi1; g3.sAi = 1; g3.sAk = s.ry1;
  g3.B = T2; g3.sBk = 1; g3.sBj = (int64_t)s.rx1 * s.ra1; g3.sBb0 = sz2;
  g3.C = y; g3.sCi = 1; g3.sCj = s.ry1; g3.sCb0 = ldy;
  g3.nb0 = nvec; g3.nbatch = nvec;
  zlaunch(g3, st);
}

void zbfun3_chain(const Shape& s, int nvec, const zc* phi1, const zc* A, const zc* phi2,
                  const zc* x, int64_t ldx, zc* y, int64_t ldy, hipStream_t st) {
  zbfun3_chain_t<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st);
}

void cbfun3_chain(const Shape& s, int nvec, const hipFloatComplex* phi1, const hipFloatComplex* A,
                  const hipFloatComplex* phi2, const hipFloatComplex* x, int64_t ldx, hipFloatComplex* y,
                  int64_t ldy, hipStream_t st) {
  zbfun3_chain_t<float>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st);
}

// y = B^H x (ttals.f90:85-108): x is (c1,i,c2), y is (b1,j,b2).
//  R1(b1,a1,i,c2) = sum_c1 conj(phi1(c1,b1,a1)) x(c1,i,c2)
//  R2(b1,j,a2,c2) = sum_{a1,i} R1(b1,a1,i,c2) conj(A(a1,i,j,a2))   (per c2)
//  y(b1,j,b2)     = sum_{a2,c2} R2(b1,j,a2,c2) conj(phi2(b2,a2,c2))
void zbfun3_transp_chain(const Shape& s, int nvec, const zc* phi1, const zc* A, const zc* phi2,
                         const zc* x, int64_t ldx, zc* y, int64_t ldy, hipStream_t st) {
  if (nvec <= 0) return;
  const int64_t sz1 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  zc* R1 = (zc*)arena(kArenaBfun, sizeof(zc) * (size_t)(sz1 + sz2) * nvec);
  zc* R2 = R1 + sz1 * nvec;
  ZGemmOp g1 = zop<double>();
  g1.M = s.rx1 * s.ra1; g1.N = s.n * s.ry2; g1.K = s.ry1;
  g1.A = phi1; g1.sAi = s.ry1; g1.sAk = 1; g1.conjA = 1;
  g1.B = x; g1.sBk = 1; g1.sBj = s.ry1; g1.sBb0 = ldx;
  g1.C = R1; g1.sCi = 1; g1.sCj = (int64_t)s.rx1 * s.ra1; g1.sCb0 = sz1;
  g1.nb0 = nvec; g1.nbatch = nvec;
  zgemm_launch(g1, st);
  ZGemmOp g2 = zop<double>();
  g2.M = s.rx1; g2.N = s.m * s.ra2; g2.K = s.ra1 * s.n;
  g2.A = R1; g2.sAi = 1; g2.sAk = s.rx1; g2.sAb0 = (int64_t)s.rx1 * s.ra1 * s.n; g2.sAb1 = sz1;
  g2.B = A; g2.sBk = 1; g2.sBj = (int64_t)s.ra1 * s.n; g2.conjB = 1;
  g2.C = R2; g2.sCi = 1; g2.sCj = s.rx1; g2.sCb0 = (int64_t)s.rx1 * s.m * s.ra2; g2.sCb1 = sz2;
  g2.nb0 = s.ry2; g2.nbatch = s.ry2 * nvec;
  zgemm_launch(g2, st);
  ZGemmOp g3 = zop<double>();
  g3.M = s.rx1 * s.m; g3.N = s.rx2; g3.K = s.ra2 * s.ry2;
  g3.A = R2; g3.sAi = 1; g3.sAk = (int64_t)s.rx1 * s.m; g3.sAb0 = sz2;
  g3.B = phi2; g3.sBk = s.rx2; g3.sBj = 1; g3.conjB = 1;
  g3.C = y; g3.sCi = 1; g3.sCj = (int64_t)s.rx1 * s.m; g3.sCb0 = ldy;
  g3.nb0 = nvec; g3.nbatch = nvec;
  zgemm_launch(g3, st);
}

// ------------------------------------------------------------------ phi
// phi_left (tt_linalg.f90:287-312):
//  R1(b1,a1,i,c2) = sum_c1 phi1(c1,b1,a1) y(c1,i,c2)                 (:298)
//  R2(b1,j,a2,c2) = sum_{a1,i} R1(b1,a1,i,c2) A(a1,i,j,a2)   per c2  (:303)
//  new(c2,b2,a2)  = sum_{b1,j} R2(b1,j,a2,c2) x(b1,j,b2)     per a2  (:307)
template <typename T>
void phi_left_chain(const Shape& s, const T* phi1, const T* A, const T* x, const T* y, T* out,
                    hipStream_t st) {
  const int64_t sz1 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  T* R1 = (T*)arena(kArenaBfun, sizeof(T) * (size_t)(sz1 + sz2));
  T* R2 = R1 + sz1;
  T* work = nullptr;
  GemmOp<T> g1 = colmajor<T>('T', 'N', s.rx1 * s.ra1, s.n * s.ry2, s.ry1, T(1), phi1, s.ry1, y,
                             s.ry1, T(0), R1, (int64_t)s.rx1 * s.ra1);
  gemm_launch(g1, st);
  GemmOp<T> g2{};
  g2.M = s.rx1; g2.N = s.m * s.ra2; g2.K = s.ra1 * s.n;
  g2.A = R1; g2.sAi = 1; g2.sAk = s.rx1; g2.sAb0 = (int64_t)s.rx1 * s.ra1 * s.n;
  g2.B = A; g2.sBk = 1; g2.sBj = (int64_t)s.ra1 * s.n;
  g2.C = R2; g2.sCi = 1; g2.sCj = s.rx1; g2.sCb0 = (int64_t)s.rx1 * s.m * s.ra2;
  g2.nb0 = s.ry2; g2.nbatch = s.ry2; g2.alpha = T(1); g2.beta = T(0); g2.ksplit = 1;
  gemm_launch(g2, st);
  GemmOp<T> g3{};
  g3.M = s.ry2; g3.N = s.rx2; g3.K = s.rx1 * s.m;
  g3.A = R2; g3.sAi = (int64_t)s.rx1 * s.m * s.ra2; g3.sAk = 1; g3.sAb0 = (int64_t)s.rx1 * s.m;
  g3.B = x; g3.sBk = 1; g3.sBj = (int64_t)s.rx1 * s.m;
  g3.C = out; g3.sCi = 1; g3.sCj = s.ry2; g3.sCb0 = (int64_t)s.ry2 * s.rx2;
  g3.nb0 = s.ra2; g3.nbatch = s.ra2; g3.alpha = T(1); g3.beta = T(0); g3.ksplit = 1;
  g3.work = work;
  gemm_launch(g3, st);
}

// phi_right (tt_linalg.f90:259-284): stages 1-2 of bfun3, then
//  new(b1,a1,c1) = sum_{i,c2} T2(b1,a1,i,c2) y(c1,i,c2)                (:279)
template <typename T>
void phi_right_chain(const Shape& s, const T* phi2, const T* A, const T* x, const T* y, T* out,
                     hipStream_t st) {
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  T* T1 = (T*)arena(kArenaBfun, sizeof(T) * (size_t)(sz1 + sz2));
  T* T2 = T1 + sz1;
  bfun3_stage12(s, 1, A, phi2, x, (int64_t)s.rx1 * s.m * s.rx2, T1, T2, st);
  GemmOp<T> g3 = colmajor<T>('N', 'T', s.rx1 * s.ra1, s.ry1, s.n * s.ry2, T(1), T2,
                             (int64_t)s.rx1 * s.ra1, y, s.ry1, T(0), out, (int64_t)s.rx1 * s.ra1);
  gemm_launch(g3, st);
}

template void phi_left_chain<double>(const Shape&, const double*, const double*, const double*,
                                     const double*, double*, hipStream_t);
template void phi_right_chain<double>(const Shape&, const double*, const double*, const double*,
                                      const double*, double*, hipStream_t);

void zphi_left_chain(const Shape& s, const zc* phi1, const zc* A, const zc* x, const zc* y,
                     zc* out, hipStream_t st) {
  const int64_t sz1 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  zc* R1 = (zc*)arena(kArenaBfun, sizeof(zc) * (size_t)(sz1 + sz2));
  zc* R2 = R1 + sz1;
  ZGemmOp g1 = zop<double>();
  g1.M = s.rx1 * s.ra1; g1.N = s.n * s.ry2; g1.K = s.ry1;
  g1.A = phi1; g1.sAi = s.ry1; g1.sAk = 1;
  g1.B = y; g1.sBk = 1; g1.sBj = s.ry1; g1.conjB = 1;
  g1.C = R1; g1.sCi = 1; g1.sCj = (int64_t)s.rx1 * s.ra1;
  zgemm_launch(g1, st);
  ZGemmOp g2 = zop<double>();
  g2.M = s.rx1; g2.N = s.m * s.ra2; g2.K = s.ra1 * s.n;
  g2.A = R1; g2.sAi = 1; g2.sAk = s.rx1; g2.sAb0 = (int64_t)s.rx1 * s.ra1 * s.n;
  g2.B = A; g2.sBk = 1; g2.sBj = (int64_t)s.ra1 * s.n;
  g2.C = R2; g2.sCi = 1; g2.sCj = s.rx1; g2.sCb0 = (int64_t)s.rx1 * s.m * s.ra2;
  g2.nb0 = s.ry2; g2.nbatch = s.ry2;
  zgemm_launch(g2, st);
  ZGemmOp g3 = zop<double>();
  g3.M = s.ry2; g3.N = s.rx2; g3.K = s.rx1 * s.m;
  g3.A = R2; g3.sAi = (int64_t)s.rx1 * s.m * s.ra2; g3.sAk = 1; g3.sAb0 = (int64_t)s.rx1 * s.m;
  g3.B = x; g3.sBk = 1; g3.sBj = (int64_t)s.rx1 * s.m;
  g3.C = out; g3.sCi = 1; g3.sCj = s.ry2; g3.sCb0 = (int64_t)s.ry2 * s.rx2;
  g3.nb0 = s.ra2; g3.nbatch = s.ra2;
  zgemm_launch(g3, st);
}

void zphi_right_chain(const Shape& s, const zc* phi2, const zc* A, const zc* x, const zc* y,
                      zc* out, hipStream_t st) {
  const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
  const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
  zc* T1 = (zc*)arena(kArenaBfun, sizeof(zc) * (size_t)(sz1 + sz2));
  zc* T2 = T1 + sz1;
  ZGemmOp g1 = zop<double>();
  g1.M = s.rx1 * s.m; g1.N = s.ra2 * s.ry2; g1.K = s.rx2;
  g1.A = x; g1.sAi = 1; g1.sAk = (int64_t)s.rx1 * s.m;
  g1.B = phi2; g1.sBk = 1; g1.sBj = s.rx2;
  g1.C = T1; g1.sCi = 1; g1.sCj = (int64_t)s.rx1 * s.m;
  zgemm_launch(g1, st);
  ZGemmOp g2 = zop<double>();
  g2.M = s.rx1; g2.N = s.ra1 * s.n; g2.K = s.m * s.ra2;
  g2.A = T1; g2.sAi = 1; g2.sAk = s.rx1; g2.sAb0 = (int64_t)s.rx1 * s.m * s.ra2;
  g2.B = A; g2.sBk = (int64_t)s.ra1 * s.n; g2.sBj = 1;
  g2.C = T2; g2.sCi = 1; g2.sCj = s.rx1; g2.sCb0 = (int64_t)s.rx1 * s.ra1 * s.n;
  g2.nb0 = s.ry2; g2.nbatch = s.ry2;
  zgemm_launch(g2, st);
  ZGemmOp g3 = zop<double>();
  g3.M = s.rx1 * s.ra1; g3.N = s.ry1; g3.K = s.n * s.ry2;
  g3.A = T2; g3.sAi = 1; g3.sAk = (int64_t)s.rx1 * s.ra1;
  g3.B = y; g3.sBk = s.ry1; g3.sBj = 1; g3.conjB = 1;
  g3.C = out; g3.sCi = 1; g3.sCj = (int64_t)s.rx1 * s.ra1;
  zgemm_launch(g3, st);
}

// phi2_left (tt_linalg.f90:329-340) and phi2_right (:315-327).
template <typename T>
void phi2_left_chain(int rx1, int rx2, int ry1, int n, int ry2, const T* phi1, const T* x,
                     const T* y, T* out, hipStream_t st) {
  T* R1 = (T*)arena(kArenaBfun, sizeof(T) * (size_t)rx1 * n * ry2);
  gemm_launch(colmajor<T>('T', 'N', rx1, n * ry2, ry1, T(1), phi1, ry1, y, ry1, T(0), R1, rx1), st);
  gemm_launch(colmajor<T>('T', 'N', ry2, rx2, rx1 * n, T(1), R1, (int64_t)rx1 * n, x,
                          (int64_t)rx1 * n, T(0), out, ry2), st);
}

template <typename T>
void phi2_right_chain(int rx1, int rx2, int ry1, int n, int ry2, const T* phi2, const T* x,
                      const T* y, T* out, hipStream_t st) {
  T* R1 = (T*)arena(kArenaBfun, sizeof(T) * (size_t)rx1 * n * ry2);
  gemm_launch(colmajor<T>('N', 'N', rx1 * n, ry2, rx2, T(1), x, (int64_t)rx1 * n, phi2, rx2, T(0),
                          R1, (int64_t)rx1 * n), st);
  gemm_launch(colmajor<T>('N', 'T', rx1, ry1, n * ry2, T(1), R1, rx1, y, ry1, T(0), out, rx1), st);
}

template void phi2_left_chain<double>(int, int, int, int, int, const double*, const double*,
                                      const double*, double*, hipStream_t);
template void phi2_right_chain<double>(int, int, int, int, int, const double*, const double*,
                                       const double*, double*, hipStream_t);

// ------------------------------------------------------------------ full matrix
// B((c1,i,c2),(b1,j,b2)) = sum_{a1,a2} phi1(c1,b1,a1) A(a1,i,j,a2) phi2(b2,a2,c2)
// (tt_eigb.f90:88-105).  One thread per entry, rows fastest for coalesced
// stores; the ra1*ra2 sum is short.
__global__ void fullmat_kernel(int r1, int nn, int r2, int ra1, int ra2, const double* __restrict__ phi1,
                               const double* __restrict__ A, const double* __restrict__ phi2,
                               double* __restrict__ B) {
  const int64_t N = (int64_t)r1 * nn * r2;
  const int64_t total = N * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t row = t % N, col = t / N;
    int c1 = (int)(row % r1), i = (int)((row / r1) % nn), c2 = (int)(row / ((int64_t)r1 * nn));
    int b1 = (int)(col % r1), j = (int)((col / r1) % nn), b2 = (int)(col / ((int64_t)r1 * nn));
    double s = 0.0;
    for (int a1 = 0; a1 < ra1; ++a1) {
      double p1 = phi1[c1 + (int64_t)r1 * (b1 + (int64_t)r1 * a1)];
      if (p1 == 0.0) continue;
      for (int a2 = 0; a2 < ra2; ++a2) {
        double a = A[a1 + ra1 * (i + (int64_t)nn * (j + (int64_t)nn * a2))];
        double p2 = phi2[b2 + (int64_t)r2 * (a2 + (int64_t)ra2 * c2)];
        s += p1 * a * p2;
      }
    }
    B[t] = s;
  }
}

__global__ void localdiag_kernel(int r1, int nn, int r2, int ra1, int ra2, const double* __restrict__ phi1,
                                 const double* __restrict__ A, const double* __restrict__ phi2,
                                 double* __restrict__ D) {
  const int64_t N = (int64_t)r1 * nn * r2;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x) {
    const int c1 = (int)(t % r1), i = (int)((t / r1) % nn), c2 = (int)(t / ((int64_t)r1 * nn));
    double s = 0.0;
    for (int a1 = 0; a1 < ra1; ++a1) {
      const double p1 = phi1[c1 + (int64_t)r1 * (c1 + (int64_t)r1 * a1)];
      for (int a2 = 0; a2 < ra2; ++a2)
        s += p1 * A[a1 + ra1 * (i + (int64_t)nn * (i + (int64_t)nn * a2))] *
             phi2[c2 + (int64_t)r2 * (a2 + (int64_t)ra2 * c2)];
    }
    D[t] = s;
  }
}

void localdiag_launch(int r1, int nn, int r2, int ra1, int ra2, const double* phi1, const double* A,
                      const double* phi2, double* D, hipStream_t st) {
  const int64_t N = (int64_t)r1 * nn * r2;
  const int blocks = (int)std::min<int64_t>(ceil_div(N, 256), 4096);
  hipLaunchKernelGGL(localdiag_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, r1, nn, r2, ra1, ra2, phi1,
                     A, phi2, D);
  TTF_HIP(hipGetLastError());
}

void fullmat_launch(int r1, int nn, int r2, int ra1, int ra2, const double* phi1, const double* A,
                    const double* phi2, double* B, hipStream_t st) {
  const int64_t N = (int64_t)r1 * nn * r2;
  int blocks = (int)std::min<int64_t>(ceil_div(N * N, 256), 8192);
  hipLaunchKernelGGL(fullmat_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, r1, nn, r2, ra1,
                     ra2, phi1, A, phi2, B);
  TTF_HIP(hipGetLastError());
}

}  // namespace ttf

// ------------------------------------------------------------------ C ABI
using namespace ttf;

static Shape mkshape(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2) {
  Shape s{rx1, m, rx2, ry1, n, ry2, ra1, ra2};
  TTF_REQUIRE(rx1 > 0 && m > 0 && rx2 > 0 && ry1 > 0 && n > 0 && ry2 > 0 && ra1 > 0 && ra2 > 0,
              "invalid sizes rx1=%d m=%d rx2=%d ry1=%d n=%d ry2=%d ra1=%d ra2=%d", rx1, m, rx2, ry1,
              n, ry2, ra1, ra2);
  return s;
}

extern "C" {

int ttf_dbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
               const double* phi1, const double* A, const double* phi2, const double* x,
               int64_t ldx, double* y, int64_t ldy) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    TTF_REQUIRE(ldx >= (int64_t)rx1 * m * rx2 || nvec <= 1, "ttf_dbfun3: ldx too small");
    TTF_REQUIRE(ldy >= (int64_t)ry1 * n * ry2 || nvec <= 1, "ttf_dbfun3: ldy too small");
    bfun3_chain<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, stream());
  });
}

int ttf_sbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
               const float* phi1, const float* A, const float* phi2, const float* x, int64_t ldx,
               float* y, int64_t ldy) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    TTF_REQUIRE(ldx >= (int64_t)rx1 * m * rx2 || nvec <= 1, "ttf_sbfun3: ldx too small");
    TTF_REQUIRE(ldy >= (int64_t)ry1 * n * ry2 || nvec <= 1, "ttf_sbfun3: ldy too small");
    bfun3_chain<float>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, stream());
  });
}

int ttf_zbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
               const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* phi2,
               const ttf_zdouble* x, int64_t ldx, ttf_zdouble* y, int64_t ldy) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    zbfun3_chain(s, nvec, (const zc*)phi1, (const zc*)A, (const zc*)phi2, (const zc*)x, ldx,
                 (zc*)y, ldy, stream());
  });
}

int ttf_zbfun3_transp(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
                      const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* phi2,
                      const ttf_zdouble* x, int64_t ldx, ttf_zdouble* y, int64_t ldy) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    zbfun3_transp_chain(s, nvec, (const zc*)phi1, (const zc*)A, (const zc*)phi2, (const zc*)x, ldx,
                        (zc*)y, ldy, stream());
  });
}

int ttf_dphi_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                  const double* phi1, const double* A, const double* x, const double* y,
                  double* out) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    phi_left_chain<double>(s, phi1, A, x, y, out, stream());
  });
}

int ttf_dphi_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                   const double* phi2, const double* A, const double* x, const double* y,
                   double* out) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    phi_right_chain<double>(s, phi2, A, x, y, out, stream());
  });
}

int ttf_zphi_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                  const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* x,
                  const ttf_zdouble* y, ttf_zdouble* out) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    zphi_left_chain(s, (const zc*)phi1, (const zc*)A, (const zc*)x, (const zc*)y, (zc*)out,
                    stream());
  });
}

int ttf_zphi_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                   const ttf_zdouble* phi2, const ttf_zdouble* A, const ttf_zdouble* x,
                   const ttf_zdouble* y, ttf_zdouble* out) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    zphi_right_chain(s, (const zc*)phi2, (const zc*)A, (const zc*)x, (const zc*)y, (zc*)out,
                     stream());
  });
}

int ttf_dphi2_left(int rx1, int rx2, int ry1, int n, int ry2, const double* phi1, const double* x,
                   const double* y, double* out) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && rx2 > 0 && ry1 > 0 && n > 0 && ry2 > 0, "ttf_dphi2_left: invalid sizes");
    phi2_left_chain<double>(rx1, rx2, ry1, n, ry2, phi1, x, y, out, stream());
  });
}

int ttf_dphi2_right(int rx1, int rx2, int ry1, int n, int ry2, const double* phi2,
                    const double* x, const double* y, double* out) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && rx2 > 0 && ry1 > 0 && n > 0 && ry2 > 0, "ttf_dphi2_right: invalid sizes");
    phi2_right_chain<double>(rx1, rx2, ry1, n, ry2, phi2, x, y, out, stream());
  });
}

int ttf_dbfun3_profile(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                       int nvec, const double* phi1, const double* A, const double* phi2,
                       const double* x, int64_t ldx, double* y, int64_t ldy, int iters,
                       float* stage_ms) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    TTF_REQUIRE(iters > 0 && nvec > 0 && stage_ms != nullptr, "ttf_dbfun3_profile: bad arguments");
    hipStream_t st = stream();
    const int64_t sz1 = (int64_t)s.rx1 * s.m * s.ra2 * s.ry2;
    const int64_t sz2 = (int64_t)s.rx1 * s.ra1 * s.n * s.ry2;
    double* T1 = (double*)arena(kArenaBfun, sizeof(double) * (size_t)(sz1 + sz2) * nvec);
    double* T2 = T1 + sz1 * nvec;
    const GemmOp<double> ops[3] = {bfun3_op1(s, nvec, phi2, x, ldx, T1),
                                   bfun3_op2(s, nvec, A, (const double*)T1, T2),
                                   bfun3_op3(s, nvec, phi1, (const double*)T2, y, ldy)};
    hipEvent_t e0, e1;
    TTF_HIP(hipEventCreate(&e0));
    TTF_HIP(hipEventCreate(&e1));
    // each stage is launched `iters` times back to back between two events
    // (the stages are idempotent), after one launch that settles its tuning
    for (int k = 0; k < 3; ++k) {
      gemm_launch(ops[k], st);
      TTF_HIP(hipEventRecord(e0, st));
      for (int it = 0; it < iters; ++it) gemm_launch(ops[k], st);
      TTF_HIP(hipEventRecord(e1, st));
      TTF_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      TTF_HIP(hipEventElapsedTime(&t, e0, e1));
      stage_ms[k] = t / (float)iters;
    }
    TTF_HIP(hipEventDestroy(e0));
    TTF_HIP(hipEventDestroy(e1));
  });
}

int ttf_dbfun3_chain_profile(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
                             const double* phi1, const double* A, const double* phi2, const double* x,
                             int64_t ldx, double* y, int64_t ldy, int iters, float* ms, int* fused) {
  return guarded([&] {
    Shape s = mkshape(rx1, m, rx2, ry1, n, ry2, ra1, ra2);
    TTF_REQUIRE(iters > 0 && nvec > 0 && ms != nullptr, "ttf_dbfun3_chain_profile: bad arguments");
    hipStream_t st = stream();
    if (fused) *fused = bfun3_fused_fits(s, nvec) && !overlaps(x, ldx, (const double*)y, ldy, nvec) ? 1 : 0;
    bfun3_chain<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st);   // settles any tuning
    hipEvent_t e0, e1;
    TTF_HIP(hipEventCreate(&e0));
    TTF_HIP(hipEventCreate(&e1));
    TTF_HIP(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it) bfun3_chain<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st);
    TTF_HIP(hipEventRecord(e1, st));
    TTF_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    TTF_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / (float)iters;
    TTF_HIP(hipEventDestroy(e0));
    TTF_HIP(hipEventDestroy(e1));
  });
}

int ttf_d2d_fullmat(int rx1, int m, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                    const double* phi2, double* B) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && m > 0 && rx2 > 0 && ra1 > 0 && ra2 > 0, "ttf_d2d_fullmat: invalid sizes");
    fullmat_launch(rx1, m, rx2, ra1, ra2, phi1, A, phi2, B, stream());
  });
}

}  // extern "C"

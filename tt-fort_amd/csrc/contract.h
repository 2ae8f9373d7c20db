// Contraction chains of the local operator and interface updates
// (implemented in contract.hip) for use by the solvers and drivers.
#pragma once

#include "gemm_mfma.h"

namespace ttf {

// Sizes of one local problem: x (rx1,m,rx2), y (ry1,n,ry2), A (ra1,n,m,ra2).
struct Shape {
  int rx1, m, rx2, ry1, n, ry2, ra1, ra2;
};

// Real flops of one bfun3 application to one vector (the three GEMM stages);
// a complex application costs 4 times this.
inline double bfun3_flops(const Shape& s) {
  return 2.0 * s.rx1 * s.m * s.rx2 * s.ra2 * s.ry2 + 2.0 * s.ra1 * s.n * s.m * s.ra2 * s.ry2 * s.rx1 +
         2.0 * s.ry1 * s.n * s.ry2 * s.rx1 * s.ra1;
}

// The whole FP64 chain in one launch (bfun3_fused.hip) when the shape fits
// its tiles (n <= 16, ry1 <= 128, ra1 n <= 16, m ra2 <= 64,
// ra2 (16 / n) <= 64, rx1 m <= 256, LDS within 160 KB) and TTF_BFUN3_FUSED
// is not 0; false leaves y untouched and the caller runs the GEMM chain.
// x and y must not overlap.
bool bfun3_fused(const Shape& s, int nvec, const double* phi1, const double* A, const double* phi2,
                 const double* x, int64_t ldx, double* y, int64_t ldy, hipStream_t st);
bool bfun3_fused_fits(const Shape& s, int nvec);

template <typename T>
void bfun3_chain(const Shape& s, int nvec, const T* phi1, const T* A, const T* phi2, const T* x,
                 int64_t ldx, T* y, int64_t ldy, hipStream_t st);
// Elements of workspace bfun3_chain_ws needs for nvec vectors: the two
// stage intermediates and the largest stage's split-K slices.
size_t bfun3_workspace(const Shape& s, int nvec);
// bfun3_chain with every intermediate in ws (ws_cap elements, an error if
// too small): no allocation and no tuning runs; the same bits as
// bfun3_chain.
template <typename T>
void bfun3_chain_ws(const Shape& s, int nvec, const T* phi1, const T* A, const T* phi2, const T* x,
                    int64_t ldx, T* y, int64_t ldy, T* ws, size_t ws_cap, hipStream_t st);
template <typename T>
void bfun3_stage12(const Shape& s, int nvec, const T* A, const T* phi2, const T* x, int64_t ldx,
                   T* T1, T* T2, hipStream_t st);

void zbfun3_chain(const Shape& s, int nvec, const hipDoubleComplex* phi1, const hipDoubleComplex* A,
                  const hipDoubleComplex* phi2, const hipDoubleComplex* x, int64_t ldx,
                  hipDoubleComplex* y, int64_t ldy, hipStream_t st);
// The same chain in complex64 (complex64 operands and intermediates, fp32
// matrix cores with fp32 accumulation).
void cbfun3_chain(const Shape& s, int nvec, const hipFloatComplex* phi1, const hipFloatComplex* A,
                  const hipFloatComplex* phi2, const hipFloatComplex* x, int64_t ldx, hipFloatComplex* y,
                  int64_t ldy, hipStream_t st);
void zbfun3_transp_chain(const Shape& s, int nvec, const hipDoubleComplex* phi1,
                         const hipDoubleComplex* A, const hipDoubleComplex* phi2,
                         const hipDoubleComplex* x, int64_t ldx, hipDoubleComplex* y, int64_t ldy,
                         hipStream_t st);

template <typename T>
void phi_left_chain(const Shape& s, const T* phi1, const T* A, const T* x, const T* y, T* out,
                    hipStream_t st);
template <typename T>
void phi_right_chain(const Shape& s, const T* phi2, const T* A, const T* x, const T* y, T* out,
                     hipStream_t st);
void zphi_left_chain(const Shape& s, const hipDoubleComplex* phi1, const hipDoubleComplex* A,
                     const hipDoubleComplex* x, const hipDoubleComplex* y, hipDoubleComplex* out,
                     hipStream_t st);
void zphi_right_chain(const Shape& s, const hipDoubleComplex* phi2, const hipDoubleComplex* A,
                      const hipDoubleComplex* x, const hipDoubleComplex* y, hipDoubleComplex* out,
                      hipStream_t st);
template <typename T>
void phi2_left_chain(int rx1, int rx2, int ry1, int n, int ry2, const T* phi1, const T* x,
                     const T* y, T* out, hipStream_t st);
template <typename T>
void phi2_right_chain(int rx1, int rx2, int ry1, int n, int ry2, const T* phi2, const T* x,
                      const T* y, T* out, hipStream_t st);

void fullmat_launch(int r1, int nn, int r2, int ra1, int ra2, const double* phi1, const double* A,
                    const double* phi2, double* B, hipStream_t st);
// Diagonal of the same local matrix: D[(c1, i, c2)] = sum over (a1, a2) of
// phi1[c1, c1, a1] A[a1, i, i, a2] phi2[c2, a2, c2] (N = r1 nn r2 values).
void localdiag_launch(int r1, int nn, int r2, int ra1, int ra2, const double* phi1, const double* A,
                      const double* phi2, double* D, hipStream_t st);

}  // namespace ttf

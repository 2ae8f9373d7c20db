// Dense FP64 linear algebra on the device for the sweep drivers: small
// symmetric eigenproblems (Rayleigh-Ritz, basis orthonormalisation), tall-
// skinny Gram products, and the thin SVD behind the rank truncation.
#pragma once

#include "ttf_common.h"

#include <vector>

namespace ttf {

// G (p x q, leading dim ldg) = X^T Y for X (N x p, ldx), Y (N x q, ldy);
// p, q <= 64.  Deterministic two-pass reduction over row chunks.
void tsgram(int N, int p, int q, const double* X, int64_t ldx, const double* Y, int64_t ldy,
            double* G, int ldg, hipStream_t st);
// The same for X = [X1 X2] (p1 + p2 columns from two matrices).
void tsgram2(int N, int p1, const double* X1, int64_t ldx1, int p2, const double* X2, int64_t ldx2, int q,
             const double* Y, int64_t ldy, double* G, int ldg, hipStream_t st);

// Eigen-decomposition of `count` symmetric n x n matrices (n <= 64), stored
// consecutively with leading dimension n at A (overwritten).  Eigenvalues go
// to w (ascending, or descending with desc = 1), eigenvectors to Q (n x n
// each).  Cyclic Jacobi in LDS, one work-group per matrix.
void syev_small(int n, int count, double* A, double* w, double* Q, int desc, hipStream_t st);

// Thin SVD of an m x n column-major matrix A (lda): A = U diag(S) VT with
// k = min(m, n) values in descending order; U (m x k, ldu), VT (k x n, ldvt).
// Blocked one-sided Jacobi with accumulated rotations, singular values from
// the final column norms.  Singular vectors belonging to zero singular
// values (norm at the rounding floor sqrt(mw) eps ||A||_F) are completed to
// an orthonormal set when `complete` is set, else left zero (enough whenever
// they are only used multiplied by their singular value).  Returns the
// number of Jacobi sweeps used.
int gesvd(int m, int n, const double* A, int64_t lda, double* U, int64_t ldu, double* S,
          double* VT, int64_t ldvt, hipStream_t st, int complete = 1);

// Singular values (descending: device S and host hs, k = min(m, n)) and the
// orthonormal singular vectors of ONE side of A: left (m x k) when
// want_left, else right (n x k), as the columns of O (ldo).  The other
// factor is the caller's product with A (O^T A or A O).  When the wanted
// side is the shorter dimension the Jacobi iteration runs on R^T from a
// Householder QR of the tall orientation with its columns in decreasing norm
// order (the preconditioning of Drmac-Veselic, the pivoting reduced to one
// sort), which keeps the sweep count low for graded spectra and needs no
// accumulated rotations.  Vectors of zero singular values are zero
// columns; *kpos counts the nonzero ones.  Returns the Jacobi sweeps used.
int svd_vectors(int m, int n, const double* A, int64_t lda, bool want_left, double* S, double* O,
                int64_t ldo, std::vector<double>& hs, int* kpos, hipStream_t st);

// Orthogonal factorisation A (m x n, ld m) = Q R with Q (m x k) orthonormal
// and R = Q^T A (k x n), k = min(m, n).  Q is the left singular basis,
// completed where A is rank deficient; any orthonormal Q with Q R = A serves
// the sweeps (the reference's dqr, tt_linalg.f90:30-51, is used only as
// such).  S: k device doubles of workspace.
void orth_factor(int m, int n, const double* A, double* Q, double* R, double* S, hipStream_t st);

// Householder QR with LAPACK's reflector convention (dgeqrf + dorgqr, the
// reference's dqr, tt_linalg.f90:30-51): A (n x m, ld n) = Q R with Q
// (n x k, ld n) orthonormal and R (k x m, ld ldr) upper trapezoidal,
// k = min(n, m).  Panels of up to 32 columns in LDS (n < 16384 rows), the
// trailing columns and the explicit Q formed by compact-WY GEMMs.
void householder_qr(int n, int m, const double* A, double* Q, double* R, int64_t ldr, hipStream_t st);

// Completes columns [k0, k) of U (m x k, ldu) to an orthonormal set.
void complete_columns(int m, int k0, int k, double* U, int64_t ldu, hipStream_t st);

// B (n x m) = A (m x n)^T, both column-major with leading dims.
void transpose(int m, int n, const double* A, int64_t lda, double* B, int64_t ldb, hipStream_t st);

}  // namespace ttf

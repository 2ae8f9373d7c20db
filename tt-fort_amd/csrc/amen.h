// Local building blocks of the AMEn linear solver (amen_local.hip) and the
// device sweep driver (amen.hip).
#pragma once

#include "solvers.h"

#include <functional>

namespace ttf {

// 'N'/'C'/'L'/'R' -> lower case (the reference accepts either case).
char norm_ptype(char p);

// Block-Jacobi preconditioner of one local step (tt_linalg.f90:363-533):
// 'n' none, 'c' rx1*rx2 blocks of n x n (layout (n, n, rx2, rx1)), 'l' rx2
// blocks of (rx1 n)^2, 'r' rx1 blocks of (n rx2)^2.  Blocks up to 128.
struct JacPrec {
  char type = 'n';
  int rx1 = 0, n = 0, rx2 = 0;
  const double* jacs = nullptr;
  // the local operator the blocks came from (for recording training pairs
  // of the learned preconditioner; may be left null)
  int ra1 = 0, ra2 = 0;
  const double* phi1 = nullptr;
  const double* A = nullptr;
  const double* phi2 = nullptr;
};
size_t jac_size(char type, int rx1, int n, int rx2);   // doubles
void jac_gen(char type, int rx1, int n, int rx2, int ra1, int ra2, const double* Phi1, const double* A,
             const double* Phi2, double* jacs, hipStream_t st);
void jac_apply(char type, int rx1, int n, int rx2, const double* jacs, const double* x, double* y,
               hipStream_t st);

// In-place inverse of one n x n column-major matrix (n <= 128), Gauss-Jordan
// with partial pivoting (the dense local solve of ttamen.f90:138-146); a
// singular matrix is an error.
void dense_inverse(int n, double* M, hipStream_t st);

// In-place inverse of nblk symmetric positive definite nb x nb blocks
// (column-major, consecutive; nb <= 256): Gauss-Jordan without pivoting for
// nb <= 128, else a 2 x 2 split through the Schur complement of the leading
// half.  Returns false (the blocks then hold no inverse) when a pivot is not
// positive, i.e. when some block is not positive definite.
bool spd_block_inverse(int nb, int64_t nblk, double* M, hipStream_t st);

struct GmresStats {
  int it = 0;         // restart cycles completed (0-based index of the last)
  int j = 0;          // Krylov dimension of the last cycle
  double relres = 0;  // |residual| / |rhs| estimate at exit
  int matvecs = 0;
};
// Solves B P sol = rhs (right preconditioning; the caller applies P to sol,
// as tt_adapt_als.f90:776-781 does) by restarted GMRES(nrestart), at most
// niters cycles, stopping when the residual estimate falls below tol times
// the initial residual norm (tt_linalg.f90:797-803).  sol starts at zero.
// With `mix`, the Krylov directions are multiplied by the fp32 operator and
// each restart residual by the FP64 one (iterative refinement).
GmresStats gmres_solve(const LocalOp& op, const JacPrec& P, const double* rhs, double* sol, int nrestart,
                       double tol, int niters, hipStream_t st, const MixedOp* mix = nullptr);

// Flexible GMRES (right preconditioning by a possibly nonlinear map M, with
// the preconditioned directions kept: x = sum_j y_j M(v_j)).  Solves
// B x = rhs from x = 0 with the same restart / stopping rules as
// gmres_solve; x is the solution itself.
typedef std::function<void(const double* v, double* z)> PrecFn;
GmresStats gmres_flex(const LocalOp& op, const PrecFn& M, const double* rhs, double* x, int nrestart, double tol,
                      int niters, hipStream_t st, const MixedOp* mix = nullptr);

// Pivoted incomplete Cholesky of the Gram matrix of the n "items" of dA
// (columns of the m x n dA for trans 'N', rows of the n x m dA for 'T'),
// returning up to numvec orthonormal directions in du (n x numvec) and the
// count produced (tt_linalg.f90:540-643).  lam (optional, numvec) receives
// the final eigenvalues.
int duchol(char trans, int m, int n, const double* dA, int numvec, double* du, double* lam, hipStream_t st);

// The first two stages of the local operator: y(b1, a1, i, c2), i.e. the
// (rx1 ra1) x (n ry2) block of A Phi2 x (tt_linalg.f90:210-240).
void bfun3_right_half(const Shape& s, const double* A, const double* phi2, const double* x, double* y,
                      hipStream_t st);
// Phi1 A x with the right interface open: y(c1, i, b2, a2), the
// (ry1 n) x (rx2 ra2) block (the contraction tt_linalg.f90:242-256 is meant
// to compute).
void bfun3_left_half(const Shape& s, const double* phi1, const double* A, const double* x, double* y,
                     hipStream_t st);

}  // namespace ttf

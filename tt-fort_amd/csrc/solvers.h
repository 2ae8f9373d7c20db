// Device-resident local solvers used by the sweep drivers.
#pragma once

#include "contract.h"
#include "linalg.h"

namespace ttf {

// The projected operator of one local step (device pointers).
struct LocalOp {
  Shape s;
  const double* phi1;
  const double* A;
  const double* phi2;
  int N() const { return s.ry1 * s.n * s.ry2; }
  void apply(int nvec, const double* x, int64_t ldx, double* y, int64_t ldy, hipStream_t st) const {
    bfun3_chain<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, st);
  }
};

// FP32 copy of a local operator for the mixed-precision path: the vector
// is rounded to float, the three-stage chain runs on the fp32 matrix cores
// with fp32 accumulation, and the result is widened back to double.  The
// iterative solvers that use it keep their bases, projections and restart
// residuals in FP64 (the restart residual rhs - B x is the FP64 refinement
// step).  Buffers live in a dedicated arena: one MixedOp at a time.
struct MixedOp {
  Shape s;
  int nvec = 1;   // vectors the staging buffers hold
  float* phi1 = nullptr;
  float* A = nullptr;
  float* phi2 = nullptr;
  float* x = nullptr;
  float* y = nullptr;
  void apply(const double* xin, double* yout, hipStream_t st) const;
  // k <= nvec vectors at once (columns of x and y with leading dimensions)
  void apply_block(int k, const double* xin, int64_t ldx, double* yout, int64_t ldy, hipStream_t st) const;
};
MixedOp make_mixed(const LocalOp& op, hipStream_t st, int nvec = 1);

struct EigStats {
  int iters = 0;
  int matvecs = 0;       // local matvecs (vectors), all precisions
  int matvecs32 = 0;     // of which on the fp32 operator
  int refreshes = 0;     // FP64 recomputations of A X (mixed precision)
  double res = 0.0;      // 2-norm of the residual norms of the wanted pairs
  double anorm = 0.0;    // operator norm estimate (largest |Ritz value|)
  int dense = 0;         // 1 if the dense path was used
  int block_prec = 0;    // 1 shifted block preconditioner, -1 tried and not
                         // positive definite (diagonal used), 0 not tried
};

// Lowest `nev` eigenpairs of the symmetric local operator.  X (N x nev, ld N,
// device) holds the initial guess on entry and the eigenvectors on exit;
// lambda (host, nev) receives the eigenvalues in ascending order.
// Stopping rule: all residual norms <= tol * anorm (the rule of the
// reference's PRIMME call, eps = eps2/resid_damp), or max_matvecs reached.
// `guard` extra block vectors accelerate convergence of the wanted pairs.
// With `mix` (an fp32 copy of op holding at least nev + guard vectors), the
// search directions are multiplied on the fp32 matrix cores while the
// residuals are above 1e-6 of the operator norm, with A X and A P
// recomputed in FP64 (and a Rayleigh-Ritz step on X) whenever the
// maintained residuals meet the tolerance, fall 1000-fold or reach that
// level; below it the solve finishes on the FP64 operator.  Every stopping
// decision is taken on FP64 residuals of FP64 products.
EigStats block_eig(const LocalOp& op, int nev, int guard, double* X, double* lambda, double tol,
                   int max_matvecs, int dense_max, hipStream_t st, const MixedOp* mix = nullptr);

}  // namespace ttf

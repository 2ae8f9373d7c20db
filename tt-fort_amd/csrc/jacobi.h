// Two-sided cyclic Jacobi for small symmetric matrices held in LDS, shared by
// the small eigensolver, SVQB orthonormalisation and Rayleigh-Ritz kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ttf {

// Scratch in LDS for one Jacobi call (n <= 2 * kJacobiMaxHalf).
constexpr int kJacobiMaxHalf = 32;
struct JacobiScratch {
  double c[kJacobiMaxHalf], s[kJacobiMaxHalf];
  int p[kJacobiMaxHalf], q[kJacobiMaxHalf];
  int flag;
};

// Diagonalises the symmetric n x n matrix A (row stride LD) in place and
// accumulates the rotations into V (V <- V J).  Disjoint pairs of one round
// of the circle schedule (over an even index count) are rotated together:
// first the columns of A and V, then the rows of A, after which the
// annihilated pair entries are set to zero exactly.  A pair is rotated when
// |a_pq| exceeds both tol_abs (the caller's rounding floor, normally
// eps * ||A||_F over the meaningful part of A) and eps * sqrt(|a_pp a_qq|);
// the loop ends after a sweep without rotation or after max_sweeps.
// All threads of the block call it; A and V must be visible to all threads.
template <int LD>
__device__ void jacobi_sym(double (*A)[LD], double (*V)[LD], int n, int max_sweeps, double tol_abs,
                           JacobiScratch* js) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n + (n & 1);
  const int half = nn / 2;
  const double eps = 2.220446049250313e-16;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) js->flag = 0;
    __syncthreads();
    for (int r = 0; r < nn - 1; ++r) {
      for (int k = tid; k < half; k += nt) {
        int a, b;
        if (k == 0) { a = r; b = nn - 1; }
        else { a = (r + k) % (nn - 1); b = (r - k + (nn - 1)) % (nn - 1); }
        const int p = min(a, b), q = max(a, b);
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p][q], app = A[p][p], aqq = A[q][q];
          if (fabs(apq) > tol_abs && fabs(apq) > eps * sqrt(fabs(app) * fabs(aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            js->flag = 1;
          }
        }
        js->p[k] = p; js->q[k] = q; js->c[k] = c; js->s[k] = s;
      }
      __syncthreads();
      for (int e = tid; e < half * n; e += nt) {
        const int k = e / n, i = e % n;
        const int p = js->p[k], q = js->q[k];
        if (q >= n || js->s[k] == 0.0) continue;
        const double c = js->c[k], s = js->s[k];
        const double aip = A[i][p], aiq = A[i][q];
        A[i][p] = c * aip - s * aiq;
        A[i][q] = s * aip + c * aiq;
        const double vip = V[i][p], viq = V[i][q];
        V[i][p] = c * vip - s * viq;
        V[i][q] = s * vip + c * viq;
      }
      __syncthreads();
      for (int e = tid; e < half * n; e += nt) {
        const int k = e / n, j = e % n;
        const int p = js->p[k], q = js->q[k];
        if (q >= n || js->s[k] == 0.0) continue;
        const double c = js->c[k], s = js->s[k];
        const double apj = A[p][j], aqj = A[q][j];
        A[p][j] = c * apj - s * aqj;
        A[q][j] = s * apj + c * aqj;
      }
      __syncthreads();
      for (int k = tid; k < half; k += nt) {
        const int p = js->p[k], q = js->q[k];
        if (q < n && js->s[k] != 0.0) { A[p][q] = 0.0; A[q][p] = 0.0; }
      }
      __syncthreads();
    }
    const int again = js->flag;
    __syncthreads();
    if (!again) break;
  }
}

// Block-wide sum of v over all threads (blockDim.x a multiple of 64, at most
// 1024); `red` holds one double per wave.  Every thread receives the sum.
__device__ inline double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

}  // namespace ttf

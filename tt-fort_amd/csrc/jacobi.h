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
// accumulates the rotations into V (V <- V J).  Each round of the circle
// schedule (over an even index count nn) rotates its nn/2 disjoint pairs
// together: one phase computes the rotations, one phase applies J^T A J by
// 2x2 blocks (rows of pair P x columns of pair Q, each block read and
// written by one thread, so the update is in place) and V J by row x pair
// slices; annihilated pair entries are set to zero exactly.  A pair is
// rotated when |a_pq| exceeds both tol_abs (the caller's rounding floor,
// normally eps * ||A||_F over the meaningful part of A) and
// eps * sqrt(|a_pp a_qq|); the loop ends after a sweep without rotation or
// after max_sweeps.  All threads of the block call it.
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
          if (fabs(apq) > tol_abs && apq * apq > eps * eps * fabs(app * aqq)) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = copysign(1.0, tau) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = rsqrt(1.0 + t * t);
            s = t * c;
            js->flag = 1;
          }
        }
        js->p[k] = p; js->q[k] = q; js->c[k] = c; js->s[k] = s;
      }
      __syncthreads();
      // A <- J^T A J by 2x2 blocks (P, Q); then V <- V J by (row, Q) slices
      const int nblk = half * half;
      for (int e = tid; e < nblk + n * half; e += nt) {
        if (e < nblk) {
          const int kp = e % half, kq = e / half;
          const double c2 = js->c[kq], s2 = js->s[kq], c1 = js->c[kp], s1 = js->s[kp];
          if (s1 == 0.0 && s2 == 0.0) continue;
          const int p = js->p[kp], q = js->q[kp], u = js->p[kq], v = js->q[kq];
          const bool hq = q < n, hv = v < n;
          double apu = A[p][u], apv = hv ? A[p][v] : 0.0;
          double aqu = hq ? A[q][u] : 0.0, aqv = (hq && hv) ? A[q][v] : 0.0;
          // columns u, v
          double t0 = c2 * apu - s2 * apv, t1 = s2 * apu + c2 * apv;
          double t2 = c2 * aqu - s2 * aqv, t3 = s2 * aqu + c2 * aqv;
          // rows p, q
          apu = c1 * t0 - s1 * t2; aqu = s1 * t0 + c1 * t2;
          apv = c1 * t1 - s1 * t3; aqv = s1 * t1 + c1 * t3;
          if (kp == kq && s1 != 0.0) { apv = 0.0; aqu = 0.0; }
          A[p][u] = apu;
          if (hv) A[p][v] = apv;
          if (hq) A[q][u] = aqu;
          if (hq && hv) A[q][v] = aqv;
        } else {
          const int e2 = e - nblk;
          const int i = e2 % n, kq = e2 / n;
          const double c = js->c[kq], s = js->s[kq];
          if (s == 0.0) continue;
          const int u = js->p[kq], v = js->q[kq];
          const double x = V[i][u], y = V[i][v];
          V[i][u] = c * x - s * y;
          V[i][v] = s * x + c * y;
        }
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

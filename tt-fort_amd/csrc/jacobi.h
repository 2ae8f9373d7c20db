// Two-sided cyclic Jacobi for small symmetric matrices held in LDS, shared by
// the small eigensolver, SVQB orthonormalisation and Rayleigh-Ritz kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ttf {

constexpr int kJacobiMaxN = 64;   // largest order (rows are indexed by lanes)

// Pair k of round r of the circle schedule over nn (even) indices: index
// nn-1 stays fixed and the others rotate, so every pair meets once in the
// nn-1 rounds of a sweep.
__device__ inline void jacobi_pair(int r, int k, int nn, int& p, int& q) {
  int a, b;
  if (k == 0) {
    a = r;
    b = nn - 1;
  } else {
    a = r + k;
    if (a >= nn - 1) a -= nn - 1;
    b = r - k;
    if (b < 0) b += nn - 1;
  }
  p = min(a, b);
  q = max(a, b);
}

// Rotation (c, s) that annihilates A[p][q] (J = [[c, s], [-s, c]] on rows
// and columns p, q), or the identity when the pair is already decoupled:
// |a_pq| at or below the caller's rounding floor tol_abs or below
// eps * sqrt(|a_pp a_qq|).  t = tan(theta) is the smaller root of
// t^2 + 2 cot(2 theta) t - 1 = 0, written as
// t = sign(d) 2 a_pq / (|d| + sqrt(d^2 + 4 a_pq^2)) with d = a_qq - a_pp.
template <int LD>
__device__ inline bool jacobi_rotation(const double (*A)[LD], int p, int q, int n, double tol_abs,
                                       double& c, double& s) {
  c = 1.0;
  s = 0.0;
  if (q >= n) return false;
  const double eps = 2.220446049250313e-16;
  const double apq = A[p][q], app = A[p][p], aqq = A[q][q];
  if (!(fabs(apq) > tol_abs && apq * apq > eps * eps * fabs(app * aqq))) return false;
  const double d = aqq - app;
  const double t = copysign(1.0, d) * (2.0 * apq) / (fabs(d) + sqrt(d * d + 4.0 * apq * apq));
  c = rsqrt(1.0 + t * t);
  s = t * c;
  return true;
}

// Diagonalises the symmetric n x n matrix held in A (row stride LD, n <= 64)
// and accumulates the rotations into V (V <- V J).  B is a second n x n
// buffer: each round of the circle schedule reads one buffer and writes the
// other, so one barrier per round suffices.  Every thread owns 2x2 blocks
// (rows of pair kp x columns of pair kq) and derives both rotations from
// the round's input itself, so no phase publishes rotations; the thread also
// rotates rows kp and kp + nn/2 of V for pair kq.  Annihilated entries are
// set to zero exactly.  The loop ends after a sweep without rotation or
// after max_sweeps.  `flags` is 3 ints of LDS.  All threads of the block
// call it; the returned pointer (A or B) holds the diagonalised matrix.
template <int LD>
__device__ double (*jacobi_sym(double (*A)[LD], double (*B)[LD], double (*V)[LD], int n, int max_sweeps,
                               double tol_abs, int* flags))[LD] {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n + (n & 1);
  const int half = nn / 2;
  if (n <= 1) return A;
  // thread -> (kp, kq0); threads with kq0 >= kqstep would repeat columns
  const int kp = tid % half, kq0 = tid / half, kqstep = max(nt / half, 1);
  double(*cur)[LD] = A;
  double(*nxt)[LD] = B;
  if (tid < 3) flags[tid] = 0;
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    int* rotated = flags + sweep % 3;
    // flags[(sweep+1) % 3] was last read before the previous sweep's barriers
    if (tid == 0) flags[(sweep + 1) % 3] = 0;
    for (int r = 0; r < nn - 1; ++r) {
      if (kq0 < kqstep && kq0 < half) {
        int p, q;
        jacobi_pair(r, kp, nn, p, q);
        double c1, s1;
        const bool rot1 = jacobi_rotation<LD>(cur, p, q, n, tol_abs, c1, s1);
        if (rot1) *rotated = 1;
        const bool hq = q < n;
        for (int kq = kq0; kq < half; kq += kqstep) {
          int u, v;
          jacobi_pair(r, kq, nn, u, v);
          double c2, s2;
          jacobi_rotation<LD>(cur, u, v, n, tol_abs, c2, s2);
          const bool hv = v < n;
          double apu = cur[p][u], apv = hv ? cur[p][v] : 0.0;
          double aqu = hq ? cur[q][u] : 0.0, aqv = (hq && hv) ? cur[q][v] : 0.0;
          // columns u, v
          const double t0 = c2 * apu - s2 * apv, t1 = s2 * apu + c2 * apv;
          const double t2 = c2 * aqu - s2 * aqv, t3 = s2 * aqu + c2 * aqv;
          // rows p, q
          apu = c1 * t0 - s1 * t2;
          aqu = s1 * t0 + c1 * t2;
          apv = c1 * t1 - s1 * t3;
          aqv = s1 * t1 + c1 * t3;
          if (kq == kp && rot1) {
            apv = 0.0;
            aqu = 0.0;
          }
          nxt[p][u] = apu;
          if (hv) nxt[p][v] = apv;
          if (hq) nxt[q][u] = aqu;
          if (hq && hv) nxt[q][v] = aqv;
          // rows kp and kp + half of V, columns of pair kq
          if (s2 != 0.0) {
            for (int i = kp; i < n; i += half) {
              const double x = V[i][u], y = V[i][v];
              V[i][u] = c2 * x - s2 * y;
              V[i][v] = s2 * x + c2 * y;
            }
          }
        }
      }
      __syncthreads();
      double(*t)[LD] = cur;
      cur = nxt;
      nxt = t;
    }
    if (!*rotated) break;
  }
  return cur;
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

// Block-wide maximum, same contract as block_sum.
__device__ inline double block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < nw; ++i) s = fmax(s, red[i]);
  return s;
}

}  // namespace ttf

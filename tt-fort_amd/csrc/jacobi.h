// Two-sided cyclic Jacobi for small symmetric matrices held in LDS, shared by
// the small eigensolver, SVQB orthonormalisation and Rayleigh-Ritz kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "wave.h"

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

// Rotation parameters (J = [[c, s], [-s, c]] on rows and columns p, q) that
// annihilate a_pq of the 2x2 block [[app, apq], [apq, aqq]] when `go`, else
// the identity.  t = tan(theta) is the smaller root of
// t^2 + 2 cot(2 theta) t - 1 = 0, t = sign(d) 2 a_pq / (|d| + sqrt(d^2 + 4 a_pq^2))
// with d = a_qq - a_pp, evaluated from the hardware reciprocal (square
// root) estimates refined by Newton steps: t to about 2^-46 (it only sets
// how well a_pq is annihilated), c = (1 + t^2)^(-1/2) by two steps so that
// c^2 + s^2 = 1 to rounding (the rotation stays orthogonal).  Branch free,
// so the rotations of two pairs interleave.
__device__ __forceinline__ void jacobi_cs(double apq, double app, double aqq, bool go, double& c, double& s) {
  // `go` only selects the inputs: (d, a) = (1, 0) gives q = 1, den = 2,
  // t = 0, hence c = 1, s = 0, with no branch for the compiler to make
  // around the arithmetic (which would serialise a thread's two rotations)
  const double a = go ? apq : 0.0;
  const double d = go ? aqq - app : 1.0;
  const double q = d * d + 4.0 * a * a;
  double rq = __builtin_amdgcn_rsq(q);
  rq = rq * (1.5 - 0.5 * q * rq * rq);
  const double den = fabs(d) + q * rq;
  double ri = __builtin_amdgcn_rcp(den);
  ri = ri * (2.0 - den * ri);
  const double t = (d >= 0.0 ? 2.0 : -2.0) * a * ri;
  const double x = 1.0 + t * t;
  double cc = __builtin_amdgcn_rsq(x);
  cc = cc * (1.5 - 0.5 * x * cc * cc);
  cc = cc * (1.5 - 0.5 * x * cc * cc);
  c = cc;
  s = t * cc;
}

// Whether pair (p, q) of the current matrix is rotated: |a_pq| above the
// caller's rounding floor tol_abs and above eps * sqrt(|a_pp a_qq|).
__device__ __forceinline__ bool jacobi_go(double apq, double app, double aqq, double tol_abs) {
  const double eps = 2.220446049250313e-16;
  return fabs(apq) > tol_abs && apq * apq > eps * eps * fabs(app * aqq);
}

// Rotations with |sin| at most this are "small": a sweep whose rotations
// are all small leaves off-diagonal entries of order n * kJacobiSmall^2 *
// ||A|| (quadratic convergence), below the rounding floor for n <= 64, so
// the next sweep would not change the result and is skipped.
constexpr double kJacobiSmall = 1e-9;

// The double-buffered form, for any n <= 64: diagonalises the symmetric
// n x n matrix held in A (row stride LD) and accumulates the rotations into
// V (V <- V J).  B is a second buffer: each round of the circle schedule reads one buffer and writes the other,
// so one barrier per round suffices.  Odd n is padded with a zero row and
// column n (in A, B and V; the buffers hold at least n + 1 rows), so every
// pair is real and no load is conditional; rotations involving the pad are
// the identity.  Thread t owns the 2x2 block (rows of pair kp) x (columns of
// pair kq), t = kp + half * kq, derives both rotations from the round's
// input itself (no phase publishes rotations) and also rotates rows kp and
// kp + half of V in the columns of pair kq.  All of a task's loads are
// issued before its arithmetic so that one LDS round trip and the two
// interleaved rotation computations make up the round.  The 2x2 blocks are
// updated in full (annihilated entries keep their rounding-level values).
// The loop ends after a sweep with no rotation above `small` (default
// kJacobiSmall) or after max_sweeps; a thread takes at most 4 tasks, so the
// block needs 4 blockDim.x >= (n / 2 + 1)^2.  `flags` is 3 ints of LDS.  All threads
// of the block call it; the returned pointer (A or B) holds the diagonalised
// matrix.
template <int LD>
__device__ double (*jacobi_sym_circle(double (*A)[LD], double (*B)[LD], double (*V)[LD], int n, int max_sweeps,
                                      double tol_abs, int* flags, double small = kJacobiSmall))[LD] {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n + (n & 1);
  const int half = nn / 2;
  if (n <= 1) return A;
  if (n & 1) {
    for (int j = tid; j <= n; j += nt) {
      A[n][j] = 0.0;
      A[j][n] = 0.0;
      B[n][j] = 0.0;
      B[j][n] = 0.0;
      V[n][j] = 0.0;
      V[j][n] = 0.0;
    }
  }
  const int ntask = half * half;   // at most 1024 = 4 x 256 tasks
  double(*cur)[LD] = A;
  double(*nxt)[LD] = B;
  if (tid < 3) flags[tid] = 0;
  // each thread's tasks (pair indices kp, kq), once: the rounds then have no
  // division on their critical path
  int kps[4], kqs[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int t = tid + w * nt;
    kps[w] = t % half;
    kqs[w] = t / half;
  }
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    int* moved = flags + sweep % 3;
    // flags[(sweep+1) % 3] was last read before the previous sweep's barriers
    if (tid == 0) flags[(sweep + 1) % 3] = 0;
    for (int r = 0; r < nn - 1; ++r) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        if (tid + w * nt >= ntask) break;
        const int kp = kps[w], kq = kqs[w];
        int p, q, u, v;
        jacobi_pair(r, kp, nn, p, q);
        jacobi_pair(r, kq, nn, u, v);
        const double app = cur[p][p], aqq = cur[q][q], apq = cur[p][q];
        const double auu = cur[u][u], avv = cur[v][v], auv = cur[u][v];
        const double apu = cur[p][u], apv = cur[p][v], aqu = cur[q][u], aqv = cur[q][v];
        const int i1 = kp + half;
        const double x0 = V[kp][u], y0 = V[kp][v], x1 = V[i1][u], y1 = V[i1][v];
        const bool go1 = jacobi_go(apq, app, aqq, tol_abs);
        const bool go2 = jacobi_go(auv, auu, avv, tol_abs);
        double c1, s1, c2, s2;
        jacobi_cs(apq, app, aqq, go1, c1, s1);
        jacobi_cs(auv, auu, avv, go2, c2, s2);
        if (kq == 0 && fabs(s1) > small) *moved = 1;
        // columns u, v
        const double t0 = c2 * apu - s2 * apv, t1 = s2 * apu + c2 * apv;
        const double t2 = c2 * aqu - s2 * aqv, t3 = s2 * aqu + c2 * aqv;
        // rows p, q
        nxt[p][u] = c1 * t0 - s1 * t2;
        nxt[p][v] = c1 * t1 - s1 * t3;
        nxt[q][u] = s1 * t0 + c1 * t2;
        nxt[q][v] = s1 * t1 + c1 * t3;
        // rows kp and kp + half of V, columns of pair kq
        V[kp][u] = c2 * x0 - s2 * y0;
        V[kp][v] = s2 * x0 + c2 * y0;
        V[i1][u] = c2 * x1 - s2 * y1;
        V[i1][v] = s2 * x1 + c2 * y1;
      }
      __syncthreads();
      double(*tt)[LD] = cur;
      cur = nxt;
      nxt = tt;
    }
    if (!*moved) break;
  }
  return cur;
}

// The same circle schedule with the matrix kept in slot order: slots 2k and
// 2k+1 hold the pair that round's rotation k acts on, and between rounds the
// contents move by one fixed permutation of the slots (the circle's
// rotation), so that after the nn - 1 rounds of a sweep every index is back
// in its starting slot.  Slot s starts with index jacobi_slot_index(s).
__device__ inline int jacobi_slot_index(int s, int nn) {
  return (s & 1) == 0 ? s >> 1 : (s == 1 ? nn - 1 : nn - 1 - (s >> 1));
}

// The slot that the contents of slot s move to between two rounds (half =
// nn / 2 pairs).
__device__ inline int jacobi_slot_next(int s, int half) {
  if (half == 1 || s == 1) return s;
  if (s == 0) return 3;
  if ((s & 1) == 0) return s - 2;
  if (s == 2 * half - 1) return 2 * half - 2;
  return s + 2;
}

// jacobi_sym for (n + 1) / 2 squared <= blockDim.x: thread (kp, kq) owns
// the 2x2 block of slot pairs kp x kq for the whole solve, with every
// address fixed once.  A round: each thread loads its block and its two
// rows of the vectors' slot pair kq; the diagonal-block threads compute the
// round's rotations and publish them; after a barrier every thread applies
// its two rotations and stores the block at the slots its contents move to.
// All loads of a round precede the barrier and all stores follow it, so the
// matrix and the vectors are updated in place.  Working copies: the matrix
// in slot order in B, the vectors (V with columns in slot order) in A; at
// the end V receives the vectors in index order and A's diagonal the
// eigenvalues, and A is returned.
template <int LD>
__device__ double (*jacobi_sym_slots(double (*A)[LD], double (*B)[LD], double (*V)[LD], int n, int max_sweeps,
                                     double tol_abs, int* flags, double small))[LD] {
  __shared__ double2 csb[kJacobiMaxN / 2];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n + (n & 1);
  const int half = nn / 2;
  if (n & 1) {
    for (int j = tid; j <= n; j += nt) {
      A[n][j] = 0.0;
      A[j][n] = 0.0;
      V[n][j] = 0.0;
      V[j][n] = 0.0;
    }
  }
  __syncthreads();
  for (int e = tid; e < nn * nn; e += nt) {
    const int s = e % nn, t = e / nn;
    B[s][t] = A[jacobi_slot_index(s, nn)][jacobi_slot_index(t, nn)];
  }
  __syncthreads();
  for (int e = tid; e < nn * nn; e += nt) {
    const int r = e % nn, s = e / nn;
    A[r][s] = V[r][jacobi_slot_index(s, nn)];
  }
  if (tid < 3) flags[tid] = 0;
  __syncthreads();
  double(*M)[LD] = B;
  double(*W)[LD] = A;
  const bool task = tid < half * half;
  const int kp = task ? tid % half : 0, kq = task ? tid / half : 0;
  const int p0 = 2 * kp, p1 = p0 + 1, q0 = 2 * kq, q1 = q0 + 1;
  const int np0 = jacobi_slot_next(p0, half), np1 = jacobi_slot_next(p1, half);
  const int nq0 = jacobi_slot_next(q0, half), nq1 = jacobi_slot_next(q1, half);
  const int r0 = kp, r1 = kp + half;   // the rows of W this thread rotates
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    int* moved = flags + sweep % 3;
    // flags[(sweep+1) % 3] was last read before the previous sweep's barriers
    if (tid == 0) flags[(sweep + 1) % 3] = 0;
    for (int r = 0; r < nn - 1; ++r) {
      double m00 = 0.0, m01 = 0.0, m10 = 0.0, m11 = 0.0, w00 = 0.0, w01 = 0.0, w10 = 0.0, w11 = 0.0;
      if (task) {
        m00 = M[p0][q0]; m01 = M[p0][q1]; m10 = M[p1][q0]; m11 = M[p1][q1];
        w00 = W[r0][q0]; w01 = W[r0][q1]; w10 = W[r1][q0]; w11 = W[r1][q1];
        if (kp == kq) {   // a diagonal block: [[a_pp, a_pq], [a_pq, a_qq]]
          double c, s;
          jacobi_cs(m01, m00, m11, jacobi_go(m01, m00, m11, tol_abs), c, s);
          csb[kp] = make_double2(c, s);
          if (fabs(s) > small) *moved = 1;
        }
      }
      __syncthreads();
      if (task) {
        const double2 rp = csb[kp], rq = csb[kq];
        const double c1 = rp.x, s1 = rp.y, c2 = rq.x, s2 = rq.y;
        // columns of pair kq, then rows of pair kp
        const double t0 = c2 * m00 - s2 * m01, t1 = s2 * m00 + c2 * m01;
        const double t2 = c2 * m10 - s2 * m11, t3 = s2 * m10 + c2 * m11;
        M[np0][nq0] = c1 * t0 - s1 * t2;
        M[np0][nq1] = c1 * t1 - s1 * t3;
        M[np1][nq0] = s1 * t0 + c1 * t2;
        M[np1][nq1] = s1 * t1 + c1 * t3;
        W[r0][nq0] = c2 * w00 - s2 * w01;
        W[r0][nq1] = s2 * w00 + c2 * w01;
        W[r1][nq0] = c2 * w10 - s2 * w11;
        W[r1][nq1] = s2 * w10 + c2 * w11;
      }
      __syncthreads();
    }
    if (!*moved) break;
  }
  // a sweep ends with every index in its starting slot
  for (int e = tid; e < nn * nn; e += nt) {
    const int r = e % nn, s = e / nn;
    V[r][jacobi_slot_index(s, nn)] = W[r][s];
  }
  __syncthreads();
  for (int s = tid; s < nn; s += nt) {
    const int i = jacobi_slot_index(s, nn);
    A[i][i] = M[s][s];
  }
  __syncthreads();
  return A;
}

// Diagonalises the symmetric n x n matrix held in A (row stride LD, n <= 64)
// and accumulates the rotations into V (V <- V J); B is a second buffer of
// the same shape.  Returns the matrix (A or B) whose diagonal holds the
// eigenvalues, in the order of V's columns.  Orders with (n + 1) / 2 squared
// <= blockDim.x take the slot-ordered kernel above, larger ones the
// double-buffered one.
template <int LD>
__device__ double (*jacobi_sym(double (*A)[LD], double (*B)[LD], double (*V)[LD], int n, int max_sweeps,
                               double tol_abs, int* flags, double small = kJacobiSmall))[LD] {
  if (n <= 1) return A;
  const int half = (n + 1) / 2;
  if (half * half <= (int)blockDim.x && n <= kJacobiMaxN)
    return jacobi_sym_slots<LD>(A, B, V, n, max_sweeps, tol_abs, flags, small);
  return jacobi_sym_circle<LD>(A, B, V, n, max_sweeps, tol_abs, flags, small);
}

// Block-wide sum of v over all threads (blockDim.x a multiple of 64, at most
// 1024); `red` holds one double per wave.  Every thread receives the sum.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum64(v);
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

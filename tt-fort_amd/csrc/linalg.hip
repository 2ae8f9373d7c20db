// Dense FP64 linear algebra kernels: tall-skinny Gram products, small
// symmetric eigenproblems (cyclic Jacobi in LDS) and the thin SVD (blocked
// one-sided Jacobi) used by the rank truncation of the sweeps.
#include "linalg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace ttf {

// ------------------------------------------------------------------ tsgram
constexpr int kTsgRows = 64;

__global__ __launch_bounds__(256) void tsgram_partial(int N, int p, int q, const double* __restrict__ X,
                                                      int64_t ldx, const double* __restrict__ Y,
                                                      int64_t ldy, double* __restrict__ part) {
  __shared__ double Xs[kTsgRows][65];
  __shared__ double Ys[kTsgRows][65];
  const int tid = threadIdx.x;
  const int pq = p * q;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  const int nchunks = (N + kTsgRows - 1) / kTsgRows;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int r0 = c * kTsgRows;
    __syncthreads();
    for (int e = tid; e < kTsgRows * 64; e += 256) {
      int r = e % kTsgRows, i = e / kTsgRows;
      bool ok = (r0 + r) < N;
      if (i < p) Xs[r][i] = ok ? X[(int64_t)(r0 + r) + ldx * i] : 0.0;
      if (i < q) Ys[r][i] = ok ? Y[(int64_t)(r0 + r) + ldy * i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      int e = tid + 256 * t;
      if (e < pq) {
        int i = e % p, j = e / p;
        double s = 0.0;
#pragma unroll 8
        for (int r = 0; r < kTsgRows; ++r) s += Xs[r][i] * Ys[r][j];
        acc[t] += s;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    int e = tid + 256 * t;
    if (e < pq) part[(int64_t)blockIdx.x * pq + e] = acc[t];
  }
}

__global__ void tsgram_reduce(int p, int q, int nparts, const double* __restrict__ part,
                              double* __restrict__ G, int ldg) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int pq = p * q;
  if (e >= pq) return;
  double s = 0.0;
  for (int w = 0; w < nparts; ++w) s += part[(int64_t)w * pq + e];
  G[(e % p) + (int64_t)ldg * (e / p)] = s;
}

void tsgram(int N, int p, int q, const double* X, int64_t ldx, const double* Y, int64_t ldy,
            double* G, int ldg, hipStream_t st) {
  TTF_REQUIRE(p >= 1 && q >= 1 && p <= 64 && q <= 64, "tsgram: p=%d q=%d out of range", p, q);
  const int nchunks = (int)ceil_div(std::max(N, 1), kTsgRows);
  const int nparts = std::min(nchunks, 256);
  double* part = (double*)arena(kArenaGram, sizeof(double) * (size_t)nparts * p * q);
  hipLaunchKernelGGL(tsgram_partial, dim3(nparts), dim3(256), 0, st, N, p, q, X, ldx, Y, ldy, part);
  TTF_HIP(hipGetLastError());
  hipLaunchKernelGGL(tsgram_reduce, dim3((unsigned)ceil_div(p * q, 256)), dim3(256), 0, st, p, q,
                     nparts, part, G, ldg);
  TTF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ Jacobi
// Cyclic Jacobi on the symmetric n x n matrix A (LDS, row stride LD) with
// eigenvector accumulation in V (LDS).  Disjoint rotation pairs of a round
// (circle schedule over an even index count) are applied together: first to
// columns of A and V, then to rows of A.  Runs until a sweep applies no
// rotation or `max_sweeps` is reached.  All threads of the block call it.
template <int LD>
__device__ void jacobi_lds(double (*A)[LD], double (*V)[LD], int n, int max_sweeps,
                           double* rc, double* rs, int* rp, int* rq, int* flag) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n + (n & 1);
  const int half = nn / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < nn - 1; ++r) {
      for (int k = tid; k < half; k += nt) {
        int a, b;
        if (k == 0) { a = r; b = nn - 1; }
        else { a = (r + k) % (nn - 1); b = (r - k + (nn - 1)) % (nn - 1); }
        int p = min(a, b), q = max(a, b);
        double c = 1.0, s = 0.0;
        if (q < n) {
          double apq = A[p][q], app = A[p][p], aqq = A[q][q];
          if (fabs(apq) > 1e-300 && fabs(apq) > 2.2e-16 * sqrt(fabs(app) * fabs(aqq))) {
            double tau = (aqq - app) / (2.0 * apq);
            double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            *flag = 1;
          }
        }
        rp[k] = p; rq[k] = q; rc[k] = c; rs[k] = s;
      }
      __syncthreads();
      // columns of A and V
      for (int e = tid; e < half * n; e += nt) {
        int k = e / n, i = e % n;
        int p = rp[k], q = rq[k];
        if (q >= n || rs[k] == 0.0) continue;
        double c = rc[k], s = rs[k];
        double aip = A[i][p], aiq = A[i][q];
        A[i][p] = c * aip - s * aiq;
        A[i][q] = s * aip + c * aiq;
        double vip = V[i][p], viq = V[i][q];
        V[i][p] = c * vip - s * viq;
        V[i][q] = s * vip + c * viq;
      }
      __syncthreads();
      // rows of A
      for (int e = tid; e < half * n; e += nt) {
        int k = e / n, j = e % n;
        int p = rp[k], q = rq[k];
        if (q >= n || rs[k] == 0.0) continue;
        double c = rc[k], s = rs[k];
        double apj = A[p][j], aqj = A[q][j];
        A[p][j] = c * apj - s * aqj;
        A[q][j] = s * apj + c * aqj;
      }
      __syncthreads();
    }
    if (*flag == 0) break;
    __syncthreads();
  }
}

// Sorts eigenpairs (A diagonal, V columns) and writes w, Q (global, ld n).
template <int LD>
__device__ void write_sorted(double (*A)[LD], double (*V)[LD], int n, int desc, double* w,
                            double* Q) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double wi = A[i][i];
    int pos = 0;
    for (int j = 0; j < n; ++j) {
      double wj = A[j][j];
      bool before = desc ? (wj > wi || (wj == wi && j < i)) : (wj < wi || (wj == wi && j < i));
      pos += before ? 1 : 0;
    }
    w[pos] = wi;
    for (int r = 0; r < n; ++r) Q[r + (int64_t)n * pos] = V[r][i];
  }
}

__global__ __launch_bounds__(256) void syev_small_kernel(int n, double* Aall, double* wall,
                                                         double* Qall, int desc) {
  __shared__ double A[64][65];
  __shared__ double V[64][65];
  __shared__ double rc[32], rs[32];
  __shared__ int rp[32], rq[32], flag;
  const double* Ag = Aall + (int64_t)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    int i = e % n, j = e / n;
    A[i][j] = 0.5 * (Ag[i + n * j] + Ag[j + n * i]);
    V[i][j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi_lds<65>(A, V, n, 40, rc, rs, rp, rq, &flag);
  __syncthreads();
  write_sorted<65>(A, V, n, desc, wall + (int64_t)blockIdx.x * n, Qall + (int64_t)blockIdx.x * n * n);
}

void syev_small(int n, int count, double* A, double* w, double* Q, int desc, hipStream_t st) {
  TTF_REQUIRE(n >= 1 && n <= 64, "syev_small: n=%d out of range (1..64)", n);
  if (count <= 0) return;
  hipLaunchKernelGGL(syev_small_kernel, dim3(count), dim3(256), 0, st, n, A, w, Q, desc);
  TTF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ SVD
// Blocked one-sided (Hestenes) Jacobi.  Work matrix W (mw x nw, ld mw,
// mw >= nw) and accumulated rotations V (nw x nw).  Columns are grouped in
// blocks of b; each round of a sweep pairs all blocks (circle schedule) and
// one work-group per pair loads the pair's 2b columns into LDS and runs
// one-sided Jacobi on them: 2x2 rotations computed from the columns' own dot
// products (which keeps small singular values relatively accurate), the
// product of the rotations accumulated in a 2b x 2b matrix that is applied
// to V's columns at the end.  The largest relative coupling between the two
// blocks seen before rotating drives the sweep loop.
constexpr int kSvdMaxP = 32;          // columns per pair (2b) at most
constexpr int kSvdLdsDoubles = 16384; // LDS budget for the pair's columns

__device__ inline void circle_pair(int r, int k, int nn, int* p, int* q) {
  int a, b;
  if (k == 0) { a = r; b = nn - 1; }
  else { a = (r + k) % (nn - 1); b = (r - k + (nn - 1)) % (nn - 1); }
  *p = min(a, b);
  *q = max(a, b);
}

__global__ __launch_bounds__(256) void svd_round_kernel(int mw, int nw, int bsz, int nb, int round,
                                                        double* __restrict__ W,
                                                        double* __restrict__ V, double small2,
                                                        unsigned long long* offmax) {
  extern __shared__ double cols_lds[];   // [2b][mw]
  __shared__ double Qp[kSvdMaxP][kSvdMaxP + 1];
  __shared__ int cols[kSvdMaxP];
  __shared__ int nv, nfirst, any;
  __shared__ unsigned long long local_off;
  const int tid = threadIdx.x;
  int bi, bj;
  circle_pair(round, blockIdx.x, nb, &bi, &bj);
  if (tid == 0) {
    int c = 0;
    for (int t = 0; t < bsz; ++t) if (bi * bsz + t < nw) cols[c++] = bi * bsz + t;
    nfirst = c;
    for (int t = 0; t < bsz; ++t) if (bj * bsz + t < nw) cols[c++] = bj * bsz + t;
    nv = c;
    local_off = 0ull;
  }
  __syncthreads();
  const int n = nv, nf = nfirst;
  if (n == 0 || nf == 0 || nf == n) return;   // a block beyond the last column
  for (int e = tid; e < n * mw; e += blockDim.x) {
    int j = e / mw, r = e % mw;
    cols_lds[(int64_t)j * mw + r] = W[(int64_t)r + (int64_t)mw * cols[j]];
  }
  for (int e = tid; e < kSvdMaxP * kSvdMaxP; e += blockDim.x)
    Qp[e / kSvdMaxP][e % kSvdMaxP] = (e / kSvdMaxP == e % kSvdMaxP) ? 1.0 : 0.0;
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  const double tol = 4.0 * eps * sqrt((double)mw);
  const int nn = n + (n & 1);
  const int g = tid >> 4, l16 = tid & 15;
  bool rotated_any = false;
  for (int isweep = 0; isweep < 12; ++isweep) {
    if (tid == 0) any = 0;
    __syncthreads();
    for (int r = 0; r < nn - 1; ++r) {
      if (g < nn / 2) {
        int p, q;
        circle_pair(r, g, nn, &p, &q);
        if (q < n) {
          double* ap = cols_lds + (int64_t)p * mw;
          double* aq = cols_lds + (int64_t)q * mw;
          double al = 0.0, be = 0.0, ga = 0.0;
          for (int i = l16; i < mw; i += 16) {
            double x = ap[i], y = aq[i];
            al += x * x; be += y * y; ga += x * y;
          }
#pragma unroll
          for (int o = 8; o > 0; o >>= 1) {
            al += __shfl_xor(al, o, 16);
            be += __shfl_xor(be, o, 16);
            ga += __shfl_xor(ga, o, 16);
          }
          // columns below the noise floor are numerically zero: never rotated
          double den = (al > small2 && be > small2) ? sqrt(al * be) : 0.0;
          double rel = den > 0.0 ? fabs(ga) / den : 0.0;
          if (isweep == 0 && p < nf && q >= nf && l16 == 0)
            atomicMax(&local_off, (unsigned long long)__double_as_longlong(rel));
          if (den > 0.0 && rel > tol) {
            double zeta = (be - al) / (2.0 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
            for (int i = l16; i < mw; i += 16) {
              double x = ap[i], y = aq[i];
              ap[i] = c * x - s * y;
              aq[i] = s * x + c * y;
            }
            for (int i = l16; i < n; i += 16) {
              double x = Qp[i][p], y = Qp[i][q];
              Qp[i][p] = c * x - s * y;
              Qp[i][q] = s * x + c * y;
            }
            if (l16 == 0) any = 1;
          }
        }
      }
      __syncthreads();
    }
    if (any) rotated_any = true;
    const int stop = (any == 0);
    __syncthreads();
    if (stop) break;
  }
  if (tid == 0) atomicMax(offmax, local_off);
  if (!rotated_any) return;
  for (int e = tid; e < n * mw; e += blockDim.x) {
    int j = e / mw, r = e % mw;
    W[(int64_t)r + (int64_t)mw * cols[j]] = cols_lds[(int64_t)j * mw + r];
  }
  __syncthreads();
  // V[:, cols] <- V[:, cols] Qp, streaming 64-row chunks of V through LDS
  double (*chunk)[kSvdMaxP + 1] = (double (*)[kSvdMaxP + 1])cols_lds;
  for (int r0 = 0; r0 < nw; r0 += 64) {
    for (int e = tid; e < 64 * n; e += blockDim.x) {
      int r = e % 64, j = e / 64;
      chunk[r][j] = (r0 + r < nw) ? V[(int64_t)(r0 + r) + (int64_t)nw * cols[j]] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < 64 * n; e += blockDim.x) {
      int r = e % 64, j = e / 64;
      if (r0 + r >= nw) continue;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += chunk[r][k] * Qp[k][j];
      V[(int64_t)(r0 + r) + (int64_t)nw * cols[j]] = s;
    }
    __syncthreads();
  }
}

__global__ void set_identity(int n, double* V) {
  int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    V[e] = (e % n == e / n) ? 1.0 : 0.0;
}

__global__ void transpose_kernel(int m, int n, const double* __restrict__ A, int64_t lda,
                                 double* __restrict__ B, int64_t ldb) {
  __shared__ double tile[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int i = bx + threadIdx.x, j = by + k;
    tile[k][threadIdx.x] = (i < m && j < n) ? A[i + lda * j] : 0.0;
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int j = by + threadIdx.x, i = bx + k;
    if (i < m && j < n) B[j + ldb * i] = tile[threadIdx.x][k];
  }
}

void transpose(int m, int n, const double* A, int64_t lda, double* B, int64_t ldb, hipStream_t st) {
  if (m <= 0 || n <= 0) return;
  dim3 grid((unsigned)ceil_div(m, 32), (unsigned)ceil_div(n, 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, m, n, A, lda, B, ldb);
  TTF_HIP(hipGetLastError());
}

__global__ void colnorm_kernel(int m, int n, const double* __restrict__ W, double* __restrict__ nrm) {
  __shared__ double red[256];
  int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double v = W[i + (int64_t)m * j];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) nrm[j] = sqrt(red[0]);
}

// out(:, k) = W(:, perm[k]) * scale[k]   (column gather with scaling)
__global__ void gather_scale_cols(int m, int k, const double* __restrict__ W, int64_t ldw,
                                  const int* __restrict__ perm, const double* __restrict__ scale,
                                  double* __restrict__ out, int64_t ldo) {
  int64_t total = (int64_t)m * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e % m), c = (int)(e / m);
    out[i + ldo * c] = W[i + ldw * perm[c]] * scale[c];
  }
}

// Orthonormal completion of columns [k0, k) of U (m x k, ld): each new
// column starts from a deterministic pseudo-random vector and is
// orthogonalised twice against all previous columns.  One work-group.
// Classical Gram-Schmidt applied twice, with all projection coefficients of
// a pass computed in parallel (one thread per previous column).
__global__ __launch_bounds__(256) void complete_basis(int m, int k0, int k, double* U, int64_t ldu) {
  __shared__ double red[256];
  __shared__ double coef;
  __shared__ double cj[4096];
  for (int c = k0; c < k; ++c) {
    double* u = U + ldu * c;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      unsigned h = (unsigned)(i * 2654435761u) ^ (unsigned)(c * 40503u + 17u);
      h ^= h >> 13; h *= 0x5bd1e995u; h ^= h >> 15;
      u[i] = ((double)(h & 0xffffff) / (double)0x1000000) - 0.5;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      for (int j = threadIdx.x; j < c; j += blockDim.x) {
        const double* v = U + ldu * j;
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += v[i] * u[i];
        cj[j] = s;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < c; ++j) s += cj[j] * U[i + ldu * j];
        u[i] -= s;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) s += u[i] * u[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) coef = 1.0 / sqrt(red[0]);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) u[i] *= coef;
    __syncthreads();
  }
}

int gesvd(int m, int n, const double* A, int64_t lda, double* U, int64_t ldu, double* S,
          double* VT, int64_t ldvt, hipStream_t st, int complete) {
  TTF_REQUIRE(m >= 1 && n >= 1, "gesvd: invalid sizes %d x %d", m, n);
  const bool tall = m >= n;
  const int mw = tall ? m : n, nw = tall ? n : m;   // work matrix mw x nw
  const int k = nw;
  TTF_REQUIRE(nw <= 4096, "gesvd: min(m,n)=%d exceeds 4096", nw);
  // arena layout: W (mw*nw) | V (nw*nw) | nrm (nw) | scale (nw) | perm (nw ints) | offmax
  size_t need = sizeof(double) * ((size_t)mw * nw + (size_t)nw * nw + 2 * (size_t)nw) +
                sizeof(int) * (size_t)nw + 64;
  char* base = (char*)arena(kArenaTrunc, need);
  double* W = (double*)base;
  double* V = W + (size_t)mw * nw;
  double* nrm = V + (size_t)nw * nw;
  double* scale = nrm + nw;
  int* perm = (int*)(scale + nw);
  unsigned long long* offmax = (unsigned long long*)(((uintptr_t)(perm + nw) + 15) & ~(uintptr_t)15);
  if (tall) {
    TTF_HIP(hipMemcpy2DAsync(W, sizeof(double) * mw, A, sizeof(double) * lda, sizeof(double) * m, n,
                             hipMemcpyDeviceToDevice, st));
  } else {
    transpose(m, n, A, lda, W, mw, st);
  }
  hipLaunchKernelGGL(set_identity, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)nw * nw, 256), 4096)),
                     dim3(256), 0, st, nw, V);
  TTF_HIP(hipGetLastError());
  // column norms bound the absolute noise floor of the cross products
  hipLaunchKernelGGL(colnorm_kernel, dim3(nw), dim3(256), 0, st, mw, nw, W, nrm);
  std::vector<double> hn(nw);
  TTF_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
  TTF_HIP(hipStreamSynchronize(st));
  double fro2 = 0.0;
  for (double v : hn) fro2 += v * v;
  int sweeps = 0;
  double small2 = 0.0;
  if (fro2 > 0.0) {
    const double eps = 2.220446049250313e-16;
    const double tol = 4.0 * eps * std::sqrt((double)mw);
    // squared column norms below (sqrt(mw) eps ||A||_F)^2 are rounding noise
    small2 = mw * eps * eps * fro2;
    // block size: the pair's 2b columns must fit the LDS budget
    int two_b = std::min<int>(kSvdMaxP, kSvdLdsDoubles / mw);
    two_b = std::max(2, two_b - (two_b & 1));
    // at least two non-empty blocks, so that every column pair is coupled
    const int bsz = std::max(1, std::min(two_b / 2, (int)ceil_div(nw, 2)));
    int nb = (int)ceil_div(nw, bsz);
    if (nb % 2) ++nb;
    if (nb < 2) nb = 2;
    const size_t lds = sizeof(double) * (size_t)std::max(two_b * mw, 64 * (kSvdMaxP + 1));
    TTF_REQUIRE(lds <= 160 * 1024, "gesvd: pair block does not fit LDS (mw=%d)", mw);
    static bool attr_set = false;
    if (!attr_set) {
      TTF_HIP(hipFuncSetAttribute((const void*)svd_round_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 10240));
      attr_set = true;
    }
    unsigned long long hoff = 0;
    for (sweeps = 1; sweeps <= 40; ++sweeps) {
      TTF_HIP(hipMemsetAsync(offmax, 0, sizeof(unsigned long long), st));
      for (int r = 0; r < nb - 1; ++r) {
        hipLaunchKernelGGL(svd_round_kernel, dim3(nb / 2), dim3(256), lds, st, mw, nw, bsz, nb, r, W,
                           V, small2, offmax);
        TTF_HIP(hipGetLastError());
      }
      TTF_HIP(hipMemcpyAsync(&hoff, offmax, sizeof(hoff), hipMemcpyDeviceToHost, st));
      TTF_HIP(hipStreamSynchronize(st));
      double off;
      std::memcpy(&off, &hoff, sizeof(off));
      if (off <= tol) break;
    }
  }
  hipLaunchKernelGGL(colnorm_kernel, dim3(nw), dim3(256), 0, st, mw, nw, W, nrm);
  TTF_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
  TTF_HIP(hipStreamSynchronize(st));
  std::vector<int> order(nw);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hn[a] > hn[b]; });
  // columns at the rounding-noise floor count as zero singular values
  const double tiny = std::max(std::sqrt(small2), 1e-300);
  std::vector<double> hs(nw), hscale(nw), hone(nw, 1.0);
  int kpos = 0;
  for (int c = 0; c < nw; ++c) {
    hs[c] = hn[order[c]];
    hscale[c] = hs[c] > tiny ? 1.0 / hs[c] : 0.0;
    if (hs[c] > tiny) kpos = c + 1;
  }
  // upload permutation and scales, then form the normalised left factor of
  // W and the permuted V
  TTF_HIP(hipMemcpyAsync(perm, order.data(), sizeof(int) * nw, hipMemcpyHostToDevice, st));
  TTF_HIP(hipMemcpyAsync(scale, hscale.data(), sizeof(double) * nw, hipMemcpyHostToDevice, st));
  TTF_HIP(hipMemcpyAsync(S, hs.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
  double* Uw = tall ? U : (double*)arena(kArenaTrunc2, sizeof(double) * (size_t)mw * k);
  const int64_t lduw = tall ? ldu : mw;
  int blocks = (int)std::min<int64_t>(ceil_div((int64_t)mw * k, 256), 8192);
  hipLaunchKernelGGL(gather_scale_cols, dim3(blocks), dim3(256), 0, st, mw, k, W, (int64_t)mw, perm,
                     scale, Uw, lduw);
  TTF_HIP(hipGetLastError());
  if (kpos < k && complete) {
    hipLaunchKernelGGL(complete_basis, dim3(1), dim3(256), 0, st, mw, kpos, k, Uw, lduw);
    TTF_HIP(hipGetLastError());
  }
  // V permuted (scale 1): Vp(:, c) = V(:, order[c])
  double* Vp = (double*)arena(kArenaTrunc3, sizeof(double) * (size_t)nw * nw);
  TTF_HIP(hipMemcpyAsync(scale, hone.data(), sizeof(double) * nw, hipMemcpyHostToDevice, st));
  blocks = (int)std::min<int64_t>(ceil_div((int64_t)nw * nw, 256), 8192);
  hipLaunchKernelGGL(gather_scale_cols, dim3(blocks), dim3(256), 0, st, nw, nw, V, (int64_t)nw, perm,
                     scale, Vp, (int64_t)nw);
  TTF_HIP(hipGetLastError());
  if (tall) {
    // A = U S V^T: VT (k x n) = Vp^T
    transpose(nw, nw, Vp, nw, VT, ldvt, st);
  } else {
    // A^T = Uw S Vp^T  =>  A = Vp S Uw^T: U = Vp (m x m), VT = Uw^T (m x n)
    TTF_HIP(hipMemcpy2DAsync(U, sizeof(double) * ldu, Vp, sizeof(double) * nw, sizeof(double) * nw, nw,
                             hipMemcpyDeviceToDevice, st));
    transpose(mw, k, Uw, mw, VT, ldvt, st);
  }
  TTF_HIP(hipStreamSynchronize(st));
  return sweeps;
}

}  // namespace ttf

// ------------------------------------------------------------------ C ABI
using namespace ttf;

extern "C" {

int ttf_dgesvd(int m, int n, const double* A, int64_t lda, double* U, int64_t ldu, double* S,
               double* VT, int64_t ldvt, int* sweeps) {
  return guarded([&] {
    int sw = gesvd(m, n, A, lda, U, ldu, S, VT, ldvt, stream());
    if (sweeps) *sweeps = sw;
  });
}

int ttf_dsyev_small(int n, int count, double* A, double* w, double* Q, int desc) {
  return guarded([&] { syev_small(n, count, A, w, Q, desc, stream()); });
}

int ttf_dgram(int N, int p, int q, const double* X, int64_t ldx, const double* Y, int64_t ldy,
              double* G, int ldg) {
  return guarded([&] { tsgram(N, p, q, X, ldx, Y, ldy, G, ldg, stream()); });
}

}  // extern "C"

// Dense FP64 linear algebra kernels: tall-skinny Gram products, small
// symmetric eigenproblems (cyclic Jacobi in LDS) and the thin SVD (blocked
// one-sided Jacobi) used by the rank truncation of the sweeps.
#include "linalg.h"
#include "gemm_mfma.h"
#include "jacobi.h"
#include "wave.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace ttf {

// ------------------------------------------------------------------ tsgram
constexpr int kTsgRows = 64;   // rows per chunk: one per lane of a wave

// G (p x q) = X^T Y for tall X (N x p), Y (N x q) in two launches: every
// work-group of tsgram_partial sums row chunks of 64 into its partial p x q
// block, and tsgram_reduce adds the partials in a fixed order (so the
// result is reproducible): the 16 lanes of a DPP row take interleaved
// partials of one output, one batch of loads each, and a row sum combines
// them.  X's columns may come from two matrices: columns [0, p1) from X,
// the rest from X2 (ld ldx2).
__global__ __launch_bounds__(256) void tsgram_partial(int N, int p, int q, const double* __restrict__ X,
                                                      int64_t ldx, int p1, const double* __restrict__ X2,
                                                      int64_t ldx2, const double* __restrict__ Y,
                                                      int64_t ldy, double* __restrict__ part) {
  // chunk rows staged in LDS (columns past p / q as zeros), partial products
  // on the matrix cores: the output's 16x16 tiles (at most 16) go to the
  // four waves in turn, each tile a chain of 64-row x 4-deep MFMAs per chunk
  __shared__ double Xs[kTsgRows][65];
  __shared__ double Ys[kTsgRows][65];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tp = (p + 15) >> 4, tq = (q + 15) >> 4, ntile = tp * tq;
  const int fi = lane & 15, fk = lane >> 4;
  d4_t acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = d4_t{0.0, 0.0, 0.0, 0.0};
  const int nchunks = (N + kTsgRows - 1) / kTsgRows;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int r0 = c * kTsgRows;
    // thread: row r of the chunk, columns i0, i0 + 4, ...; every load in
    // flight before the LDS stores (a wave reads 64 consecutive rows)
    const int r = tid & 63, i0 = tid >> 6;
    const bool ok = (r0 + r) < N;
    double xv[16], yv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = i0 + 4 * u;
      const double* xc = i < p1 ? X + ldx * i : X2 + ldx2 * (i - p1);
      xv[u] = (ok && i < p) ? xc[r0 + r] : 0.0;
      yv[u] = (ok && i < q) ? Y[(int64_t)(r0 + r) + ldy * i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      Xs[r][i0 + 4 * u] = xv[u];
      Ys[r][i0 + 4 * u] = yv[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = wv + 4 * u;
      if (t < ntile) {   // uniform over the wave
        const int ti = t % tp, tj = t / tp;
#pragma unroll 4
        for (int k0 = 0; k0 < kTsgRows; k0 += 4)
          acc[u] = mfma16(Xs[k0 + fk][16 * ti + fi], Ys[k0 + fk][16 * tj + fi], acc[u]);
      }
    }
  }
  const int pq = p * q;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int t = wv + 4 * u;
    if (t < ntile) {
      const int ti = t % tp, tj = t / tp;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * ti + mfma_row<double>(lane, r), j = 16 * tj + fi;
        if (i < p && j < q) part[(int64_t)blockIdx.x * pq + i + p * j] = acc[u][r];
      }
    }
  }
}

constexpr int kTsgSub = 16;   // lanes per output in the reduction (one DPP row)

__global__ __launch_bounds__(256) void tsgram_reduce(int p, int q, int nparts, const double* __restrict__ part,
                                                     double* __restrict__ G, int ldg) {
  const int pq = p * q, tid = threadIdx.x;
  const int e = blockIdx.x * (256 / kTsgSub) + tid / kTsgSub, k = tid % kTsgSub;
  // each lane's partials (nparts <= 256: at most 16) all loaded before the
  // sum, in a fixed order
  double v[256 / kTsgSub];
#pragma unroll
  for (int u = 0; u < 256 / kTsgSub; ++u) {
    const int w = k + kTsgSub * u;
    v[u] = (e < pq && w < nparts) ? part[(int64_t)w * pq + e] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < 256 / kTsgSub; ++u) s += v[u];
  s = row_sum16(s);
  if (k == 0 && e < pq) G[(e % p) + (int64_t)ldg * (e / p)] = s;
}

void tsgram2(int N, int p1, const double* X1, int64_t ldx1, int p2, const double* X2, int64_t ldx2, int q,
             const double* Y, int64_t ldy, double* G, int ldg, hipStream_t st) {
  const int p = p1 + p2;
  TTF_REQUIRE(p1 >= 0 && p2 >= 0 && p >= 1 && q >= 1 && p <= 64 && q <= 64, "tsgram: p=%d+%d q=%d out of range",
              p1, p2, q);
  const int nchunks = (int)ceil_div(std::max(N, 1), kTsgRows);
  const int nparts = std::min(nchunks, 256);   // tsgram_reduce takes at most 256 partials
  double* part = (double*)arena(kArenaGram, sizeof(double) * (size_t)nparts * p * q);
  hipLaunchKernelGGL(tsgram_partial, dim3(nparts), dim3(256), 0, st, N, p, q, X1, ldx1, p1, p2 > 0 ? X2 : X1,
                     p2 > 0 ? ldx2 : ldx1, Y, ldy, part);
  TTF_HIP(hipGetLastError());
  hipLaunchKernelGGL(tsgram_reduce, dim3((unsigned)ceil_div(p * q, 256 / kTsgSub)), dim3(256), 0, st, p, q,
                     nparts, part, G, ldg);
  TTF_HIP(hipGetLastError());
}

void tsgram(int N, int p, int q, const double* X, int64_t ldx, const double* Y, int64_t ldy,
            double* G, int ldg, hipStream_t st) {
  tsgram2(N, p, X, ldx, 0, X, ldx, q, Y, ldy, G, ldg, st);
}

// ------------------------------------------------------------------ small eig
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
  __shared__ double A[64][65], B[64][65];
  __shared__ double V[64][65];
  __shared__ int flags[3];
  __shared__ double red[16];
  const double* Ag = Aall + (int64_t)blockIdx.x * n * n;
  double f2 = 0.0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    int i = e % n, j = e / n;
    double a = 0.5 * (Ag[i + n * j] + Ag[j + n * i]);
    A[i][j] = a;
    V[i][j] = (i == j) ? 1.0 : 0.0;
    f2 += a * a;
  }
  f2 = block_sum(f2, red);
  double(*D)[65] = jacobi_sym<65>(A, B, V, n, 40, 2.220446049250313e-16 * sqrt(f2), flags);
  write_sorted<65>(D, V, n, desc, wall + (int64_t)blockIdx.x * n, Qall + (int64_t)blockIdx.x * n * n);
}

void syev_small(int n, int count, double* A, double* w, double* Q, int desc, hipStream_t st) {
  TTF_REQUIRE(n >= 1 && n <= 64, "syev_small: n=%d out of range (1..64)", n);
  if (count <= 0) return;
  hipLaunchKernelGGL(syev_small_kernel, dim3(count), dim3(256), 0, st, n, A, w, Q, desc);
  TTF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ SVD
// One-sided (Hestenes) Jacobi on the columns of a work matrix X (mw x nw,
// ld mw, mw >= nw), optionally accumulating the rotations in V (nw x nw).
// Columns are grouped in blocks of b; each round of a sweep pairs all
// blocks (circle schedule) and one work-group per pair takes the pair's
// n = 2b columns P into LDS:
//   1. G = P^T P by MFMA; a pair whose couplings |g_ij| / sqrt(g_ii g_jj)
//      are all below half the tolerance is left untouched;
//   2. a bounded number of Hestenes sweeps over the pair's columns: 2x2
//      rotations from the columns' own dot products (which keeps small
//      singular values accurate), their product accumulated in Qp; the
//      largest coupling measured here (the same quantity that decides a
//      rotation) drives the sweep loop, so a sweep without rotations ends it;
//   3. P back to X and, with V, V(:, pair) <- V(:, pair) Qp by MFMA.
// Columns whose squared norm is below the rounding floor small2 are treated
// as zero and never rotated.
constexpr int kSvdMaxP = 32;          // columns per pair (2b) at most
constexpr int kSvdLdsDoubles = 16384; // LDS budget for the pair's columns
static_assert((kSvdMaxP & (kSvdMaxP - 1)) == 0, "kSvdMaxP must be a power of two");

__device__ inline void circle_pair(int r, int k, int nn, int* p, int* q) {
  int a, b;
  if (k == 0) { a = r; b = nn - 1; }
  else { a = (r + k) % (nn - 1); b = (r - k + (nn - 1)) % (nn - 1); }
  *p = min(a, b);
  *q = max(a, b);
}

// In-place X (rows x n, column stride ld, in LDS) <- X Qp with Qp(k, j) =
// Qc[j][k]; each wave takes 16-row slabs and loads the slab's A fragments
// before writing any of its outputs.
__device__ inline void lds_apply_rotation(double* X, int rows, int ld, int n,
                                          const double (*Qc)[kSvdMaxP + 1]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int nt = (n + 15) >> 4, nk = (n + 3) >> 2;
  for (int r0 = wv * 16; r0 < rows; r0 += 16 * nwv) {
    const int row = r0 + (lane & 15);
    double a[kSvdMaxP / 4];
#pragma unroll
    for (int t = 0; t < kSvdMaxP / 4; ++t) {
      const int k = 4 * t + (lane >> 4);
      a[t] = (t < nk && k < n && row < rows) ? X[k * ld + row] : 0.0;
    }
    for (int tj = 0; tj < nt; ++tj) {
      const int cj = tj * 16 + (lane & 15);
      d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < kSvdMaxP / 4; ++t) {
        if (t < nk) {
          const int k = 4 * t + (lane >> 4);
          const double bq = (k < n && cj < n) ? Qc[cj][k] : 0.0;
          acc = mfma16(a[t], bq, acc);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int orow = r0 + mfma_row<double>(lane, r);
        if (orow < rows && cj < n) X[cj * ld + orow] = acc[r];
      }
    }
  }
}

// G (n x n, n <= 32) = P^T P for the LDS columns P (rows x n, column stride
// ld), by MFMA: T = nt^2 tiles of 16x16, the block's 4 waves split tiles x
// rows.  `scratch` holds 1024 doubles.  Ends with a block barrier.
__device__ inline void lds_gram(const double* P, int rows, int ld, int n, double (*G)[kSvdMaxP + 1],
                                double* scratch) {
  // the first four waves compute; the others only take the barriers
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool comp = wv < 4;
  const int nt = (n + 15) >> 4, T = nt * nt, parts = 4 / T;
  const int tile = wv % T, part = wv / T;
  const int ti = tile % nt, tj = tile / nt;
  const int kchunk = (((rows + 3) / 4 + parts - 1) / parts) * 4;
  const int kbeg = part * kchunk, kend = min(rows, kbeg + kchunk);
  const int ci = ti * 16 + (lane & 15), cj = tj * 16 + (lane & 15), kk = lane >> 4;
  const bool va = ci < n, vb = cj < n;
  const double* pa = P + (va ? ci : 0) * ld;
  const double* pb = P + (vb ? cj : 0) * ld;
  d4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0, acc2 = acc0, acc3 = acc0;
  int k0 = comp ? kbeg : kend;
  // 32 rows a step: all 16 loads in flight before the 8 MFMAs (columns past
  // n read column 0 and are zeroed by a select, not a branch)
  for (; k0 + 32 <= kend; k0 += 32) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = k0 + 4 * u + kk;
      a[u] = pa[r];
      b[u] = pb[r];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = va ? a[u] : 0.0;
      b[u] = vb ? b[u] : 0.0;
    }
    acc0 = mfma16(a[0], b[0], acc0);
    acc1 = mfma16(a[1], b[1], acc1);
    acc2 = mfma16(a[2], b[2], acc2);
    acc3 = mfma16(a[3], b[3], acc3);
    acc0 = mfma16(a[4], b[4], acc0);
    acc1 = mfma16(a[5], b[5], acc1);
    acc2 = mfma16(a[6], b[6], acc2);
    acc3 = mfma16(a[7], b[7], acc3);
  }
  for (; k0 < kend; k0 += 8) {
    const int r0 = k0 + kk, r1 = k0 + 4 + kk;
    const double a0 = (va && r0 < kend) ? pa[r0] : 0.0, b0 = (vb && r0 < kend) ? pb[r0] : 0.0;
    const double a1 = (va && r1 < kend) ? pa[r1] : 0.0, b1 = (vb && r1 < kend) ? pb[r1] : 0.0;
    acc0 = mfma16(a0, b0, acc0);
    acc1 = mfma16(a1, b1, acc1);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    acc0[r] += acc2[r];
    acc1[r] += acc3[r];
  }
  if (T == 4) {
    if (comp)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        G[ti * 16 + mfma_row<double>(lane, r)][tj * 16 + (lane & 15)] = acc0[r] + acc1[r];
    __syncthreads();
  } else {   // one tile, rows split over the 4 waves: partial tiles summed in order
    if (comp)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        scratch[part * 256 + mfma_row<double>(lane, r) * 16 + (lane & 15)] = acc0[r] + acc1[r];
    __syncthreads();
    for (int e = threadIdx.x; e < 256; e += blockDim.x)
      G[e >> 4][e & 15] = ((scratch[e] + scratch[256 + e]) + scratch[512 + e]) + scratch[768 + e];
    __syncthreads();
  }
}

// Column block copies between global memory (column stride gld, columns
// cols[0..n)) and LDS (column stride ld), column by column with each
// thread's rows: no division in the address arithmetic, and each thread
// issues the accesses of 16 columns together.  When the rows fit in half the
// block, the two halves take alternate groups of 16 columns, so that a pair
// of 32 columns of up to blockDim / 2 rows moves in one round trip.
__device__ inline void cols_copy_split(int rows, int& t, int& step, int& c_first, int& c_step) {
  const int nt = blockDim.x;
  const bool split = 2 * rows <= nt;
  step = split ? nt / 2 : nt;
  const bool upper = split && (int)threadIdx.x >= step;
  t = upper ? (int)threadIdx.x - step : (int)threadIdx.x;
  c_first = upper ? 16 : 0;
  c_step = split ? 32 : 16;
}

__device__ inline void cols_to_lds(const double* __restrict__ src, int64_t gld, const int* cols, int rows,
                                   int n, double* dst, int ld) {
  int t, step, c_first, c_step;
  cols_copy_split(rows, t, step, c_first, c_step);
  for (int c0 = c_first; c0 < n; c0 += c_step) {
    for (int r = t; r < rows; r += step) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = (c0 + u < n) ? src[(int64_t)r + gld * cols[c0 + u]] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (c0 + u < n) dst[(c0 + u) * ld + r] = v[u];
    }
  }
}

__device__ inline void lds_to_cols(const double* src, int ld, int rows, int n, double* __restrict__ dst,
                                   int64_t gld, const int* cols) {
  int t, step, c_first, c_step;
  cols_copy_split(rows, t, step, c_first, c_step);
  for (int c0 = c_first; c0 < n; c0 += c_step) {
    for (int r = t; r < rows; r += step) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = (c0 + u < n) ? src[(c0 + u) * ld + r] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (c0 + u < n) dst[(int64_t)r + gld * cols[c0 + u]] = v[u];
    }
  }
}

// Maximum of a non-negative v over the work-group into *m (an LDS word
// cleared beforehand, read after a barrier): wave maxima first, so that one
// lane per wave does the atomic.
__device__ inline void block_max_into(double v, unsigned long long* m) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) atomicMax(m, (unsigned long long)__double_as_longlong(v));
}

// A column pair's 2x2 Hestenes step inside the pair block: returns the
// measured coupling (0 for a column at the noise floor) and rotates the two
// LDS columns when it exceeds tol.  The 16 lanes of one DPP row (one row
// segment each) share a pair and reduce the dot products in registers; with
// NR > 0 each thread keeps its NR elements per column in registers between
// the dot products and the update (mw <= 16 NR).  FULL (mw == LPP NR: every
// lane's NR rows exist) drops the per-row tests, which otherwise become a
// branch around every load and store.
template <int NR, int LPP = 16, bool FULL = false>
__device__ inline double hestenes_pair(double* ap, double* aq, int mw, int l16, double small2, double tol,
                                       double* c_out, double* s_out) {
  static_assert(LPP == 16 || LPP == 32, "lanes per pair: 16 or 32");
  double al = 0.0, be = 0.0, ga = 0.0;
  double xr[NR > 0 ? NR : 1], yr[NR > 0 ? NR : 1];
  if constexpr (NR > 0) {
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      const int i = l16 + LPP * t;
      xr[t] = (FULL || i < mw) ? ap[i] : 0.0;
      yr[t] = (FULL || i < mw) ? aq[i] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      al += xr[t] * xr[t];
      be += yr[t] * yr[t];
      ga += xr[t] * yr[t];
    }
  } else {
    for (int i = l16; i < mw; i += LPP) {
      const double x = ap[i], y = aq[i];
      al += x * x; be += y * y; ga += x * y;
    }
  }
  al = row_sum16(al);
  be = row_sum16(be);
  ga = row_sum16(ga);
  if constexpr (LPP == 32) {   // the pair's two DPP rows: both get the same total
    al += __shfl_xor(al, 16);
    be += __shfl_xor(be, 16);
    ga += __shfl_xor(ga, 16);
  }
  *s_out = 0.0;
  if (!(al > small2 && be > small2)) return 0.0;
  // coupling |ga| / sqrt(al be) from the refined reciprocal square root
  const double ab = al * be;
  double ri = __builtin_amdgcn_rsq(ab);
  ri = ri * (1.5 - 0.5 * ab * ri * ri);
  const double rel = fabs(ga) * ri;
  if (rel <= tol) return rel;
  // the rotation that orthogonalises the two columns is the Jacobi rotation
  // of their 2x2 Gram matrix [[al, ga], [ga, be]]
  double c, s;
  jacobi_cs(ga, al, be, true, c, s);
  if constexpr (NR > 0) {
#pragma unroll
    for (int t2 = 0; t2 < NR; ++t2) {
      const int i = l16 + LPP * t2;
      if (FULL || i < mw) {
        ap[i] = c * xr[t2] - s * yr[t2];
        aq[i] = s * xr[t2] + c * yr[t2];
      }
    }
  } else {
    for (int i = l16; i < mw; i += LPP) {
      const double x = ap[i], y = aq[i];
      ap[i] = c * x - s * y;
      aq[i] = s * x + c * y;
    }
  }
  *c_out = c;
  *s_out = s;
  return rel;
}

template <int NR, int LPP = 16, bool FULL = false>
__global__ __launch_bounds__(16 * LPP) void svd_pair_kernel(int mw, int nw, int bsz, int nb, int round,
                                                       double* __restrict__ W, double* __restrict__ V,
                                                       double small2, double tol, int inner_sweeps,
                                                       unsigned long long* offmax) {
  extern __shared__ double lds[];   // pair columns: column j at lds + j * ldp
  __shared__ double G[kSvdMaxP][kSvdMaxP + 1];    // Gram of the pair's columns
  __shared__ double Qc[kSvdMaxP][kSvdMaxP + 1];   // Qp by columns: Qc[j][i] = Qp(i, j)
  __shared__ double dinv[kSvdMaxP];
  __shared__ int cols[kSvdMaxP];
  __shared__ int nv, nfirst, any, rotated;
  __shared__ unsigned long long cmax, rmax;
  const int tid = threadIdx.x;
  int bi, bj;
  circle_pair(round, blockIdx.x, nb, &bi, &bj);
  if (tid == 0) {
    int c = 0;
    for (int t = 0; t < bsz; ++t) if (bi * bsz + t < nw) cols[c++] = bi * bsz + t;
    nfirst = c;
    for (int t = 0; t < bsz; ++t) if (bj * bsz + t < nw) cols[c++] = bj * bsz + t;
    nv = c;
    cmax = 0ull;
    rmax = 0ull;
    rotated = 0;
  }
  __syncthreads();
  const int n = nv, nf = nfirst;
  if (n == 0 || nf == 0 || nf == n) return;   // a block beyond the last column
  const int ldp = mw | 1;                     // odd stride: conflict-free fragment loads
  cols_to_lds(W, mw, cols, mw, n, lds, ldp);
  __syncthreads();
  // ---- 1. Gram and the skip test
  lds_gram(lds, mw, ldp, n, G, &Qc[0][0]);
  for (int i = tid; i < n; i += blockDim.x) {
    const double g = G[i][i];
    dinv[i] = g > small2 ? 1.0 / sqrt(g) : 0.0;
  }
  __syncthreads();
  {
    double cm = 0.0;
    for (int e = tid; e < kSvdMaxP * kSvdMaxP; e += blockDim.x) {
      const int i = e % kSvdMaxP, j = e / kSvdMaxP;   // kSvdMaxP is a power of two
      if (i < j && j < n) cm = fmax(cm, fabs(G[i][j]) * dinv[i] * dinv[j]);
    }
    block_max_into(cm, &cmax);
  }
  __syncthreads();
  if (__longlong_as_double((long long)cmax) <= 0.5 * tol) return;
  // ---- 2. Hestenes rotations on the pair's columns; 16 groups of LPP
  // threads take one column pair each per round of the circle schedule
  const bool accum = (V != nullptr);
  if (accum)
    for (int e = tid; e < kSvdMaxP * kSvdMaxP; e += blockDim.x) {
      const int j = e / kSvdMaxP, i = e % kSvdMaxP;
      Qc[j][i] = (i == j) ? 1.0 : 0.0;
    }
  __syncthreads();
  {
    const int nn = n + (n & 1);
    const int g = tid / LPP, l16 = tid % LPP;
    double seen = 0.0;
    for (int isw = 0; isw < inner_sweeps; ++isw) {
      if (tid == 0) any = 0;
      __syncthreads();
      for (int r = 0; r < nn - 1; ++r) {
        if (g < nn / 2) {
          int p, q;
          jacobi_pair(r, g, nn, p, q);   // circle_pair's pairs, without the divisions
          if (q < n) {
            double c, s;
            const double rel = hestenes_pair<NR, LPP, FULL>(lds + p * ldp, lds + q * ldp, mw, l16, small2, tol, &c, &s);
            seen = fmax(seen, rel);
            if (s != 0.0) {
              if (accum) {   // the rotations are only needed to update V
#pragma unroll
                for (int i = l16; i < kSvdMaxP; i += LPP) {
                  const double u = Qc[p][i], v = Qc[q][i];
                  Qc[p][i] = c * u - s * v;
                  Qc[q][i] = s * u + c * v;
                }
              }
              if (l16 == 0) any = 1;
            }
          }
        }
        __syncthreads();
      }
      const int stop = (any == 0);
      if (!stop && tid == 0) rotated = 1;
      __syncthreads();
      if (stop) break;
    }
    block_max_into(seen, &rmax);
  }
  __syncthreads();
  if (tid == 0) atomicMax(offmax, rmax);
  if (!rotated) return;
  // ---- 3. the rotated columns back; V(:, pair) <- V(:, pair) Qp
  lds_to_cols(lds, ldp, mw, n, W, mw, cols);
  if (V == nullptr) return;
  __syncthreads();
  const int ldv = nw | 1;
  cols_to_lds(V, nw, cols, nw, n, lds, ldv);
  __syncthreads();
  lds_apply_rotation(lds, nw, ldv, n, Qc);
  __syncthreads();
  lds_to_cols(lds, ldv, nw, n, V, nw, cols);
}

// ---- Householder QR (R factor only) for the preconditioning of the Jacobi
// SVD.  One panel of pc <= 32 columns is factorised in LDS by one
// work-group (LAPACK's dgeqr2 reflectors: H = I - tau v v^T, v(0) = 1,
// beta = -sign(x0) ||x||); the panel's R goes back in place, the explicit
// unit-lower V and Y = V T^T (T from the forward compact-WY recurrence of
// dlarft) go to Vb and Yb, so that the trailing columns C are updated by two
// GEMMs: C <- C - Y (V^T C).
constexpr int kQrPanel = kSvdMaxP;

__global__ __launch_bounds__(256) void qr_panel_kernel(int rows, int pc, double* __restrict__ X, int64_t ld,
                                                       double* __restrict__ Vb, double* __restrict__ Yb,
                                                       int64_t sX, int64_t sVY) {
  extern __shared__ double P[];   // panel column c at P + c * lp
  // blockIdx.y: the matrix of a batch (offsets sX into X, sVY into Vb, Yb)
  X += blockIdx.y * sX;
  Vb += blockIdx.y * sVY;
  Yb += blockIdx.y * sVY;
  __shared__ double T[kQrPanel][kQrPanel + 1];
  __shared__ double S[kQrPanel][kQrPanel + 1];
  __shared__ double tau[kQrPanel], red[16];
  __shared__ int colidx[kQrPanel];
  const int tid = threadIdx.x;
  const int lp = rows | 1;
  if (tid < kQrPanel) colidx[tid] = tid;
  __syncthreads();
  cols_to_lds(X, ld, colidx, rows, pc, P, lp);
  __syncthreads();
  const int g = tid >> 4, l16 = tid & 15;
  for (int j = 0; j < pc; ++j) {
    double* x = P + j * lp;
    double s = 0.0;
    for (int i = j + 1 + tid; i < rows; i += blockDim.x) s += x[i] * x[i];
    s = block_sum(s, red);
    const double x0 = (j < rows) ? x[j] : 0.0;
    double t = 0.0, beta = x0, scal = 0.0;
    if (s > 0.0) {
      const double nrm = sqrt(x0 * x0 + s);
      beta = x0 >= 0.0 ? -nrm : nrm;
      t = (beta - x0) / beta;
      scal = 1.0 / (x0 - beta);
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < rows; i += blockDim.x) x[i] *= scal;
    if (tid == 0) {
      if (j < rows) x[j] = beta;
      tau[j] = t;
    }
    __syncthreads();
    if (t != 0.0) {
      // rows four strides at a time (four partial sums), so that a lane's
      // loads are in flight together instead of one round trip per row
      for (int c = j + 1 + g; c < pc; c += 16) {
        double* y = P + c * lp;
        double w0 = (l16 == 0) ? y[j] : 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
        int i = j + 1 + l16;
        for (; i + 48 < rows; i += 64) {
          w0 += x[i] * y[i];
          w1 += x[i + 16] * y[i + 16];
          w2 += x[i + 32] * y[i + 32];
          w3 += x[i + 48] * y[i + 48];
        }
        for (; i < rows; i += 16) w0 += x[i] * y[i];
        double w = row_sum16((w0 + w1) + (w2 + w3));
        w *= t;
        if (l16 == 0) y[j] -= w;
        i = j + 1 + l16;
        for (; i + 48 < rows; i += 64) {
          const double x0 = x[i], x1 = x[i + 16], x2 = x[i + 32], x3 = x[i + 48];
          const double y0 = y[i], y1 = y[i + 16], y2 = y[i + 32], y3 = y[i + 48];
          y[i] = y0 - w * x0;
          y[i + 16] = y1 - w * x1;
          y[i + 32] = y2 - w * x2;
          y[i + 48] = y3 - w * x3;
        }
        for (; i < rows; i += 16) y[i] -= w * x[i];
      }
    }
    __syncthreads();
  }
  // R (upper part of the panel's first pc rows) back in place
  for (int e = tid; e < pc * pc; e += blockDim.x) {
    const int i = e % pc, c = e / pc;
    if (i <= c && i < rows) X[i + ld * c] = P[c * lp + i];
  }
  __syncthreads();
  // explicit unit-lower V in the panel
  for (int c = 0; c < pc; ++c)
    for (int i = tid; i <= c && i < rows; i += blockDim.x) P[c * lp + i] = (i == c) ? 1.0 : 0.0;
  __syncthreads();
  lds_gram(P, rows, lp, pc, S, &T[0][0]);   // S = V^T V (T used as scratch)
  // T: T(j, j) = tau_j, T(0:j, j) = -tau_j T(0:j, 0:j) S(0:j, j)
  for (int e = tid; e < kQrPanel * (kQrPanel + 1); e += blockDim.x) (&T[0][0])[e] = 0.0;
  __syncthreads();
  for (int j = 0; j < pc; ++j) {
    if (tid < j) {
      double acc = 0.0;
      for (int l = tid; l < j; ++l) acc += T[tid][l] * S[l][j];
      T[tid][j] = -tau[j] * acc;
    }
    if (tid == 0) T[j][j] = tau[j];
    __syncthreads();
  }
  for (int c = 0; c < pc; ++c)
    for (int i = tid; i < rows; i += blockDim.x) Vb[i + (int64_t)rows * c] = P[c * lp + i];
  __syncthreads();
  // Y = V T^T: lds_apply_rotation computes X Qp with Qp(k, j) = Qc[j][k] = T[j][k]
  lds_apply_rotation(P, rows, lp, pc, T);
  __syncthreads();
  for (int c = 0; c < pc; ++c)
    for (int i = tid; i < rows; i += blockDim.x) Yb[i + (int64_t)rows * c] = P[c * lp + i];
}

// X (q x q, ld q) = P R^T for the upper triangle R of W (ld ldw): row i of
// R^T goes to row rowperm[i] of X.
__global__ void upper_transpose_kernel(int q, const double* __restrict__ W, int64_t ldw,
                                       const int* __restrict__ rowperm, double* __restrict__ X) {
  const int64_t total = (int64_t)q * q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % q), j = (int)(e / q);
    X[rowperm[i] + (int64_t)q * j] = (j <= i) ? W[j + ldw * i] : 0.0;
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

// out(:, k) = W(:, perm[k]) * scale[k]   (column gather with scaling; no
// scaling when scale is null)
__global__ void gather_scale_cols(int m, int k, const double* __restrict__ W, int64_t ldw,
                                  const int* __restrict__ perm, const double* __restrict__ scale,
                                  double* __restrict__ out, int64_t ldo) {
  int64_t total = (int64_t)m * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e % m), c = (int)(e / m);
    out[i + ldo * c] = W[i + ldw * perm[c]] * (scale ? scale[c] : 1.0);
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

void complete_columns(int m, int k0, int k, double* U, int64_t ldu, hipStream_t st) {
  if (k0 >= k) return;
  TTF_REQUIRE(k <= 4096, "complete_columns: k=%d exceeds 4096", k);
  hipLaunchKernelGGL(complete_basis, dim3(1), dim3(256), 0, st, m, k0, k, U, ldu);
  TTF_HIP(hipGetLastError());
}

// ---- Cholesky QR of nearly orthonormal columns
// G (q x q, ld q) = R^T R by right-looking Cholesky in LDS and the inverse of
// the upper factor: Rinv (q x q, ld q, upper).  LDS holds R in the upper
// triangle of a (q+1) x (q+1) array and column j of R^-1 in its row j+1 (the
// strictly lower part), so both fit at once.  *info = k+1 when the k-th
// pivot is not positive.  One work-group, q <= kCholMax.
constexpr int kCholMax = 128;

__global__ __launch_bounds__(256) void chol_inv_kernel(int q, const double* __restrict__ G,
                                                       double* __restrict__ Rinv, int* __restrict__ info) {
  extern __shared__ double L[];
  __shared__ int fail;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ld = q + 1;
  for (int e = tid; e < q * q; e += nt) {
    const int i = e % q, j = e / q;
    if (i <= j) L[i * ld + j] = 0.5 * (G[i + (int64_t)q * j] + G[j + (int64_t)q * i]);
  }
  if (tid == 0) fail = 0;
  for (int k = 0; k < q; ++k) {
    __syncthreads();
    if (tid == 0) {
      const double d = L[k * ld + k];
      if (d > 0.0) L[k * ld + k] = sqrt(d);
      else fail = k + 1;
    }
    __syncthreads();
    if (fail) break;
    const double rkk = L[k * ld + k];
    for (int j = k + 1 + tid; j < q; j += nt) L[k * ld + j] /= rkk;
    __syncthreads();
    const int w = q - k - 1;
    for (int e = tid; e < w * w; e += nt) {
      const int i = k + 1 + e % w, j = k + 1 + e / w;
      if (i <= j) L[i * ld + j] -= L[k * ld + i] * L[k * ld + j];
    }
  }
  __syncthreads();
  if (fail) {
    if (tid == 0) *info = fail;
    return;
  }
  // column j of X = R^-1 by back substitution, kept in row j+1 of L
  for (int j = tid; j < q; j += nt) {
    double* x = L + (j + 1) * ld;
    x[j] = 1.0 / L[j * ld + j];
    for (int i = j - 1; i >= 0; --i) {
      double acc = 0.0;
      for (int l = i + 1; l <= j; ++l) acc += L[i * ld + l] * x[l];
      x[i] = -acc / L[i * ld + i];
    }
  }
  __syncthreads();
  for (int e = tid; e < q * q; e += nt) {
    const int i = e % q, j = e / q;
    Rinv[e] = (i <= j) ? L[(j + 1) * ld + i] : 0.0;
  }
  if (tid == 0) *info = 0;
}

__global__ void scale_cols_kernel(int m, int k, double* __restrict__ U, int64_t ldu, const double* __restrict__ sc) {
  const int64_t total = (int64_t)m * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % m), c = (int)(e / m);
    U[i + ldu * c] *= sc[c];
  }
}

// Two passes of Cholesky QR on the p x k columns of U (ld ldu), which must
// be nearly orthonormal already; returns false (U partly updated) if a
// Cholesky factorisation breaks down.
bool cholqr2(int p, int k, double* U, int64_t ldu, hipStream_t st) {
  if (k <= 0) return true;
  TTF_REQUIRE(k <= kCholMax, "cholqr2: %d columns exceed %d", k, kCholMax);
  static std::atomic<bool> attr_set{false};   // concurrent first calls only repeat the call
  if (!attr_set) {
    TTF_HIP(hipFuncSetAttribute((const void*)chol_inv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * (kCholMax + 1) * (kCholMax + 1))));
    attr_set = true;
  }
  double* G = (double*)arena(kArenaSvdLong, sizeof(double) * ((size_t)2 * k * k + (size_t)p * k) + 64);
  double* Ri = G + (size_t)k * k;
  double* T = Ri + (size_t)k * k;
  int* info = (int*)(T + (size_t)p * k);
  const size_t lds = sizeof(double) * (size_t)(k + 1) * (k + 1);
  for (int pass = 0; pass < 2; ++pass) {
    if (k <= 64) tsgram(p, k, k, U, ldu, U, ldu, G, k, st);
    else gemm_launch(colmajor<double>('T', 'N', k, k, p, 1.0, U, ldu, U, ldu, 0.0, G, k), st);
    hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(256), lds, st, k, G, Ri, info);
    TTF_HIP(hipGetLastError());
    int hinfo = 0;
    read_back(&hinfo, info, sizeof(int), st);
    if (hinfo != 0) return false;
    gemm_launch(colmajor<double>('N', 'N', p, k, k, 1.0, U, ldu, Ri, k, 0.0, T, p), st);
    TTF_HIP(hipMemcpy2DAsync(U, sizeof(double) * ldu, T, sizeof(double) * p, sizeof(double) * p, k,
                             hipMemcpyDeviceToDevice, st));
  }
  return true;
}

// Columns per block pair of the one-sided Jacobi (TTF_SVD_PAIR overrides,
// 4..32; default 32): smaller pairs give more work-groups per round and
// shorter inner sweeps, but more rounds.
static int svd_pair_columns() {
  static int v = [] {
    const char* e = std::getenv("TTF_SVD_PAIR");
    int x = e ? std::atoi(e) : kSvdMaxP;
    return std::min(std::max(x, 4), (int)kSvdMaxP);
  }();
  return v;
}

// Lanes per column pair in the Jacobi SVD (TTF_SVD_LANES: 16 or 32, the
// default): 32 halves each thread's rows and gives every SIMD two waves, so
// that one wave's dependent steps overlap the other's (a whole wave per
// pair, 64, measured slower).
static int svd_pair_lanes() {
  static const int v = [] {
    const char* e = std::getenv("TTF_SVD_LANES");
    return (e != nullptr && std::atoi(e) == 16) ? 16 : 32;
  }();
  return v;
}

// Hestenes sweeps per pair and round (TTF_SVD_INNER overrides; default 1).
static int svd_inner_sweeps() {
  static int v = [] {
    const char* e = std::getenv("TTF_SVD_INNER");
    int x = e ? std::atoi(e) : 1;
    return std::min(std::max(x, 1), 30);
  }();
  return v;
}

namespace {

struct JacobiResult {
  int sweeps = 0;
  double small2 = 0.0;   // squared rounding floor of the column norms
};

// One-sided Jacobi on the columns of X (mw x nw, ld mw) until a sweep makes
// no rotation; V (nw x nw) accumulates the rotations when not null.
JacobiResult jacobi_columns(int mw, int nw, double* X, double* V, hipStream_t st) {
  JacobiResult res;
  double* nrm = (double*)arena(kArenaSmall, sizeof(double) * (size_t)nw + 64);
  unsigned long long* offmax = (unsigned long long*)(nrm + nw);
  if (V) {
    hipLaunchKernelGGL(set_identity, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)nw * nw, 256), 4096)),
                       dim3(256), 0, st, nw, V);
    TTF_HIP(hipGetLastError());
  }
  // column norms bound the absolute noise floor of the cross products
  hipLaunchKernelGGL(colnorm_kernel, dim3(nw), dim3(256), 0, st, mw, nw, X, nrm);
  std::vector<double> hn(nw);
  read_back(hn.data(), nrm, sizeof(double) * nw, st);
  double fro2 = 0.0;
  for (double v : hn) fro2 += v * v;
  if (!(fro2 > 0.0) || nw < 2) return res;
  const double eps = 2.220446049250313e-16;
  // couplings are measured to about sqrt(mw) eps (dot-product rounding);
  // the tolerance sits well above that noise
  const double tol = std::max(8.0 * std::sqrt((double)mw), 32.0) * eps;
  // squared column norms below (sqrt(mw) eps ||X||_F)^2 are rounding noise
  res.small2 = mw * eps * eps * fro2;
  // block size: the pair's 2b columns (odd stride mw | 1) must fit the LDS budget
  const int ldp = mw | 1;
  int two_b = std::min<int>({kSvdMaxP, kSvdLdsDoubles / ldp, svd_pair_columns()});
  two_b -= (two_b & 1);
  TTF_REQUIRE(two_b >= 2, "jacobi: %d rows exceed the pair LDS budget", mw);
  // at least two non-empty blocks, so that every column pair is coupled
  const int bsz = std::max(1, std::min(two_b / 2, (int)ceil_div(nw, 2)));
  int nb = (int)ceil_div(nw, bsz);
  if (nb % 2) ++nb;
  if (nb < 2) nb = 2;
  const size_t lds = sizeof(double) * (size_t)(2 * bsz) * ldp;
  // registers per thread and column in the Hestenes steps (16 threads a pair)
  const int nr = (int)ceil_div(mw, 16);
  auto kern = nr <= 4 ? svd_pair_kernel<4> : nr <= 8 ? svd_pair_kernel<8>
            : nr <= 16 ? svd_pair_kernel<16> : nr <= 32 ? svd_pair_kernel<32> : svd_pair_kernel<0>;
  if (mw == 64) kern = svd_pair_kernel<4, 16, true>;
  else if (mw == 128) kern = svd_pair_kernel<8, 16, true>;
  else if (mw == 256) kern = svd_pair_kernel<16, 16, true>;
  int threads = 256;
  if (nr <= 16 && svd_pair_lanes() == 32) {
    kern = nr <= 4 ? svd_pair_kernel<2, 32> : nr <= 8 ? svd_pair_kernel<4, 32> : svd_pair_kernel<8, 32>;
    if (mw == 64) kern = svd_pair_kernel<2, 32, true>;
    else if (mw == 128) kern = svd_pair_kernel<4, 32, true>;
    else if (mw == 256) kern = svd_pair_kernel<8, 32, true>;
    threads = 512;
  }
  static std::atomic<bool> attr_set{false};   // concurrent first calls only repeat the call
  if (!attr_set) {
    for (auto f : {svd_pair_kernel<4>, svd_pair_kernel<8>, svd_pair_kernel<16>, svd_pair_kernel<32>,
                   svd_pair_kernel<0>, svd_pair_kernel<2, 32>, svd_pair_kernel<4, 32>, svd_pair_kernel<8, 32>,
                   svd_pair_kernel<2, 32, true>, svd_pair_kernel<4, 32, true>, svd_pair_kernel<8, 32, true>,
                   svd_pair_kernel<4, 16, true>, svd_pair_kernel<8, 16, true>, svd_pair_kernel<16, 16, true>})
      TTF_HIP(hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(double) * kSvdLdsDoubles)));
    attr_set = true;
  }
  const int inner = svd_inner_sweeps();
  unsigned long long hoff = 0;
  // Convergence is quadratic at the end: after a sweep whose largest
  // coupling (measured before its rotations) was at most kQuadStop, the
  // couplings left are of order nw * kQuadStop^2, below tol, so the sweep
  // that would only confirm it is not run.
  const double kQuadStop = std::min(1e-9, 0.1 * std::sqrt(tol / nw));
  // cyclic one-sided Jacobi needs about 40 sweeps on unpreconditioned
  // columns with a 1e-13 graded spectrum (10 after the QR preconditioning)
  for (res.sweeps = 1; res.sweeps <= 60; ++res.sweeps) {
    TTF_HIP(hipMemsetAsync(offmax, 0, sizeof(unsigned long long), st));
    for (int r = 0; r < nb - 1; ++r) {
      hipLaunchKernelGGL(kern, dim3(nb / 2), dim3(threads), lds, st, mw, nw, bsz, nb, r, X, V, res.small2, tol,
                         inner, offmax);
      TTF_HIP(hipGetLastError());
    }
    read_back(&hoff, offmax, sizeof(hoff), st);
    double off;
    std::memcpy(&off, &hoff, sizeof(off));
    if (off <= tol || off <= kQuadStop) break;
  }
  return res;
}

// Sorted column norms of X (mw x k, ld mw): hs (host, descending), the
// permutation (device perm) and scales 1/s (0 below the rounding floor) in
// scale; returns the number of columns above the floor.
int sorted_norms(int mw, int k, const double* X, double small2, std::vector<double>& hs, int* perm,
                 double* scale, hipStream_t st) {
  double* nrm = (double*)arena(kArenaSmall, sizeof(double) * (size_t)k);
  hipLaunchKernelGGL(colnorm_kernel, dim3(k), dim3(256), 0, st, mw, k, X, nrm);
  std::vector<double> hn(k);
  read_back(hn.data(), nrm, sizeof(double) * k, st);
  std::vector<int> order(k);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hn[a] > hn[b]; });
  const double tiny = std::max(std::sqrt(small2), 1e-300);
  hs.assign(k, 0.0);
  std::vector<double> hscale(k);
  int kpos = 0;
  for (int c = 0; c < k; ++c) {
    hs[c] = hn[order[c]];
    hscale[c] = hs[c] > tiny ? 1.0 / hs[c] : 0.0;
    if (hs[c] > tiny) kpos = c + 1;
  }
  TTF_HIP(hipMemcpyAsync(perm, order.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
  TTF_HIP(hipMemcpyAsync(scale, hscale.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
  return kpos;
}

// R factors of nb matrices W_b = W + b sW (p x q, ld p, p >= q, overwritten):
// R_b is left in the upper triangle of W_b(0:q, 0:q).  Panels of pc <= 32
// columns (as many as fit the LDS budget with p rows) in one launch for the
// whole batch, trailing columns updated by two batched GEMMs.
void qr_r_factor_batched(int p, int q, double* W, int nb, int64_t sW, hipStream_t st) {
  const int lp = p | 1;
  const int pcmax = std::min(kQrPanel, kSvdLdsDoubles / lp);
  TTF_REQUIRE(pcmax >= 1, "qr: %d rows exceed the panel LDS budget", p);
  TTF_REQUIRE(nb >= 1 && nb <= 65535, "qr: batch of %d", nb);
  static std::atomic<bool> attr_set{false};   // concurrent first calls only repeat the call
  if (!attr_set) {
    TTF_HIP(hipFuncSetAttribute((const void*)qr_panel_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kSvdLdsDoubles)));
    attr_set = true;
  }
  const int64_t sVY = (int64_t)p * kQrPanel, sZ = (int64_t)kQrPanel * q;
  double* Vb = (double*)arena(kArenaTrunc3, sizeof(double) * (size_t)nb * (2 * sVY + sZ));
  double* Yb = Vb + (size_t)nb * sVY;
  double* Z = Yb + (size_t)nb * sVY;
  for (int k0 = 0; k0 < q; k0 += pcmax) {
    const int rows = p - k0, pc = std::min(pcmax, q - k0);
    const size_t lds = sizeof(double) * (size_t)pc * (rows | 1);
    double* Xp = W + k0 + (int64_t)p * k0;
    hipLaunchKernelGGL(qr_panel_kernel, dim3(1, nb), dim3(256), lds, st, rows, pc, Xp, (int64_t)p, Vb, Yb, sW,
                       sVY);
    TTF_HIP(hipGetLastError());
    const int nrest = q - k0 - pc;
    if (nrest > 0) {
      double* Cm = W + k0 + (int64_t)p * (k0 + pc);
      GemmOp<double> g1 = colmajor<double>('T', 'N', pc, nrest, rows, 1.0, Vb, rows, Cm, p, 0.0, Z, pc);
      g1.nb0 = g1.nbatch = nb;
      g1.sAb0 = sVY; g1.sBb0 = sW; g1.sCb0 = sZ;
      gemm_launch(g1, st);
      GemmOp<double> g2 = colmajor<double>('N', 'N', rows, nrest, pc, -1.0, Yb, rows, Z, pc, 1.0, Cm, p);
      g2.nb0 = g2.nbatch = nb;
      g2.sAb0 = sVY; g2.sBb0 = sZ; g2.sCb0 = sW;
      gemm_launch(g2, st);
    }
  }
}

// Row chunks of X (rows x q, ld ldx) into nb consecutive h x q matrices
// (ld h), the last one padded with zero rows.
__global__ void tsqr_chunks_kernel(int rows, int q, const double* __restrict__ X, int64_t ldx, int h, int nb,
                                   double* __restrict__ C) {
  const int64_t total = (int64_t)nb * h * q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % h), j = (int)((e / h) % q), b = (int)(e / ((int64_t)h * q));
    const int r = b * h + i;
    C[e] = r < rows ? X[r + ldx * j] : 0.0;
  }
}

// The chunks' R factors (upper q x q of each h x q chunk) stacked into S
// ((nb q) x q, ld nb q), zeros below their diagonals.
__global__ void tsqr_stack_kernel(int q, int h, int nb, const double* __restrict__ C, double* __restrict__ S) {
  const int64_t ls = (int64_t)nb * q, total = ls * q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % ls), j = (int)(e / ls), b = r / q, i = r % q;
    S[e] = i <= j ? C[(int64_t)b * h * q + i + (int64_t)h * j] : 0.0;
  }
}

// Upper q x q triangle of S (ld lds) into W (ld ldw).
__global__ void tsqr_copy_r_kernel(int q, const double* __restrict__ S, int64_t lds, double* __restrict__ W,
                                   int64_t ldw) {
  const int64_t total = (int64_t)q * q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % q), j = (int)(e / q);
    if (i <= j) W[i + ldw * j] = S[i + lds * j];
  }
}

// R factor of the QR factorisation of W (p x q, ld p, p >= q, overwritten):
// R is left in the upper triangle of W(0:q, 0:q).  Matrices whose direct
// factorisation would take more than four panels (rows too long for wide
// LDS panels), with q <= 128, take the tall-skinny tree: row chunks of
// h = min(4q, 256) >= 2q rows are factorised as one batch, their R factors
// stacked (at most half as many rows) and the stack factorised again, until
// it is at most 2h rows.  Each level is Householder QR (backward stable);
// the R obtained differs from the direct one only by the signs of its rows.
void qr_r_factor(int p, int q, double* W, hipStream_t st) {
  const int h = std::max(q, std::min(4 * q, 256));
  const int pc_direct = std::max(1, std::min(kQrPanel, kSvdLdsDoubles / (p | 1)));
  if (q > 128 || p <= 2 * h || ceil_div(q, pc_direct) <= 4) {
    qr_r_factor_batched(p, q, W, 1, 0, st);
    return;
  }
  // workspace for all levels: chunks (<= (rows + h) q) and stacks (<= rows q / 2)
  size_t need = 0;
  for (int64_t rows = p; rows > 2 * h;) {
    const int64_t nb = ceil_div(rows, h);
    need += (size_t)(nb * h * q + nb * q * q);
    rows = nb * q;
  }
  double* ws = (double*)arena(kArenaTsqr, sizeof(double) * need);
  const double* cur = W;
  int64_t rows = p, ldc = p;
  while (rows > 2 * h) {
    const int nb = (int)ceil_div(rows, h);
    double* C = ws;
    ws += (size_t)nb * h * q;
    const int64_t tot = (int64_t)nb * h * q;
    hipLaunchKernelGGL(tsqr_chunks_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(tot, 256), 8192)), dim3(256),
                       0, st, (int)rows, q, cur, ldc, h, nb, C);
    TTF_HIP(hipGetLastError());
    qr_r_factor_batched(h, q, C, nb, (int64_t)h * q, st);
    double* S = ws;
    ws += (size_t)nb * q * q;
    const int64_t tots = (int64_t)nb * q * q;
    hipLaunchKernelGGL(tsqr_stack_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(tots, 256), 8192)), dim3(256),
                       0, st, q, h, nb, C, S);
    TTF_HIP(hipGetLastError());
    cur = S;
    rows = (int64_t)nb * q;
    ldc = rows;
  }
  double* last = const_cast<double*>(cur);
  qr_r_factor_batched((int)rows, q, last, 1, 0, st);
  hipLaunchKernelGGL(tsqr_copy_r_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)q * q, 256), 4096)),
                     dim3(256), 0, st, q, last, ldc, W, (int64_t)p);
  TTF_HIP(hipGetLastError());
}

}  // namespace

int gesvd(int m, int n, const double* A, int64_t lda, double* U, int64_t ldu, double* S,
          double* VT, int64_t ldvt, hipStream_t st, int complete) {
  TTF_REQUIRE(m >= 1 && n >= 1, "gesvd: invalid sizes %d x %d", m, n);
  const bool tall = m >= n;
  const int mw = tall ? m : n, nw = tall ? n : m;   // work matrix mw x nw
  const int k = nw;
  TTF_REQUIRE(nw <= 4096, "gesvd: min(m,n)=%d exceeds 4096", nw);
  // arena layout: W (mw*nw) | V (nw*nw) | scale (nw) | perm (nw ints)
  size_t need = sizeof(double) * ((size_t)mw * nw + (size_t)nw * nw + (size_t)nw) + sizeof(int) * (size_t)nw;
  char* base = (char*)arena(kArenaTrunc, need);
  double* W = (double*)base;
  double* V = W + (size_t)mw * nw;
  double* scale = V + (size_t)nw * nw;
  int* perm = (int*)(scale + nw);
  if (tall) {
    TTF_HIP(hipMemcpy2DAsync(W, sizeof(double) * mw, A, sizeof(double) * lda, sizeof(double) * m, n,
                             hipMemcpyDeviceToDevice, st));
  } else {
    transpose(m, n, A, lda, W, mw, st);
  }
  JacobiResult jr = jacobi_columns(mw, nw, W, V, st);
  std::vector<double> hs;
  const int kpos = sorted_norms(mw, nw, W, jr.small2, hs, perm, scale, st);
  TTF_HIP(hipMemcpyAsync(S, hs.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
  // normalised left factor of W and the permuted V
  double* Uw = tall ? U : (double*)arena(kArenaTrunc2, sizeof(double) * (size_t)mw * k);
  const int64_t lduw = tall ? ldu : mw;
  int blocks = (int)std::min<int64_t>(ceil_div((int64_t)mw * k, 256), 8192);
  hipLaunchKernelGGL(gather_scale_cols, dim3(blocks), dim3(256), 0, st, mw, k, W, (int64_t)mw, perm,
                     scale, Uw, lduw);
  TTF_HIP(hipGetLastError());
  if (kpos < k && complete) complete_columns(mw, kpos, k, Uw, lduw, st);
  // V permuted (scale 1): Vp(:, c) = V(:, order[c])
  double* Vp = (double*)arena(kArenaTrunc3, sizeof(double) * (size_t)nw * nw);
  std::vector<double> hone(nw, 1.0);
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
  return jr.sweeps;
}

void orth_factor(int m, int n, const double* A, double* Q, double* R, double* S, hipStream_t st) {
  const int k = std::min(m, n);
  std::vector<double> hs;
  int kpos = 0;
  svd_vectors(m, n, A, m, true, S, Q, m, hs, &kpos, st);
  complete_columns(m, kpos, k, Q, m, st);
  GemmOp<double> g = colmajor<double>('T', 'N', k, n, m, 1.0, Q, m, A, m, 0.0, R, k);
  gemm_launch(g, st);
}

namespace {
// X(n x k) = the first k columns of the n x n identity (ld n)
__global__ void identity_cols_kernel(int n, int k, double* __restrict__ X) {
  const int64_t total = (int64_t)n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    X[e] = (e % n == e / n) ? 1.0 : 0.0;
}

// R (k x m, ld ldr) = upper triangle of W (n x m, ld n), zero below
__global__ void upper_copy_kernel(int k, int m, const double* __restrict__ W, int64_t ldw, double* __restrict__ R,
                                  int64_t ldr) {
  const int64_t total = (int64_t)k * m;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % k), j = (int)(e / k);
    R[i + ldr * j] = (i <= j) ? W[i + ldw * j] : 0.0;
  }
}
}  // namespace

void householder_qr(int n, int m, const double* A, double* Q, double* R, int64_t ldr, hipStream_t st) {
  TTF_REQUIRE(n >= 1 && m >= 1, "householder_qr: invalid sizes %d x %d", n, m);
  const int k = std::min(n, m);
  const int lp = n | 1;
  const int pcmax = std::min(kQrPanel, kSvdLdsDoubles / lp);
  TTF_REQUIRE(pcmax >= 1, "householder_qr: %d rows exceed the panel LDS budget", n);
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    TTF_HIP(hipFuncSetAttribute((const void*)qr_panel_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kSvdLdsDoubles)));
    attr_set = true;
  }
  // workspace: W (n x m) | every panel's V and Y (rows x pc each, <= n x k) | Z (pc x max(m, k))
  const size_t vy = (size_t)n * k;
  double* W = (double*)arena(kArenaTrunc2, sizeof(double) * ((size_t)n * m + 2 * vy + (size_t)kQrPanel * std::max(m, k)));
  double* Vall = W + (size_t)n * m;
  double* Yall = Vall + vy;
  double* Z = Yall + vy;
  TTF_HIP(hipMemcpyAsync(W, A, sizeof(double) * (size_t)n * m, hipMemcpyDeviceToDevice, st));
  std::vector<int64_t> off;   // panel offsets into Vall / Yall
  int64_t used = 0;
  for (int k0 = 0; k0 < k; k0 += pcmax) {
    const int rows = n - k0, pc = std::min(pcmax, k - k0);
    double* Vb = Vall + used;
    double* Yb = Yall + used;
    off.push_back(used);
    used += (int64_t)rows * pc;
    const size_t lds = sizeof(double) * (size_t)pc * (rows | 1);
    hipLaunchKernelGGL(qr_panel_kernel, dim3(1, 1), dim3(256), lds, st, rows, pc, W + k0 + (int64_t)n * k0,
                       (int64_t)n, Vb, Yb, (int64_t)0, (int64_t)0);
    TTF_HIP(hipGetLastError());
    const int nrest = m - k0 - pc;
    if (nrest > 0) {
      // trailing columns: C <- H^T C = C - Y (V^T C)
      double* Cm = W + k0 + (int64_t)n * (k0 + pc);
      gemm_launch(colmajor<double>('T', 'N', pc, nrest, rows, 1.0, Vb, rows, Cm, n, 0.0, Z, pc), st);
      gemm_launch(colmajor<double>('N', 'N', rows, nrest, pc, -1.0, Yb, rows, Z, pc, 1.0, Cm, n), st);
    }
  }
  hipLaunchKernelGGL(upper_copy_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)k * m, 256), 8192)),
                     dim3(256), 0, st, k, m, W, (int64_t)n, R, ldr);
  TTF_HIP(hipGetLastError());
  // Q = H_1 ... H_P I(:, 0:k): panels applied last to first, each as
  // (I - V T V^T) X = X - V (Y^T X) with Y = V T^T, on rows and columns k0:
  hipLaunchKernelGGL(identity_cols_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)n * k, 256), 8192)),
                     dim3(256), 0, st, n, k, Q);
  TTF_HIP(hipGetLastError());
  for (int p = (int)off.size() - 1; p >= 0; --p) {
    const int k0 = p * pcmax, rows = n - k0, pc = std::min(pcmax, k - k0), cols = k - k0;
    const double* Vb = Vall + off[p];
    const double* Yb = Yall + off[p];
    double* Qs = Q + k0 + (int64_t)n * k0;
    gemm_launch(colmajor<double>('T', 'N', pc, cols, rows, 1.0, Yb, rows, Qs, n, 0.0, Z, pc), st);
    gemm_launch(colmajor<double>('N', 'N', rows, cols, pc, -1.0, Vb, rows, Z, pc, 1.0, Qs, n), st);
  }
}

int svd_vectors(int m, int n, const double* A, int64_t lda, bool want_left, double* S, double* O,
                int64_t ldo, std::vector<double>& hs, int* kpos, hipStream_t st) {
  TTF_REQUIRE(m >= 1 && n >= 1, "svd_vectors: invalid sizes %d x %d", m, n);
  const bool tall = m >= n;
  const int p = tall ? m : n, q = tall ? n : m;     // T = A or A^T, p x q
  TTF_REQUIRE(q <= 4096, "svd_vectors: min(m,n)=%d exceeds 4096", q);
  const bool short_side = want_left ? !tall : tall;  // vectors of dimension q
  const bool precond = short_side && (p | 1) <= kSvdLdsDoubles;
  if (!short_side && q <= kCholMax && (p | 1) <= kSvdLdsDoubles && q < p) {
    // vectors of the long side from the short side's: with T = A (tall) or
    // A^T (wide) and T^T T = Vs S^2 Vs^T, the long-side vectors are
    // T Vs S^-1 for the nonzero values, orthonormalised again by Cholesky QR
    // (in column order, so the leading columns span the dominant subspace)
    double* Vs = (double*)arena(kArenaSvdLong2, sizeof(double) * ((size_t)q * q + (size_t)q));
    double* sc = Vs + (size_t)q * q;
    const int sw = svd_vectors(m, n, A, lda, !want_left, S, Vs, q, hs, kpos, st);
    if (tall) gemm_launch(colmajor<double>('N', 'N', m, q, n, 1.0, A, lda, Vs, q, 0.0, O, ldo), st);
    else gemm_launch(colmajor<double>('T', 'N', n, q, m, 1.0, A, lda, Vs, q, 0.0, O, ldo), st);
    std::vector<double> hsc(q, 0.0);
    for (int c = 0; c < *kpos; ++c) hsc[c] = 1.0 / hs[c];
    TTF_HIP(hipMemcpyAsync(sc, hsc.data(), sizeof(double) * q, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)p * q, 256), 8192)),
                       dim3(256), 0, st, p, q, O, ldo, sc);
    TTF_HIP(hipGetLastError());
    if (cholqr2(p, *kpos, O, ldo, st)) return sw;
    // breakdown (not expected for these nearly orthonormal columns): the
    // one-sided Jacobi on the long columns below
  }
  if (short_side && !precond) {
    // very long rows: the plain SVD, then the wanted factor
    double* Uf = (double*)arena(kArenaDriver, sizeof(double) * ((size_t)m * q + (size_t)q * n));
    double* VTf = Uf + (size_t)m * q;
    gesvd(m, n, A, lda, Uf, m, S, VTf, q, st, 0);
    hs.assign(q, 0.0);
    read_back(hs.data(), S, sizeof(double) * q, st);
    if (want_left) {
      TTF_HIP(hipMemcpy2DAsync(O, sizeof(double) * ldo, Uf, sizeof(double) * m, sizeof(double) * m, q,
                               hipMemcpyDeviceToDevice, st));
    } else {
      transpose(q, n, VTf, q, O, ldo, st);
    }
    int kp = 0;
    for (int c = 0; c < q; ++c) if (hs[c] > 0.0) kp = c + 1;
    *kpos = kp;
    return 0;
  }
  size_t need = sizeof(double) * ((size_t)p * q + (size_t)q * q + (size_t)q) + sizeof(int) * (size_t)q;
  char* base = (char*)arena(kArenaTrunc, need);
  double* W = (double*)base;
  double* Xr = W + (size_t)p * q;
  double* scale = Xr + (size_t)q * q;
  int* perm = (int*)(scale + q);
  if (tall) {
    TTF_HIP(hipMemcpy2DAsync(W, sizeof(double) * p, A, sizeof(double) * lda, sizeof(double) * m, n,
                             hipMemcpyDeviceToDevice, st));
  } else {
    transpose(m, n, A, lda, W, p, st);
  }
  double* X = W;
  int mw = p;
  if (precond) {
    // T P = Q R with P ordering T's columns by decreasing norm; the left
    // singular vectors of P R^T are T's right singular vectors.  The order
    // matters for unfoldings whose column index runs over several graded
    // blocks (an interface index weighted by its singular values, times a
    // mode index): in the order of decreasing norms R is graded too and the
    // Jacobi iteration on R^T needs about half the sweeps.
    double* nrm = (double*)arena(kArenaSmall, sizeof(double) * (size_t)q);
    hipLaunchKernelGGL(colnorm_kernel, dim3(q), dim3(256), 0, st, p, q, W, nrm);
    TTF_HIP(hipGetLastError());
    std::vector<double> hn(q);
    read_back(hn.data(), nrm, sizeof(double) * q, st);
    std::vector<int> order(q);
    for (int c = 0; c < q; ++c) order[c] = c;
    for (double& v : hn) v = v >= 0.0 ? v : -1.0;   // NaN last, keeps the order strict
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hn[a] > hn[b]; });
    TTF_HIP(hipMemcpyAsync(perm, order.data(), sizeof(int) * q, hipMemcpyHostToDevice, st));
    double* Wp = (double*)arena(kArenaTrunc2, sizeof(double) * (size_t)p * q);
    const int gblocks = (int)std::min<int64_t>(ceil_div((int64_t)p * q, 256), 8192);
    hipLaunchKernelGGL(gather_scale_cols, dim3(gblocks), dim3(256), 0, st, p, q, W, (int64_t)p, perm,
                       (const double*)nullptr, Wp, (int64_t)p);
    TTF_HIP(hipGetLastError());
    qr_r_factor(p, q, Wp, st);
    hipLaunchKernelGGL(upper_transpose_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)q * q, 256), 4096)),
                       dim3(256), 0, st, q, Wp, (int64_t)p, (const int*)perm, Xr);
    TTF_HIP(hipGetLastError());
    X = Xr;
    mw = q;
  }
  JacobiResult jr = jacobi_columns(mw, q, X, nullptr, st);
  *kpos = sorted_norms(mw, q, X, jr.small2, hs, perm, scale, st);
  TTF_HIP(hipMemcpyAsync(S, hs.data(), sizeof(double) * q, hipMemcpyHostToDevice, st));
  const int blocks = (int)std::min<int64_t>(ceil_div((int64_t)mw * q, 256), 8192);
  hipLaunchKernelGGL(gather_scale_cols, dim3(blocks), dim3(256), 0, st, mw, q, X, (int64_t)mw, perm, scale,
                     O, ldo);
  TTF_HIP(hipGetLastError());
  return jr.sweeps;
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

int ttf_dsvd_side(int m, int n, const double* A, int64_t lda, char side, double* S, double* O,
                  int64_t ldo, int* kpos, int* sweeps) {
  return guarded([&] {
    TTF_REQUIRE(side == 'L' || side == 'l' || side == 'R' || side == 'r', "ttf_dsvd_side: side '%c'", side);
    const bool left = side == 'L' || side == 'l';
    TTF_REQUIRE(ldo >= (left ? m : n), "ttf_dsvd_side: ldo=%lld too small", (long long)ldo);
    std::vector<double> hs;
    int kp = 0;
    int sw = svd_vectors(m, n, A, lda, left, S, O, ldo, hs, &kp, stream());
    TTF_HIP(hipStreamSynchronize(stream()));
    if (kpos) *kpos = kp;
    if (sweeps) *sweeps = sw;
  });
}

int ttf_dqr(int n, int m, double* A, double* R, int64_t ldr) {
  return guarded([&] {
    TTF_REQUIRE(n >= 1 && m >= 1 && ldr >= std::min(n, m), "ttf_dqr: invalid sizes n=%d m=%d ldr=%lld", n, m,
                (long long)ldr);
    hipStream_t st = stream();
    const int k = std::min(n, m);
    double* Q = (double*)arena(kArenaTrunc3, sizeof(double) * (size_t)n * k);
    householder_qr(n, m, A, Q, R, ldr, st);
    TTF_HIP(hipMemcpyAsync(A, Q, sizeof(double) * (size_t)n * k, hipMemcpyDeviceToDevice, st));
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

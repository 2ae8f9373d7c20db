// Device-resident block eigensolver for the local problems of tt_eigb.
//
// The reference hands each local problem to PRIMME with a matvec callback
// (tt_eigb.f90:251-283; method DYNAMIC, maxBasis 100, minRestart B+1,
// maxMatvecs 500, eps = eps2/resid_damp).  Here the whole solve stays on the
// device: LOBPCG with an orthonormal basis S = [X W P] (one of the methods
// PRIMME itself offers; the reference has it wired in, commented, at
// tt_eigb.f90:254), block size nev + guard vectors.
//
// Per iteration: residuals W = AX - X diag(theta); convergence test
// ||r_i|| <= tol * anorm on the host (the only host synchronisation);
// W orthogonalised twice against [X P] and orthonormalised from W^T W by a
// pivoted Cholesky factorisation (or SVQB, its eigen-decomposition; near-null
// directions dropped either way); one block
// matvec AW; Rayleigh-Ritz on S^T A S in a single work-group (cyclic Jacobi
// in LDS), which also builds the next search direction P in the small
// space, orthogonal to the new Ritz vectors; then [X P] <- S [C Z].
// Dropped (zero) columns carry a large diagonal penalty in the projected
// matrix so they can never produce a spurious Ritz value.
#include "amen.h"
#include "solvers.h"
#include "jacobi.h"
#include "wave.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace ttf {

namespace {

constexpr int kMaxM = 64;   // largest projected basis (3 * block size)
constexpr double kEps = 2.220446049250313e-16;

// SVQB: G (k x k) = W^T W -> T = U diag(lambda^-1/2) over directions with
// lambda > 1e-14 lambda_max (others get a zero column and mask 0).
__global__ __launch_bounds__(256) void svqb_kernel(int k, const double* __restrict__ G,
                                                   double* __restrict__ T, int* __restrict__ mask) {
  __shared__ double A[kMaxM][kMaxM + 1], B[kMaxM][kMaxM + 1], V[kMaxM][kMaxM + 1];
  __shared__ double dsc[kMaxM], red[16];
  __shared__ int flags[3];
  double f2 = 0.0;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    int i = e % k, j = e / k;
    double gi = G[i + k * i], gj = G[j + k * j];
    // scale to unit diagonal first (the "S" of SVQB) for a well-conditioned eig
    double di = gi > 0 ? 1.0 / sqrt(gi) : 0.0, dj = gj > 0 ? 1.0 / sqrt(gj) : 0.0;
    double a = 0.5 * (G[i + k * j] + G[j + k * i]) * di * dj;
    A[i][j] = a;
    V[i][j] = (i == j) ? 1.0 : 0.0;
    f2 += a * a;
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    double gi = G[i + k * i];
    dsc[i] = gi > 0 ? 1.0 / sqrt(gi) : 0.0;
  }
  f2 = block_sum(f2, red);
  double(*D)[kMaxM + 1] = jacobi_sym<kMaxM + 1>(A, B, V, k, 40, kEps * sqrt(f2), flags);
  double lmax = 0.0;
  for (int i = 0; i < k; ++i) lmax = fmax(lmax, D[i][i]);
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    int i = e % k, j = e / k;
    double l = D[j][j];
    bool keep = l > 1e-14 * lmax && lmax > 0.0;
    T[i + k * j] = keep ? dsc[i] * V[i][j] / sqrt(l) : 0.0;
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    double l = D[j][j];
    mask[j] = (l > 1e-14 * lmax && lmax > 0.0) ? 1 : 0;
  }
}

// The same contract by pivoted Cholesky, in one wave: with the unit-diagonal
// A = D^-1/2 G D^-1/2 (D = diag G), P^T A P = L L^T with the largest
// remaining diagonal as pivot, stopped at the first pivot below 1e-14 (rank
// r); T = D^-1/2 P L^-T on the r kept directions, so that W = Wr T has
// T^T G T = I, and zero columns (mask 0) after them.  W spans what SVQB's
// W spans, to rounding, and the Rayleigh-Ritz step depends on the span
// only; the factorisation takes k short steps where the eigen-decomposition
// takes several Jacobi sweeps.
// With q > 0, G (ld ldg) is [Q Wr]^T Wr for Q with q orthonormal (or zero)
// columns, and the factorised Gram is that of Wr - Q (Q^T Wr), which is
// Wr^T Wr - (Q^T Wr)^T (Q^T Wr): the projection's second pass and the
// orthonormalisation then share one pass over the vectors.  A direction
// whose squared norm falls below 1e-8 of Wr's own in that subtraction has
// lost its relative accuracy to cancellation and is dropped.
constexpr int kMaxB = 32;   // block columns handled by cholqb_kernel
__global__ __launch_bounds__(64) void cholqb_kernel(int k, const double* __restrict__ G, int ldg, int q,
                                                    double* __restrict__ T, int* __restrict__ mask) {
  __shared__ double A[kMaxB][kMaxB + 1];   // Schur complements (permuted order)
  __shared__ double L[kMaxB][kMaxB + 1];   // Cholesky factor (permuted order)
  __shared__ double Gs[3 * kMaxB][kMaxB + 1];   // G (q + k rows)
  __shared__ double sc[kMaxB], lc[kMaxB];
  __shared__ int perm[kMaxB];
  __shared__ int piv, rank_;
  const int lane = threadIdx.x;
  // G into LDS, four loads in flight per lane (clamped in-range addresses)
  {
    const int rows = q + k, tot = rows * k;
    for (int e0 = 0; e0 < tot; e0 += 4 * 64) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + lane + 64 * u, tot - 1);
        v[u] = G[(e % rows) + (int64_t)ldg * (e / rows)];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + lane + 64 * u;
        if (e < tot) Gs[e % rows][e / rows] = v[u];
      }
    }
  }
  __syncthreads();
  // the block's Gram into L (scratch until the factorisation)
  for (int e = lane; e < k * k; e += 64) {
    const int i = e % k, j = e / k;
    double g = Gs[q + i][j];
    for (int l = 0; l < q; ++l) g -= Gs[l][i] * Gs[l][j];
    L[i][j] = g;
  }
  __syncthreads();
  for (int i = lane; i < k; i += 64) {
    const double g = L[i][i];
    sc[i] = (g > 0.0 && g >= 1e-8 * Gs[q + i][i]) ? 1.0 / sqrt(g) : 0.0;
    perm[i] = i;
  }
  if (lane == 0) rank_ = k;
  __syncthreads();
  for (int e = lane; e < k * k; e += 64) {
    const int i = e % k, j = e / k;
    A[i][j] = 0.5 * (L[i][j] + L[j][i]) * sc[i] * sc[j];
  }
  __syncthreads();
  for (int e = lane; e < k * k; e += 64) L[e % k][e / k] = 0.0;
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (lane == 0) {
      int p = j;
      double best = A[j][j];
      for (int i = j + 1; i < k; ++i)
        if (A[i][i] > best) { best = A[i][i]; p = i; }
      if (!(best > 1e-14)) rank_ = j;   // also stops on NaN
      piv = p;
    }
    __syncthreads();
    if (rank_ <= j) break;   // the same shared value for every lane
    const int p = piv;
    if (p != j) {   // swap indices p and j: A's rows and columns, L's rows, perm
      for (int i = lane; i < k; i += 64) {
        const double t = A[i][p]; A[i][p] = A[i][j]; A[i][j] = t;
      }
      __syncthreads();
      for (int i = lane; i < k; i += 64) {
        const double t = A[p][i]; A[p][i] = A[j][i]; A[j][i] = t;
      }
      for (int c = lane; c < j; c += 64) {
        const double t = L[p][c]; L[p][c] = L[j][c]; L[j][c] = t;
      }
      if (lane == 0) { const int t = perm[p]; perm[p] = perm[j]; perm[j] = t; }
      __syncthreads();
    }
    const double ljj = sqrt(A[j][j]);
    for (int i = j + lane; i < k; i += 64) {
      const double l = i == j ? ljj : A[i][j] / ljj;
      lc[i] = l;
      L[i][j] = l;
    }
    __syncthreads();
    const int t = k - j - 1;
    for (int e = lane; e < t * t; e += 64) {
      const int i = j + 1 + e % t, c = j + 1 + e / t;
      A[i][c] -= lc[i] * lc[c];
    }
    __syncthreads();
  }
  const int r = rank_;
  // column c of L^-T by back substitution (lane c), scattered to the
  // original row order with the scaling
  for (int c = lane; c < k; c += 64) {
    if (c < r) {
      A[c][c] = 1.0 / L[c][c];   // A's upper triangle now holds L^-T
      for (int i = c - 1; i >= 0; --i) {
        double s = 0.0;
        for (int l = i + 1; l <= c; ++l) s += L[l][i] * A[l][c];
        A[i][c] = -s / L[i][i];
      }
    }
    for (int i = 0; i < k; ++i) {
      const int o = perm[i];
      T[o + k * c] = (c < r && i <= c) ? sc[o] * A[i][c] : 0.0;
    }
    mask[c] = c < r ? 1 : 0;
  }
}

// Orthonormalisation of the residual block (TTF_EIG_ORTH: "chol", the
// default, or "svqb").
static bool eig_orth_svqb() {
  static const bool v = [] {
    const char* e = std::getenv("TTF_EIG_ORTH");
    return e != nullptr && std::string(e) == "svqb";
  }();
  return v;
}

// G: the k x k Gram (ld k), or with q > 0 (the Cholesky form only) the
// (q + k) x k product [Q Wr]^T Wr (ld ldg) of cholqb_kernel.
void orth_block(int k, const double* G, double* T, int* mask, hipStream_t st, int ldg = 0, int q = 0) {
  if (eig_orth_svqb()) {
    TTF_REQUIRE(q == 0, "block_eig: the SVQB orthonormalisation takes a plain Gram");
    hipLaunchKernelGGL(svqb_kernel, dim3(1), dim3(256), 0, st, k, G, T, mask);
  } else {
    TTF_REQUIRE(k <= kMaxB && q <= 2 * kMaxB, "block_eig: block of %d columns exceeds %d", k, kMaxB);
    hipLaunchKernelGGL(cholqb_kernel, dim3(1), dim3(64), 0, st, k, G, q > 0 ? ldg : k, q, T, mask);
  }
  TTF_HIP(hipGetLastError());
}

// Rayleigh-Ritz in the small space.  H (m x m) = S^T A S for the basis
// S = [X P W] (nx, np, np columns, m = nx + 2 np); columns [0, nx) are the
// current Ritz vectors X, the validity of the P and W columns is given by
// maskP_in and maskW.  Writes theta (bs), CZ (m x 2bs: Ritz coefficients C,
// then P-direction coefficients Z), maskP (bs, validity of the new P) and
// amax = max |valid Ritz value|.
// The LDS arrays hold MR rows (m + 1 for the odd-order pad) with row stride
// LD: rr_kernel<33, 34> serves m <= 32 (the even stride keeps each
// thread's pair of Jacobi block entries 16-byte aligned for one load; the
// 24x24 Jacobi measured 12% faster than with the stride 65) and
// rr_kernel<kMaxM, kMaxM + 1> the rest (launch_rr picks).
template <int MR, int LD>
__global__ __launch_bounds__(256) void rr_kernel(int m, int bs, int nx, const double* __restrict__ H,
                                                 const int* __restrict__ maskP_in, const int* __restrict__ maskW,
                                                 double* __restrict__ theta, double* __restrict__ CZ,
                                                 int* __restrict__ maskP, double* __restrict__ amax,
                                                 int lower = 0) {
  __shared__ double A[MR][LD], V[MR][LD], Y[MR][LD];
  __shared__ double Z[MR][MR / 2 + 1];
  __shared__ double w[MR], dg[MR], red[16];
  __shared__ int flags[3];
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int e = tid; e < m * m; e += blockDim.x) mx = fmax(mx, fabs(H[e]));
  mx = block_max(mx, red);
  const double big = 1e6 * (1.0 + mx);
  // the rounding floor of the rotations is set by the valid block only: the
  // penalty rows are exactly decoupled and never rotate
  // basis columns [X (nx) | P (np) | W (np)]: X always valid, P and W by mask
  const int np = (m - nx) / 2;
  auto colok = [&](int i) {
    return i < nx ? 1 : (i < nx + np ? maskP_in[i - nx] : maskW[i - nx - np]);
  };
  // H(i, j) = S_i^T (A S_j), symmetrised.  With `lower` (A W from the
  // fp32 operator) each coupling off the diagonal blocks is taken from the
  // product of the earlier group instead (X before P before W): the X-W
  // coupling W^T (A X), of the size of the residual, then never comes from
  // A W, whose rounding would swamp it once the residual is that small.
  auto grp = [&](int i) { return i < nx ? 0 : (i < nx + np ? 1 : 2); };
  double f2 = 0.0;
  for (int e = tid; e < m * m; e += blockDim.x) {
    int i = e % m, j = e / m;
    bool ok = colok(i) && colok(j);
    const int gi = grp(i), gj = grp(j);
    double h = (!lower || gi == gj) ? 0.5 * (H[i + m * j] + H[j + m * i])
                                    : (gi > gj ? H[i + m * j] : H[j + m * i]);
    A[i][j] = ok ? h : ((i == j) ? big : 0.0);
    V[i][j] = (i == j) ? 1.0 : 0.0;
    if (ok) f2 += h * h;
  }
  f2 = block_sum(f2, red);
  // Y is the second Jacobi buffer; the diagonal is copied out before Y is
  // reused for the sorted vectors
  // A sweep whose rotations all have |sin| <= 1e-7 leaves off-diagonal
  // entries of order 1e-14 ||H||: Ritz values within that of H's
  // eigenvalues, far inside the solver's tolerance, and the next iteration's
  // Rayleigh-Ritz starts again from the exact projection.
  double(*D)[LD] = jacobi_sym<LD>(A, Y, V, m, 40, kEps * sqrt(f2), flags, 1e-7);
  for (int i = tid; i < m; i += blockDim.x) dg[i] = D[i][i];
  __syncthreads();
  // ascending order
  for (int i = tid; i < m; i += blockDim.x) {
    double wi = dg[i];
    int pos = 0;
    for (int j = 0; j < m; ++j) {
      double wj = dg[j];
      pos += (wj < wi || (wj == wi && j < i)) ? 1 : 0;
    }
    w[pos] = wi;
    for (int r = 0; r < m; ++r) Y[r][pos] = V[r][i];
  }
  __syncthreads();
  if (tid == 0) {
    double mx = 0.0;
    for (int i = 0; i < m; ++i)
      if (w[i] < 0.5 * big) mx = fmax(mx, fabs(w[i]));
    *amax = mx;
  }
  for (int i = tid; i < bs; i += blockDim.x) theta[i] = w[i];
  // Z0 = C with the X rows removed, then orthogonalised against C
  for (int e = tid; e < m * bs; e += blockDim.x) {
    int r = e % m, c = e / m;
    Z[r][c] = (r < nx) ? 0.0 : Y[r][c];
  }
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    // Z -= C (C^T Z): coefficient matrix computed per (i, c)
    __shared__ double P[MR / 2][MR / 2 + 1];
    for (int e = tid; e < bs * bs; e += blockDim.x) {
      int i = e % bs, c = e / bs;
      double s = 0.0;
      for (int r = 0; r < m; ++r) s += Y[r][i] * Z[r][c];
      P[i][c] = s;
    }
    __syncthreads();
    for (int e = tid; e < m * bs; e += blockDim.x) {
      int r = e % m, c = e / m;
      double s = 0.0;
      for (int i = 0; i < bs; ++i) s += Y[r][i] * P[i][c];
      Z[r][c] -= s;
    }
    __syncthreads();
  }
  // modified Gram-Schmidt on the columns of Z (twice) by the first wave,
  // lane r owning row r (m <= 64): dot products are wave sums in registers
  // (DPP rows, then lane reads) and each lane touches only its own row, so
  // the wave needs no barrier.  Lanes past m use row m, a spare row (m < MR)
  // that nothing else reads, and hold zeros: the loads and stores carry no
  // per-lane condition (which would make each a branch with a full wait),
  // and the column being orthogonalised stays in a register.
  if (tid < 64) {
    const bool live = tid < m;
    const int row = live ? tid : m;
    for (int pass = 0; pass < 2; ++pass) {
      for (int c = 0; c < bs; ++c) {
        const double zc0 = Z[row][c];
        double zc = live ? zc0 : 0.0;
        for (int j = 0; j < c; ++j) {
          const double zj0 = Z[row][j];
          const double zj = live ? zj0 : 0.0;
          const double s = wave_sum64(zj * zc);
          zc -= s * zj;
        }
        const double s = wave_sum64(zc * zc);
        double nrm = sqrt(s);
        double inv = nrm > 1e-10 ? 1.0 / nrm : 0.0;
        Z[row][c] = zc * inv;
        if (tid == 0 && pass == 1) maskP[c] = nrm > 1e-10 ? 1 : 0;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < m * 2 * bs; e += blockDim.x) {
    int r = e % m, c = e / m;
    CZ[r + m * c] = (c < bs) ? Y[r][c] : Z[r][c - bs];
  }
}

// Rayleigh-Ritz launch on the basis order m (arguments of rr_kernel after m).
static void launch_rr(hipStream_t st, int m, int bs, int nx, const double* H, const int* maskP_in, const int* maskW,
                      double* theta, double* CZ, int* maskP, double* amax, int lower) {
  if (m <= 32)
    hipLaunchKernelGGL((rr_kernel<33, 34>), dim3(1), dim3(256), 0, st, m, bs, nx, H, maskP_in, maskW, theta, CZ,
                       maskP, amax, lower);
  else
    hipLaunchKernelGGL((rr_kernel<kMaxM, kMaxM + 1>), dim3(1), dim3(256), 0, st, m, bs, nx, H, maskP_in, maskW,
                       theta, CZ, maskP, amax, lower);
}

// W = AX - X diag(theta); each work-group takes a range of rows and stores
// its per-column sums of squares to part (nparts x bs) ...
// With D (the local matrix's diagonal), W holds the preconditioned residual
// (D - theta_c)^-1 r_c (Davidson's diagonal preconditioner; |D - theta_c|
// clamped below at delta (1 + |theta_c|)), while the sums of squares, and so
// the convergence test, are of the plain residual.
__global__ __launch_bounds__(256) void residual_kernel(int N, int bs, const double* __restrict__ X,
                                                       const double* __restrict__ AX, int64_t ld,
                                                       const double* __restrict__ theta, double* __restrict__ W,
                                                       double* __restrict__ part, const double* __restrict__ D,
                                                       double delta) {
  __shared__ double red[4][kMaxM];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int per = (N + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * per, r1 = min(N, r0 + per);
  // eight columns at a time, each row's loads of all eight issued together
  for (int c0 = 0; c0 < bs; c0 += 8) {
    double s[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] = 0.0;
    for (int i = r0 + tid; i < r1; i += 256) {
      double ax[8], xx[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = c0 + u;
        ax[u] = c < bs ? AX[i + ld * c] : 0.0;
        xx[u] = c < bs ? X[i + ld * c] : 0.0;
      }
      const double di = D ? D[i] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = c0 + u;
        if (c < bs) {
          const double th = theta[c];
          const double w = ax[u] - th * xx[u];
          if (D) {
            const double floor_c = delta * (1.0 + fabs(th));
            double den = di - th;
            if (fabs(den) < floor_c) den = den >= 0.0 ? floor_c : -floor_c;
            W[i + ld * c] = w / den;
          } else {
            W[i + ld * c] = w;
          }
          s[u] += w * w;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (c0 + u < bs) {   // uniform
        const double t = wave_sum64(s[u]);
        if (lane == 0) red[wv][c0 + u] = t;
      }
    }
  }
  __syncthreads();
  if (tid < bs) part[(int64_t)blockIdx.x * bs + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

// ... and the residual norms nrm[c] = sqrt(sum of the partials): the 16
// lanes of a DPP row take interleaved partials of column c and a row sum
// combines them (a fixed order).
__global__ __launch_bounds__(256) void residual_norms(int bs, int nparts, const double* __restrict__ part,
                                                      double* __restrict__ nrm) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int c = t / 16, k = t % 16;
  // nparts <= 256: at most 16 partials per lane, all loaded before the sum
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int w = k + 16 * u;
    v[u] = (c < bs && w < nparts) ? part[(int64_t)w * bs + c] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < 16; ++u) s += v[u];
  s = row_sum16(s);
  if (k == 0 && c < bs) nrm[c] = sqrt(s);
}

__global__ void random_cols(int N, int c0, int c1, double* X, int64_t ld, unsigned seed) {
  int64_t total = (int64_t)N * (c1 - c0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int c = c0 + (int)(e / N);
    int64_t i = e % N;
    unsigned h = (unsigned)(i * 2654435761u) ^ (seed + 0x9e3779b9u * (unsigned)(c + 1));
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
    X[i + ld * c] = (double)(h & 0xffffff) / (double)0x1000000 - 0.5;
  }
}

// Diagonal blocks of the local matrix over the left interface and the mode
// (the 'l' blocks of tt_linalg.f90:434-462), shifted: block b2 < r2 has
// order nb = r1 n and entries
//   M((c1, i), (b1, j)) = sum_a1 Phi1(c1, b1, a1) sum_a2 A(a1, i, j, a2) Phi2(b2, a2, b2)
// less sigma on the diagonal; element (row, col) of block b2 at
// M[row + nb (col + nb b2)], row = c1 + r1 i, col = b1 + r1 j.
__global__ void eig_lblocks_kernel(int r1, int n, int r2, int ra1, int ra2, const double* __restrict__ Phi1,
                                   const double* __restrict__ A, const double* __restrict__ Phi2, double sigma,
                                   double* __restrict__ M) {
  const int64_t nb = (int64_t)r1 * n, total = nb * nb * r2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e % nb, col = (e / nb) % nb, b2 = e / (nb * nb);
    const int c1 = (int)(row % r1), i = (int)(row / r1), b1 = (int)(col % r1), j = (int)(col / r1);
    double s = 0.0;
    for (int a1 = 0; a1 < ra1; ++a1) {
      double t = 0.0;
      for (int a2 = 0; a2 < ra2; ++a2)
        t += A[a1 + (int64_t)ra1 * (i + (int64_t)n * (j + (int64_t)n * a2))] *
             Phi2[b2 + (int64_t)r2 * (a2 + (int64_t)ra2 * b2)];
      s += Phi1[c1 + (int64_t)r1 * (b1 + (int64_t)r1 * a1)] * t;
    }
    M[e] = row == col ? s - sigma : s;
  }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 8192); }

void copy_cols(const double* src, int64_t lds, double* dst, int64_t ldd, int rows, int ncols,
               hipStream_t st) {
  if (ncols <= 0) return;
  TTF_HIP(hipMemcpy2DAsync(dst, sizeof(double) * ldd, src, sizeof(double) * lds, sizeof(double) * rows,
                           ncols, hipMemcpyDeviceToDevice, st));
}

// out (N x q) = S (N x p) * C (p x q)   [overwrite]
void mul_small(int N, int p, int q, const double* S, int64_t lds, const double* Cm, int ldc, double* out,
               int64_t ldo, double beta, hipStream_t st) {
  GemmOp<double> g = colmajor<double>('N', 'N', N, q, p, 1.0, S, lds, Cm, ldc, beta, out, ldo);
  gemm_launch(g, st);
}

}  // namespace

// Davidson diagonal preconditioning of the LOBPCG search directions, with
// the relative clamp delta = 1e-2 (TTF_EIG_PREC overrides; 0 leaves the
// directions unpreconditioned).  It changes the search directions only: the
// stopping rule and the Rayleigh-Ritz values are those of the plain method.
static double eig_prec_delta() {
  static const double v = [] {
    const char* e = std::getenv("TTF_EIG_PREC");
    return e ? std::max(0.0, std::atof(e)) : 1e-2;
  }();
  return v;
}

// Inverses of the shifted 'l' blocks of op into inv ((ry1 n)^2 ry2 doubles);
// false when some shifted block is not positive definite.
static bool shifted_lblock_inverse(const LocalOp& op, double sigma, double* inv, hipStream_t st) {
  const int64_t nb = (int64_t)op.s.ry1 * op.s.n;
  hipLaunchKernelGGL(eig_lblocks_kernel, dim3(grid_for(nb * nb * op.s.ry2)), dim3(256), 0, st, op.s.ry1, op.s.n,
                     op.s.ry2, op.s.ra1, op.s.ra2, op.phi1, op.A, op.phi2, sigma, inv);
  TTF_HIP(hipGetLastError());
  return spd_block_inverse((int)nb, op.s.ry2, inv, st);
}

// The shifted block preconditioner of the LOBPCG residuals: W = (M_k -
// sigma)^-1 r on the rows of each right interface index k, M_k the 'l'
// diagonal blocks (order ry1 n <= 256), sigma = theta_0 - delta max|theta|
// below the start block's lowest Ritz value theta_0 (delta = 0.01;
// TTF_EIG_BLOCK_SHIFT overrides, TTF_EIG_BLOCK=0 turns it off).  When
// sigma lies below the local matrix's lowest eigenvalue, every M_k - sigma
// is positive definite (Cauchy interlacing); when some block is not, the
// blocks are formed again with 5 delta, and after that the solve falls back
// to the diagonal preconditioner.
static double eig_block_shift() {
  static const double v = [] {
    const char* off = std::getenv("TTF_EIG_BLOCK");
    if (off && std::atoi(off) == 0) return 0.0;
    const char* e = std::getenv("TTF_EIG_BLOCK_SHIFT");
    return e ? std::max(0.0, std::atof(e)) : 0.01;
  }();
  return v;
}
// HIP graphs for the FP64 iterations, on with TTF_EIG_GRAPHS=1.  Off by
// default: on the Heisenberg d = 64, rmax 128 sweep recording an iteration
// (about 20 launches) and updating the kept executable cost more than the
// replays save over a local solve's 30-40 iterations (2.24 against 2.21 s
// per sweep); the results are the same bits either way.
static bool eig_graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("TTF_EIG_GRAPHS");
    return e && std::atoi(e) != 0;
  }();
  return on;
}
constexpr int kBlockPrecMaxOrder = 256;
constexpr size_t kBlockPrecMaxBytes = size_t(1) << 30;

EigStats block_eig(const LocalOp& op, int nev, int guard, double* Xout, double* lambda, double tol,
                   int max_matvecs, int dense_max, hipStream_t st, const MixedOp* mix) {
  EigStats stats;
  const int N = op.N();
  TTF_REQUIRE(nev >= 1 && nev <= N, "block_eig: nev=%d invalid for N=%d", nev, N);
  // ---------------- dense path: explicit local matrix + Jacobi
  if (N <= kMaxM && (N <= dense_max || N < 3 * (nev + guard))) {
    double* B = (double*)arena(kArenaSolver, sizeof(double) * ((size_t)N * N * 2 + N));
    double* Q = B + (size_t)N * N;
    double* w = Q + (size_t)N * N;
    fullmat_launch(op.s.ry1, op.s.n, op.s.ry2, op.s.ra1, op.s.ra2, op.phi1, op.A, op.phi2, B, st);
    syev_small(N, 1, B, w, Q, 0, st);
    copy_cols(Q, N, Xout, N, N, nev, st);
    std::vector<double> hw(N);
    read_back(hw.data(), w, sizeof(double) * N, st);
    for (int i = 0; i < nev; ++i) lambda[i] = hw[i];
    stats.dense = 1;
    stats.anorm = std::max(std::fabs(hw[0]), std::fabs(hw[N - 1]));
    return stats;
  }
  // ---------------- LOBPCG
  int bs = std::min(nev + guard, N / 3);
  bs = std::max(bs, nev);
  TTF_REQUIRE(3 * bs <= kMaxM && 3 * bs <= N, "block_eig: block %d too large for N=%d", bs, N);
  const int64_t ld = N;
  const size_t colsz = (size_t)N;
  // two basis sets (ping-pong), each S = [X P W] (N x 3bs) followed by
  // AS = [AX AP AW]; the Rayleigh-Ritz update writes the next set's [X P]
  // and [AX AP] in one batched product; Wr stages the residual block
  const size_t set_sz = colsz * 6 * bs;
  double* base = (double*)arena(kArenaSolver, sizeof(double) * (2 * set_sz + colsz * bs));
  double* Wr = base + 2 * set_sz;
  auto Sset = [&](int k) { return base + (size_t)k * set_sz; };
  // small device buffers
  const int M2 = kMaxM * kMaxM;
  double* sm = (double*)arena(kArenaSmall, sizeof(double) * (4 * M2 + 4 * kMaxM) + sizeof(int) * 4 * kMaxM);
  double* Gm = sm;
  double* Tm = Gm + M2;
  double* Hm = Tm + M2;
  double* CZ = Hm + M2;
  double* theta = CZ + M2;
  double* nrm = theta + kMaxM;
  double* amax = nrm + kMaxM;
  int* maskW = (int*)(amax + 2 * kMaxM);
  int* maskPb[2] = {maskW + kMaxM, maskW + 2 * kMaxM};   // ping-pong validity of P
  TTF_HIP(hipMemsetAsync(maskPb[0], 0, sizeof(int) * 2 * kMaxM, st));
  for (int k = 0; k < 2; ++k) {   // P and AP start at zero (and masked out)
    TTF_HIP(hipMemsetAsync(Sset(k) + colsz * bs, 0, sizeof(double) * colsz * bs, st));
    TTF_HIP(hipMemsetAsync(Sset(k) + colsz * 4 * bs, 0, sizeof(double) * colsz * bs, st));
  }

  // initial block in set 0: guess + guard vectors, SVQB, A X, Rayleigh-Ritz
  // on X into set 1
  {
    double* X = Sset(0);
    double* AX = X + colsz * 3 * bs;
    copy_cols(Xout, N, Wr, ld, N, nev, st);
    hipLaunchKernelGGL(random_cols, dim3(grid_for((int64_t)N * (bs - nev))), dim3(256), 0, st, N, nev,
                       bs, Wr, ld, 12345u + (unsigned)N);
    TTF_HIP(hipGetLastError());
    for (int attempt = 0; attempt < 3; ++attempt) {
      tsgram(N, bs, bs, Wr, ld, Wr, ld, Gm, bs, st);
      orth_block(bs, Gm, Tm, maskW, st);
      mul_small(N, bs, bs, Wr, ld, Tm, bs, X, ld, 0.0, st);
      std::vector<int> hm(bs);
      read_back(hm.data(), maskW, sizeof(int) * bs, st);
      bool bad = false;
      for (int i = 0; i < bs; ++i) bad = bad || !hm[i];
      if (!bad) break;
      // a rank-deficient start: replace the dropped directions by random ones
      copy_cols(X, ld, Wr, ld, N, bs, st);
      for (int i = 0; i < bs; ++i)
        if (!hm[i]) {
          hipLaunchKernelGGL(random_cols, dim3(grid_for(N)), dim3(256), 0, st, N, i, i + 1, Wr, ld,
                             777u * (attempt + 1) + i);
        }
    }
    op.apply(bs, X, ld, AX, ld, st);
    tsgram(N, bs, bs, X, ld, AX, ld, Hm, bs, st);
    launch_rr(st, bs, bs, bs, Hm, maskPb[0], maskW, theta, CZ, maskPb[1], amax, 0);
    double* Xn = Sset(1);
    mul_small(N, bs, bs, X, ld, CZ, bs, Xn, ld, 0.0, st);
    mul_small(N, bs, bs, AX, ld, CZ, bs, Xn + colsz * 3 * bs, ld, 0.0, st);
    TTF_HIP(hipMemsetAsync(maskPb[1], 0, sizeof(int) * kMaxM, st));
  }
  // the shifted block preconditioner, from the start block's Ritz values
  const double* Binv = nullptr;
  const int pnb = op.s.ry1 * op.s.n, pnblk = op.s.ry2;
  {
    const double delta = eig_block_shift();
    const size_t pbytes = sizeof(double) * (size_t)pnb * pnb * pnblk;
    if (delta > 0.0 && pnb >= 2 && pnb <= kBlockPrecMaxOrder && pbytes <= kBlockPrecMaxBytes) {
      std::vector<double> h0(2 * kMaxM + 1);
      read_back(h0.data(), theta, sizeof(double) * (2 * kMaxM + 1), st);
      const double scale = std::max(std::fabs(h0[0]), h0[2 * kMaxM]);
      double* Mb = (double*)arena(kArenaEigPrec, pbytes);
      // the shift delta, then 5 delta when a block was not positive definite
      stats.block_prec = -1;
      for (double dl : {delta, 5.0 * delta})
        if (shifted_lblock_inverse(op, h0[0] - dl * scale, Mb, st)) {
          Binv = Mb;
          stats.block_prec = 1;
          break;
        }
    }
  }
  int matvecs = bs;
  int cur = 1;
  double anorm = 0.0;
  TTF_REQUIRE(mix == nullptr || mix->nvec >= bs, "block_eig: fp32 operator holds %d vectors, need %d",
              mix ? mix->nvec : 0, bs);
  // mixed precision: A X of the current set is exact (FP64) right after a
  // refresh, until the next fp32 A W enters the Rayleigh-Ritz update
  bool fresh = true, force_sync = false;
  // once an exact residual is within kSwitch of the fp32 rounding level,
  // the search directions go back to the FP64 operator for the rest of the
  // solve: the last digits are won in FP64, as by the FP64 solver
  bool fp64_finish = false;
  constexpr double kSwitch = 1e-6;
  double r_ref = -1.0;
  // FP64 refresh of the current set: A X and A P recomputed in one block
  // product (X and P are adjacent columns of S, A X and A P of A S), a
  // Rayleigh-Ritz step on X alone into the other set; P and its validity
  // carry over (a rotation within span X keeps P orthogonal to it)
  auto refresh = [&]() {
    double* X = Sset(cur);
    double* AX = X + colsz * 3 * bs;
    op.apply(2 * bs, X, ld, AX, ld, st);
    matvecs += 2 * bs;
    ++stats.refreshes;
    tsgram(N, bs, bs, X, ld, AX, ld, Hm, bs, st);
    launch_rr(st, bs, bs, bs, Hm, maskPb[cur & 1], maskW, theta, CZ, maskPb[(cur + 1) & 1], amax, 0);
    double* Xn = Sset(1 - cur);
    mul_small(N, bs, bs, X, ld, CZ, bs, Xn, ld, 0.0, st);
    mul_small(N, bs, bs, AX, ld, CZ, bs, Xn + colsz * 3 * bs, ld, 0.0, st);
    TTF_HIP(hipMemcpyAsync(Xn + colsz * bs, X + colsz * bs, sizeof(double) * colsz * bs, hipMemcpyDeviceToDevice,
                           st));
    TTF_HIP(hipMemcpyAsync(Xn + colsz * 4 * bs, X + colsz * 4 * bs, sizeof(double) * colsz * bs,
                           hipMemcpyDeviceToDevice, st));
    TTF_HIP(hipMemcpyAsync(maskPb[(cur + 1) & 1], maskPb[cur & 1], sizeof(int) * kMaxM, hipMemcpyDeviceToDevice,
                           st));
    cur = 1 - cur;
    fresh = true;
    force_sync = true;   // the next residuals are exact: read them
  };
  std::vector<double> hbuf(2 * kMaxM + 1);
  // the residual norms are read back (a host synchronisation) at the first
  // iterations and then every kSyncEvery: in between, iterations are queued
  // without waiting.  When the residual's decrease since the last read-back
  // predicts convergence sooner, the next read-back comes at the predicted
  // iteration, so fewer iterations run past convergence (at most
  // kSyncEvery - 1 either way).
  constexpr int kSyncEvery = 4;
  int next_sync = 0, it_prev = -1;
  double r_prev = 0.0;
  // the residual's partial sums share the Gram arena, which tsgram may grow
  // (and so move) between iterations: the pointer is fetched at each launch
  // the optional diagonal preconditioner (TTF_EIG_PREC = clamp delta > 0)
  const double prec_delta = eig_prec_delta();
  double* Dg = nullptr;
  if (prec_delta > 0.0 && !Binv) {
    Dg = (double*)arena(kArenaSolver2, sizeof(double) * (size_t)N);
    localdiag_launch(op.s.ry1, op.s.n, op.s.ry2, op.s.ra1, op.s.ra2, op.phi1, op.A, op.phi2, Dg, st);
  }
  const int res_grid = (int)std::min<int64_t>(256, ceil_div(N, 256));
  const size_t res_bytes = sizeof(double) * (size_t)res_grid * bs;
  // An iteration is part A (residual block, its norms, preconditioner) and
  // part B (projections, orthonormalisation, A W, Rayleigh-Ritz, next
  // basis set), with the host's convergence decision between them; part B
  // of set c is followed at once by part A of the next set.
  // With norms false the residual norms are not summed: only an iteration
  // that reads them back (a sync point) needs them.
  auto partA = [&](int c, bool norms = true) {
    double* S = Sset(c);
    double* X = S, *W = S + colsz * 2 * bs;
    double* AX = S + colsz * 3 * bs;
    double* res_part = (double*)arena(kArenaGram, res_bytes);
    // the preconditioned residual goes to Wr; with the block preconditioner
    // the plain residual is staged in the W slot, free until the SVQB
    hipLaunchKernelGGL(residual_kernel, dim3(res_grid), dim3(256), 0, st, N, bs, X, AX, ld, theta, Binv ? W : Wr,
                       res_part, Dg, prec_delta);
    if (Binv) {
      GemmOp<double> g = colmajor<double>('N', 'N', pnb, bs, pnb, 1.0, Binv, pnb, W, ld, 0.0, Wr, ld);
      g.nb0 = pnblk;
      g.nbatch = pnblk;
      g.sAb0 = (int64_t)pnb * pnb;
      g.sBb0 = pnb;
      g.sCb0 = pnb;
      gemm_launch(g, st);
    }
    if (norms) {
      hipLaunchKernelGGL(residual_norms, dim3((unsigned)ceil_div(16 * bs, 256)), dim3(256), 0, st, bs, res_grid,
                         res_part, nrm);
      TTF_HIP(hipGetLastError());
    }
  };
  auto partB = [&](int c, bool aw32) {
    double* S = Sset(c);
    double* X = S, *W = S + colsz * 2 * bs;
    double* AS = S + colsz * 3 * bs;
    double* AW = AS + colsz * 2 * bs;
    // Wr <- Wr - [X P] ([X P]^T Wr), twice (P is zero before the first
    // update).  With the Cholesky orthonormalisation the second pass's Gram
    // is [X P Wr]^T Wr, one pass over the vectors, from which the
    // orthonormalisation kernel forms the projected block's own Gram
    tsgram(N, 2 * bs, bs, X, ld, Wr, ld, Gm, 2 * bs, st);
    gemm_launch(colmajor<double>('N', 'N', N, bs, 2 * bs, -1.0, X, ld, Gm, 2 * bs, 1.0, Wr, ld), st);
    if (eig_orth_svqb()) {
      tsgram(N, 2 * bs, bs, X, ld, Wr, ld, Gm, 2 * bs, st);
      gemm_launch(colmajor<double>('N', 'N', N, bs, 2 * bs, -1.0, X, ld, Gm, 2 * bs, 1.0, Wr, ld), st);
      tsgram(N, bs, bs, Wr, ld, Wr, ld, Gm, bs, st);
      orth_block(bs, Gm, Tm, maskW, st);
    } else {
      tsgram2(N, 2 * bs, X, ld, bs, Wr, ld, bs, Wr, ld, Gm, 3 * bs, st);
      gemm_launch(colmajor<double>('N', 'N', N, bs, 2 * bs, -1.0, X, ld, Gm, 3 * bs, 1.0, Wr, ld), st);
      orth_block(bs, Gm, Tm, maskW, st, 3 * bs, 2 * bs);
    }
    // W = Wr T, then A W
    mul_small(N, bs, bs, Wr, ld, Tm, bs, W, ld, 0.0, st);
    if (aw32) mix->apply_block(bs, W, ld, AW, ld, st);
    else op.apply(bs, W, ld, AW, ld, st);
    // Rayleigh-Ritz on S = [X P W]
    const int m = 3 * bs;
    tsgram(N, m, m, S, ld, AS, ld, Hm, m, st);
    launch_rr(st, m, bs, bs, Hm, maskPb[c & 1], maskW, theta, CZ, maskPb[(c + 1) & 1], amax, aw32 ? 1 : 0);
    // next set: [X P] = S [C Z] and [AX AP] = AS [C Z], one batched product
    double* Sn = Sset(1 - c);
    GemmOp<double> g = colmajor<double>('N', 'N', N, 2 * bs, m, 1.0, S, ld, CZ, m, 0.0, Sn, ld);
    g.nb0 = 2;
    g.nbatch = 2;
    g.sAb0 = (int64_t)colsz * 3 * bs;   // S -> AS
    g.sBb0 = 0;
    g.sCb0 = (int64_t)colsz * 3 * bs;   // next S -> next AS
    gemm_launch(g, st);
  };
  // The FP64 path replays B(c) + A(1 - c) from two captured HIP graphs, one
  // per basis set, after one iteration of eager launches has tuned every
  // GEMM shape and sized every arena (a graph keeps its pointers).  The
  // kernels and their arguments are the eager ones, so the results are the
  // same bits.
  bool graphs = eig_graphs_enabled() && !mix;
  // Instantiated graphs are kept per thread between local solves: a solve
  // of the same shape records its own iteration and updates the kept
  // executable in place (hipGraphExecUpdate), which costs far less than a
  // new instantiation; another shape, or an update the runtime refuses,
  // instantiates afresh.
  struct Kept {
    int64_t key[4] = {-1, -1, -1, -1};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
  };
  static thread_local Kept kept;
  const int64_t key[4] = {N, bs, Binv ? pnb : 0, Dg ? 1 : 0};
  if (!std::equal(key, key + 4, kept.key)) {
    for (auto& ge : kept.exec)
      if (ge) {
        (void)hipGraphExecDestroy(ge);
        ge = nullptr;
      }
    std::copy(key, key + 4, kept.key);
  }
  hipGraphExec_t* gexec = kept.exec;
  bool fresh_graph[2] = {false, false};   // updated or instantiated for this solve
  auto release = [&]() {};
  // Records B(c) + A(1 - c) into gexec[c]; false (nothing launched, the
  // stream back to normal) when the recording fails, e.g. because a call in
  // it needs a synchronisation: the solve then stays on eager launches.
  auto capture = [&](int c) -> bool {
    hipGraph_t gr = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    gemm_no_tuning(true);
    bool ok = true;
    try {
      partB(c, false);
      partA(1 - c);
    } catch (...) {
      ok = false;
    }
    gemm_no_tuning(false);
    const hipError_t e1 = hipStreamEndCapture(st, &gr);
    ok = ok && e1 == hipSuccess && gr;
    if (ok && gexec[c]) {
      hipGraphExecUpdateResult res;
      hipGraphNode_t bad = nullptr;
      if (hipGraphExecUpdate(gexec[c], gr, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(gexec[c]);
        gexec[c] = nullptr;
      }
    }
    if (ok && !gexec[c]) ok = hipGraphInstantiate(&gexec[c], gr, nullptr, nullptr, 0) == hipSuccess;
    if (gr) (void)hipGraphDestroy(gr);
    if (!ok) {
      if (gexec[c]) (void)hipGraphExecDestroy(gexec[c]);
      gexec[c] = nullptr;
      (void)hipGetLastError();
      ttf_clear_error();   // the failed recording is not an error of the solve
    }
    fresh_graph[c] = ok;
    return ok;
  };
  try {
    partA(cur);
    for (int it = 0;; ++it) {
      const bool at_cap = matvecs + bs > max_matvecs;
      if (it < 3 || it >= next_sync || at_cap || force_sync) {
        force_sync = false;
        read_back(hbuf.data(), theta, sizeof(double) * (2 * kMaxM + 1), st);
        anorm = std::max(anorm, hbuf[2 * kMaxM]);   // amax follows nrm in the buffer
        double res2 = 0.0, rmax = 0.0;
        for (int i = 0; i < nev; ++i) {
          res2 += hbuf[kMaxM + i] * hbuf[kMaxM + i];
          rmax = std::max(rmax, hbuf[kMaxM + i]);
        }
        stats.iters = it;
        stats.res = std::sqrt(res2);
        const double target = tol * std::max(anorm, 1e-300);
        const bool done = rmax <= target || at_cap;
        const double rswitch = kSwitch * std::max(anorm, 1e-300);
        if (mix && !fresh) {
          // decide on FP64 residuals: refresh at the tolerance or the cap,
          // at the switch level, and whenever the maintained residuals fell
          // 1000-fold since the last exact ones (the fp32 errors in A X
          // scale with the residual at the last refresh)
          if (done || rmax <= 1e-3 * r_ref || rmax <= rswitch) {
            refresh();
            partA(cur);
            continue;
          }
        }
        if (done) break;
        // next read-back: kSyncEvery iterations on, or where the linear
        // rate since the last one reaches the tolerance, if sooner
        int step = kSyncEvery;
        if (it_prev >= 0 && it > it_prev && rmax < r_prev && rmax > 0.0) {
          const double rate = std::log(rmax / r_prev) / (it - it_prev);   // < 0
          const double need = std::log(target / rmax) / rate;                // > 0
          if (need < kSyncEvery) step = std::max(1, (int)std::ceil(need));
        }
        r_prev = rmax;
        it_prev = it;
        next_sync = it + step;
        if (fresh) {
          r_ref = rmax;
          if (mix && rmax <= rswitch) fp64_finish = true;
        }
      }
      const bool aw32 = mix && !fp64_finish;
      if (graphs && it >= 1 && !fresh_graph[cur] && !capture(cur)) graphs = false;
      if (graphs && it >= 1) {
        TTF_HIP(hipGraphLaunch(gexec[cur], st));
      } else {
        partB(cur, aw32);
        // the residual norms of the next iteration are summed only when it
        // is a sync point (the same test as at the top of the loop)
        const int next = it + 1;
        partA(1 - cur, next < 3 || next >= next_sync || matvecs + 2 * bs > max_matvecs);
      }
      if (aw32) {
        stats.matvecs32 += bs;
        fresh = false;
      }
      matvecs += bs;
      cur = 1 - cur;
    }
  } catch (...) {
    release();
    throw;
  }
  release();
  copy_cols(Sset(cur), ld, Xout, N, N, nev, st);
  for (int i = 0; i < nev; ++i) lambda[i] = hbuf[i];
  stats.matvecs = matvecs;
  stats.anorm = anorm;
  return stats;
}

}  // namespace ttf

// ---------------------------------------------------------------- C ABI (test hook)
extern "C" int ttf_eig_block_prec(int ry1, int n, int ry2, int ra1, int ra2, const double* phi1, const double* A,
                                  const double* phi2, double sigma, double* inv, int* positive_definite) {
  using namespace ttf;
  return guarded([&] {
    TTF_REQUIRE(ry1 >= 1 && n >= 1 && ry2 >= 1 && ra1 >= 1 && ra2 >= 1 && phi1 && A && phi2 && inv &&
                    positive_definite,
                "ttf_eig_block_prec: invalid arguments");
    TTF_REQUIRE((int64_t)ry1 * n <= kBlockPrecMaxOrder, "ttf_eig_block_prec: block order %lld exceeds %d",
                (long long)ry1 * n, kBlockPrecMaxOrder);
    LocalOp op{Shape{ry1, n, ry2, ry1, n, ry2, ra1, ra2}, phi1, A, phi2};
    *positive_definite = shifted_lblock_inverse(op, sigma, inv, stream()) ? 1 : 0;
  });
}

// ztt_ksl on the device: the projector-splitting KSL time step of
// tt_ksl.f90:431-668 for dY/dt = A Y on fixed TT ranks (complex), with the
// local exponentials of explib.f90:53-125 / expokit.f:2403-2791 (ZGEXPV).
//
//   * local operator: zbfun3 (and its adjoint for the norm estimate);
//   * Krylov exponential w = exp(t B) v: Arnoldi basis on the device
//     (classical Gram-Schmidt applied twice), the small (m+2) x (m+2)
//     augmented Hessenberg exponential by the degree-6 irreducible Pade form
//     with scaling and squaring (expokit.f:767-935, ZGPADM) in one
//     work-group, and ZGEXPV's error estimate and step-size control steered
//     by the host;
//   * operator norm for the step-size heuristics: Hager's 1-norm estimate
//     with the operator and its adjoint (the reference's default is Higham's
//     block version of the same estimator, normest.f90:146-180);
//   * QR of the cores: Gram-Schmidt (twice) in LDS with an orthonormal
//     completion of rank-deficient blocks;
//   * the S-step (zsfun_matvec, tt_ksl.f90:130-136) by two complex GEMMs.
#include "solvers.h"
#include "arnoldi_gs.h"
#include "wave.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

namespace ttf {

namespace {

using zc = hipDoubleComplex;

__device__ __forceinline__ zc zmul(zc a, zc b) { return make_hipDoubleComplex(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ zc zconjmul(zc a, zc b) {   // conj(a) b
  return make_hipDoubleComplex(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ zc zadd(zc a, zc b) { return make_hipDoubleComplex(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zc zsub(zc a, zc b) { return make_hipDoubleComplex(a.x - b.x, a.y - b.y); }

int grid_for(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 4096); }

// ---------------------------------------------------------------- multi-dot
// part[blk * k + i] = sum over the block's rows of conj(V(r, i)) u(r)
constexpr int kMdotRows = 1024;
__global__ __launch_bounds__(256) void zmdot_partial(int64_t N, int k, const zc* __restrict__ V, int64_t ldv,
                                                    const zc* __restrict__ u, zc* __restrict__ part) {
  __shared__ zc us[kMdotRows];
  const int64_t r0 = (int64_t)blockIdx.x * kMdotRows;
  const int rows = (int)std::min<int64_t>(kMdotRows, N - r0);
  for (int r = threadIdx.x; r < rows; r += blockDim.x) us[r] = u[r0 + r];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i < k; i += 4) {
    const zc* v = V + ldv * i + r0;
    double re = 0.0, im = 0.0;
    for (int r = lane; r < rows; r += 64) {
      const zc a = v[r], b = us[r];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_xor(re, o);
      im += __shfl_xor(im, o);
    }
    if (lane == 0) part[(int64_t)blockIdx.x * k + i] = make_hipDoubleComplex(re, im);
  }
}

__global__ void zmdot_reduce(int k, int nparts, const zc* __restrict__ part, zc* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  double re = 0.0, im = 0.0;
  for (int b = 0; b < nparts; ++b) {
    re += part[(int64_t)b * k + i].x;
    im += part[(int64_t)b * k + i].y;
  }
  out[i] = make_hipDoubleComplex(re, im);
}

// out (k, device) = V(:, 0:k)^H u, deterministic
void zmdot(int64_t N, int k, const zc* V, int64_t ldv, const zc* u, zc* out, hipStream_t st) {
  const int nparts = (int)ceil_div(N, kMdotRows);
  zc* part = (zc*)arena(kArenaGram, sizeof(zc) * (size_t)nparts * k);
  hipLaunchKernelGGL(zmdot_partial, dim3(nparts), dim3(256), 0, st, N, k, V, ldv, u, part);
  TTF_HIP(hipGetLastError());
  hipLaunchKernelGGL(zmdot_reduce, dim3((unsigned)ceil_div(k, 64)), dim3(64), 0, st, k, nparts, part, out);
  TTF_HIP(hipGetLastError());
}

__global__ void zscale_kernel(int64_t n, zc a, const zc* __restrict__ x, zc* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = zmul(a, x[e]);
}

// u /= sqrt(nrm2[0].x), or u = 0 when that norm is at most tol (Arnoldi
// breakdown): the normalisation of a Krylov step without a host round trip.
__global__ void zscale_by_norm_kernel(int64_t n, const zc* __restrict__ nrm2, double tol, zc* __restrict__ u) {
  const double h = sqrt(fmax(nrm2[0].x, 0.0));
  const double inv = h > tol ? 1.0 / h : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = make_hipDoubleComplex(u[i].x * inv, u[i].y * inv);
}

void zscale(int64_t n, zc a, const zc* x, zc* y, hipStream_t st) {
  hipLaunchKernelGGL(zscale_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, a, x, y);
  TTF_HIP(hipGetLastError());
}

// C (M x N, ldc) = alpha op(A) op(B) + beta C with column-major operands;
// ta / tb in {'N', 'C'} (conjugate transpose)
void zgemm_cm(char ta, char tb, int M, int N, int K, zc alpha, const zc* A, int64_t lda, const zc* B, int64_t ldb,
              zc beta, zc* C, int64_t ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  ZGemmOp g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.B = B; g.C = C;
  if (ta == 'N') { g.sAi = 1; g.sAk = lda; } else { g.sAi = lda; g.sAk = 1; g.conjA = 1; }
  if (tb == 'N') { g.sBk = 1; g.sBj = ldb; } else { g.sBk = ldb; g.sBj = 1; g.conjB = 1; }
  g.sCi = 1; g.sCj = ldc;
  g.nb0 = 1; g.nbatch = 1;
  g.alpha = alpha; g.beta = beta;
  zgemm_launch(g, st);
}

double znorm_dev(int64_t N, const zc* v, zc* scratch, hipStream_t st) {
  zmdot(N, 1, v, N, v, scratch, st);
  zc h;
  read_back(&h, scratch, sizeof(zc), st);
  return std::sqrt(std::max(h.x, 0.0));
}

// ---------------------------------------------------------------- complex QR
// Q (rows x rnew) with orthonormal columns spanning the columns of A (rows x
// cols, column-major, in LDS): Gram-Schmidt twice, columns in order, a column
// whose remainder is below 1e-13 of its norm counted as dependent; missing
// directions completed from deterministic pseudo-random vectors.
constexpr int kQrLds = 8192;   // complex entries of A held in LDS

__global__ __launch_bounds__(256) void zqr_gs_kernel(int rows, int cols, int rnew, const zc* __restrict__ A,
                                                     zc* __restrict__ Q) {
  extern __shared__ zc L[];   // column c at L + c * rows, then the basis compacted in place
  __shared__ zc h[64];
  __shared__ double red[8];
  __shared__ int nb;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < rows * cols; e += nt) L[e] = A[e];
  if (tid == 0) nb = 0;
  __syncthreads();
  auto block_norm2 = [&](const zc* a) -> double {
    double s = 0.0;
    for (int r = tid; r < rows; r += nt) s += a[r].x * a[r].x + a[r].y * a[r].y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (nt >> 6); ++w) t += red[w];
    return t;
  };
  // a -= Q Q^H a over the first b basis columns, twice.  The b <= 64 dot
  // products take a quad of lanes each (every lane a quarter of the rows,
  // then a quad sum in registers), so the 256 threads share them.
  auto orthogonalise = [&](zc* a, int b) {
    for (int pass = 0; pass < 2; ++pass) {
      {
        const int i = tid >> 2, part = tid & 3;
        double re = 0.0, im = 0.0;
        if (i < b) {
          const zc* qi = L + (size_t)i * rows;
          for (int r = part; r < rows; r += 4) {
            re += qi[r].x * a[r].x + qi[r].y * a[r].y;
            im += qi[r].x * a[r].y - qi[r].y * a[r].x;
          }
        }
        re = quad_sum(re);
        im = quad_sum(im);
        if (i < b && part == 0) h[i] = make_hipDoubleComplex(re, im);
      }
      __syncthreads();
      for (int r = tid; r < rows; r += nt) {
        zc s = a[r];
        for (int i = 0; i < b; ++i) s = zsub(s, zmul(h[i], L[(size_t)i * rows + r]));
        a[r] = s;
      }
      __syncthreads();
    }
  };
  for (int j = 0; j < cols; ++j) {
    const int b = nb;
    if (b >= rnew) break;
    zc* a = L + (size_t)j * rows;
    const double n0 = block_norm2(a);
    orthogonalise(a, b);
    const double n1 = block_norm2(a);
    const bool accept = n1 > 1e-26 * n0 && n1 > 1e-300;
    if (accept) {
      const double inv = 1.0 / sqrt(n1);
      zc* dst = L + (size_t)b * rows;
      for (int r = tid; r < rows; r += nt) dst[r] = make_hipDoubleComplex(a[r].x * inv, a[r].y * inv);
    }
    __syncthreads();
    if (accept && tid == 0) nb = b + 1;
    __syncthreads();
  }
  // completion: columns nb..rnew-1 (their LDS slots are free)
  for (int c = nb; c < rnew; ++c) {
    zc* a = L + (size_t)c * rows;
    for (int r = tid; r < rows; r += nt) {
      unsigned hsh = (unsigned)(r * 2654435761u) ^ (unsigned)(c * 40503u + 977u);
      hsh ^= hsh >> 13; hsh *= 0x5bd1e995u; hsh ^= hsh >> 15;
      a[r] = make_hipDoubleComplex((double)(hsh & 0xffff) / 65536.0 - 0.5, (double)(hsh >> 16) / 65536.0 - 0.5);
    }
    __syncthreads();
    orthogonalise(a, c);
    const double n1 = block_norm2(a);
    const double inv = n1 > 0.0 ? 1.0 / sqrt(n1) : 0.0;
    for (int r = tid; r < rows; r += nt) a[r] = make_hipDoubleComplex(a[r].x * inv, a[r].y * inv);
    __syncthreads();
  }
  for (int e = tid; e < rows * rnew; e += nt) Q[e] = L[e];
}

// The same factorisation by 8-column panels: each panel is projected
// against the basis accepted so far by block products that all 256 threads
// share (Gram-Schmidt twice: H = Q^H P, P -= Q H), then its columns are
// taken in order against the panel's own accepted columns.  A column is
// dependent when its remainder is below 1e-13 of its norm before the panel's
// projections; missing directions are completed as above.  Per column this
// needs two barriers in place of about ten, and the long updates run on all
// threads instead of one thread per row.
constexpr int kQrPanelCols = 8;

// Panel update P(:, c) -= sum_i q_i H(i, c) for the nb basis columns: a
// thread takes one row and CPT panel columns (per = kQrPanelCols / CPT
// threads share a row), so each q_i(r) it loads serves all of them.  Columns
// at or beyond jp read column jp - 1 instead and are not stored: the loads
// carry no condition (a conditional LDS load becomes a branch with a full
// wait, which serialises the loop).
template <int CPT>
__device__ __forceinline__ void zqr_panel_update(zc* L, const zc* H, int rows, int nb, int j0, int jp) {
  constexpr int per = kQrPanelCols / CPT;
  for (int e = threadIdx.x; e < rows * per; e += blockDim.x) {
    const int r = e % rows, c0 = e / rows;
    int hc[CPT];
    zc acc[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      hc[k] = 64 * min(c0 + per * k, jp - 1);
      acc[k] = make_hipDoubleComplex(0.0, 0.0);
    }
    for (int i = 0; i < nb; ++i) {
      const zc q = L[(size_t)i * rows + r];
      zc h[CPT];
#pragma unroll
      for (int k = 0; k < CPT; ++k) h[k] = H[i + hc[k]];
#pragma unroll
      for (int k = 0; k < CPT; ++k) acc[k] = zadd(acc[k], zmul(h[k], q));
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int c = c0 + per * k;
      if (c < jp) {
        zc* p = L + (size_t)(j0 + c) * rows + r;
        *p = zsub(*p, acc[k]);
      }
    }
  }
}

// One work-group per launch, so the register budget is that of one wave per
// SIMD (amdgpu_waves_per_eu(1, 1)): the compiler then keeps a loop's LDS
// loads in flight together instead of reusing one register per load.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void zqr_panel_gs_kernel(
    int rows, int cols, int rnew, const zc* __restrict__ A, zc* __restrict__ Q) {
  extern __shared__ zc L[];   // column c at L + c * rows; the basis is compacted into the leading columns
  __shared__ zc H[64 * kQrPanelCols];   // H(i, c) at i + 64 c
  __shared__ double n0[kQrPanelCols];
  __shared__ double red[4][kQrPanelCols];
  __shared__ int nb_s;
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
  for (int e = tid; e < rows * cols; e += nt) L[e] = A[e];
  if (tid == 0) nb_s = 0;
  __syncthreads();
  // squared norms of panel columns j0..j0+jp-1 into out (wave sums, then the
  // four waves' partials added by each reader)
  auto panel_norms = [&](int j0, int jp, double* out) {
    double s[kQrPanelCols];
#pragma unroll
    for (int c = 0; c < kQrPanelCols; ++c) s[c] = 0.0;
    // columns at or beyond jp repeat column jp - 1 (unconditional loads)
    for (int r = tid; r < rows; r += nt)
#pragma unroll
      for (int c = 0; c < kQrPanelCols; ++c) {
        const zc v = L[(size_t)(j0 + min(c, jp - 1)) * rows + r];
        s[c] += v.x * v.x + v.y * v.y;
      }
#pragma unroll
    for (int c = 0; c < kQrPanelCols; ++c) {
      const double t = wave_sum64(s[c]);
      if (lane == 0) red[wv][c] = t;
    }
    __syncthreads();
    if (tid < kQrPanelCols) out[tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    __syncthreads();
  };
  for (int j0 = 0; j0 < cols; j0 += kQrPanelCols) {
    const int nb = nb_s;
    if (nb >= rnew) break;
    const int jp = min(kQrPanelCols, cols - j0);
    panel_norms(j0, jp, n0);
    // project the panel against the nb accepted columns, twice
    for (int pass = 0; pass < 2 && nb > 0; ++pass) {
      // H(i, c) = q_i^H p_c: each DPP row (16 lanes) takes one basis column
      // i and all jp panel columns, a sixteenth of the rows per lane, and
      // sums over the row; 16 basis columns per round
      {
        const int g = tid >> 4, l16 = tid & 15;
        for (int i0 = 0; i0 < nb; i0 += nt >> 4) {
          const int i = i0 + g;
          const bool oki = i < nb;
          const zc* qi = L + (size_t)(oki ? i : 0) * rows;
          double re[kQrPanelCols], im[kQrPanelCols];
#pragma unroll
          for (int c = 0; c < kQrPanelCols; ++c) re[c] = im[c] = 0.0;
          // columns at or beyond jp repeat column jp - 1 (unconditional
          // loads; their sums are not stored)
          for (int r = l16; r < rows; r += 16) {
            const zc q = qi[r];
            zc v[kQrPanelCols];
#pragma unroll
            for (int c = 0; c < kQrPanelCols; ++c) v[c] = L[(size_t)(j0 + min(c, jp - 1)) * rows + r];
#pragma unroll
            for (int c = 0; c < kQrPanelCols; ++c) {
              re[c] += q.x * v[c].x + q.y * v[c].y;
              im[c] += q.x * v[c].y - q.y * v[c].x;
            }
          }
#pragma unroll
          for (int c = 0; c < kQrPanelCols; ++c) {
            re[c] = row_sum16(re[c]);
            im[c] = row_sum16(im[c]);
          }
          if (oki && l16 < jp) {
            // lane l16 stores column l16 (a static index through a select chain)
            double rv = re[0], iv = im[0];
#pragma unroll
            for (int c = 1; c < kQrPanelCols; ++c)
              if (l16 == c) { rv = re[c]; iv = im[c]; }
            H[i + 64 * l16] = make_hipDoubleComplex(rv, iv);
          }
        }
      }
      __syncthreads();
      // p_c(r) -= sum_i q_i(r) H(i, c); nt / rows threads share a row (at
      // most one per panel column)
      {
        const int per = nt / rows;
        if (per >= 8) zqr_panel_update<1>(L, H, rows, nb, j0, jp);
        else if (per >= 4) zqr_panel_update<2>(L, H, rows, nb, j0, jp);
        else if (per >= 2) zqr_panel_update<4>(L, H, rows, nb, j0, jp);
        else zqr_panel_update<8>(L, H, rows, nb, j0, jp);
      }
      __syncthreads();
    }
    // the panel's columns in order, against the columns this panel accepted
    // (at most kQrPanelCols - 1): one wave per dot product
    for (int c = 0; c < jp; ++c) {
      const int b = nb_s, b0 = nb;   // accepted so far; of them, this panel's are b0..b-1
      if (b >= rnew) break;
      zc* a = L + (size_t)(j0 + c) * rows;
      for (int pass = 0; pass < 2 && b > b0; ++pass) {
        for (int i = b0 + wv; i < b; i += nt >> 6) {
          const zc* qi = L + (size_t)i * rows;
          double re = 0.0, im = 0.0;
          for (int r = lane; r < rows; r += 64) {
            re += qi[r].x * a[r].x + qi[r].y * a[r].y;
            im += qi[r].x * a[r].y - qi[r].y * a[r].x;
          }
          re = wave_sum64(re);
          im = wave_sum64(im);
          if (lane == 0) H[i - b0] = make_hipDoubleComplex(re, im);
        }
        __syncthreads();
        for (int r = tid; r < rows; r += nt) {
          zc s0 = a[r];
          for (int i = b0; i < b; ++i) s0 = zsub(s0, zmul(H[i - b0], L[(size_t)i * rows + r]));
          a[r] = s0;
        }
        __syncthreads();
      }
      double sq = 0.0;
      for (int r = tid; r < rows; r += nt) sq += a[r].x * a[r].x + a[r].y * a[r].y;
      sq = wave_sum64(sq);
      if (lane == 0) red[wv][0] = sq;
      __syncthreads();
      const double n1 = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
      const bool accept = n1 > 1e-26 * n0[c] && n1 > 1e-300;
      if (accept) {
        // slot b is this column's own or an earlier, finished one
        const double inv = 1.0 / sqrt(n1);
        zc* dst = L + (size_t)b * rows;
        for (int r = tid; r < rows; r += nt) dst[r] = make_hipDoubleComplex(a[r].x * inv, a[r].y * inv);
      }
      __syncthreads();
      if (accept && tid == 0) nb_s = b + 1;
      __syncthreads();
    }
  }
  // completion: columns nb..rnew-1 (their LDS slots are free), each against
  // all the columns before it, twice
  for (int c = nb_s; c < rnew; ++c) {
    zc* a = L + (size_t)c * rows;
    for (int r = tid; r < rows; r += nt) {
      unsigned hsh = (unsigned)(r * 2654435761u) ^ (unsigned)(c * 40503u + 977u);
      hsh ^= hsh >> 13; hsh *= 0x5bd1e995u; hsh ^= hsh >> 15;
      a[r] = make_hipDoubleComplex((double)(hsh & 0xffff) / 65536.0 - 0.5, (double)(hsh >> 16) / 65536.0 - 0.5);
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = tid >> 2; i < c; i += nt >> 2) {
        const int part = tid & 3;
        const zc* qi = L + (size_t)i * rows;
        double re = 0.0, im = 0.0;
        for (int r = part; r < rows; r += 4) {
          re += qi[r].x * a[r].x + qi[r].y * a[r].y;
          im += qi[r].x * a[r].y - qi[r].y * a[r].x;
        }
        re = quad_sum(re);
        im = quad_sum(im);
        if (part == 0) H[i] = make_hipDoubleComplex(re, im);
      }
      __syncthreads();
      for (int r = tid; r < rows; r += nt) {
        zc s0 = a[r];
        for (int i = 0; i < c; ++i) s0 = zsub(s0, zmul(H[i], L[(size_t)i * rows + r]));
        a[r] = s0;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int r = tid; r < rows; r += nt) s += a[r].x * a[r].x + a[r].y * a[r].y;
    s = wave_sum64(s);
    if (lane == 0) red[wv][0] = s;
    __syncthreads();
    const double n1 = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    const double inv = n1 > 0.0 ? 1.0 / sqrt(n1) : 0.0;
    for (int r = tid; r < rows; r += nt) a[r] = make_hipDoubleComplex(a[r].x * inv, a[r].y * inv);
    __syncthreads();
  }
  for (int e = tid; e < rows * rnew; e += nt) Q[e] = L[e];
}

// ---------------------------------------------------------------- small expm
// E (m x m) = exp(t H) for the m x m column-major H (m <= kExpmMax): the
// irreducible (6, 6) Pade form +/-(I + 2 q^-1 p) with scaling and squaring,
// one work-group, all matrices in LDS.
constexpr int kExpmMax = 40;

__global__ __launch_bounds__(256) void zexpm_kernel(int m, double t, const zc* __restrict__ Hg, zc* __restrict__ Eg,
                                                    int* __restrict__ nsout) {
  __shared__ zc H[kExpmMax * kExpmMax], H2[kExpmMax * kExpmMax], P[kExpmMax * kExpmMax],
      Qm[kExpmMax * kExpmMax], W[kExpmMax * kExpmMax];
  __shared__ double coef[8], rowsum[kExpmMax];
  __shared__ int ns_s, piv_s;
  const int tid = threadIdx.x, nt = blockDim.x, mm = m * m;
  constexpr int ideg = 6;
  for (int e = tid; e < mm; e += nt) H[e] = Hg[e];
  __syncthreads();
  for (int i = tid; i < m; i += nt) {
    double s = 0.0;
    for (int j = 0; j < m; ++j) s += hypot(H[i + m * j].x, H[i + m * j].y);
    rowsum[i] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double hn = 0.0;
    for (int i = 0; i < m; ++i) hn = fmax(hn, rowsum[i]);
    hn = fabs(t * hn);
    ns_s = hn > 0.0 ? max(0, (int)(log(hn) / log(2.0)) + 2) : 0;
    coef[0] = 1.0;
    for (int k = 1; k <= ideg; ++k) coef[k] = coef[k - 1] * (double)(ideg + 1 - k) / (double)(k * (2 * ideg + 1 - k));
  }
  __syncthreads();
  const int ns = ns_s;
  const double scale = t / ldexp(1.0, ns);
  // C = alpha * A * B (all m x m in LDS); the inner products take four k
  // at a time into two partial sums, so that their loads are in flight
  // together
  auto matmul = [&](const zc* A, const zc* B, zc* C, double alpha) {
    for (int e = tid; e < mm; e += nt) {
      const int i = e % m, j = e / m;
      double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
      int k = 0;
      for (; k + 4 <= m; k += 4) {
        zc a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = A[i + m * (k + u)];
          b[u] = B[k + u + m * j];
        }
#pragma unroll
        for (int u = 0; u < 4; u += 2) {
          re0 += a[u].x * b[u].x - a[u].y * b[u].y;
          im0 += a[u].x * b[u].y + a[u].y * b[u].x;
          re1 += a[u + 1].x * b[u + 1].x - a[u + 1].y * b[u + 1].y;
          im1 += a[u + 1].x * b[u + 1].y + a[u + 1].y * b[u + 1].x;
        }
      }
      for (; k < m; ++k) {
        const zc a = A[i + m * k], b = B[k + m * j];
        re0 += a.x * b.x - a.y * b.y;
        im0 += a.x * b.y + a.y * b.x;
      }
      C[e] = make_hipDoubleComplex(alpha * (re0 + re1), alpha * (im0 + im1));
    }
    __syncthreads();
  };
  matmul(H, H, H2, scale * scale);
  for (int e = tid; e < mm; e += nt) {
    const bool d = (e % m) == (e / m);
    P[e] = make_hipDoubleComplex(d ? coef[ideg - 1] : 0.0, 0.0);
    Qm[e] = make_hipDoubleComplex(d ? coef[ideg] : 0.0, 0.0);
  }
  __syncthreads();
  // Horner: alternate between q and p
  zc* p = P;
  zc* q = Qm;
  zc* f = W;
  int odd = 1;
  for (int k = ideg - 1; k > 0; --k) {
    zc* used = odd ? q : p;
    matmul(used, H2, f, 1.0);
    for (int i = tid; i < m; i += nt) f[i + m * i].x += coef[k - 1];
    __syncthreads();
    if (odd) { zc* tq = q; q = f; f = tq; } else { zc* tp = p; p = f; f = tp; }
    odd = 1 - odd;
  }
  if (odd) { matmul(q, H, f, scale); zc* tq = q; q = f; f = tq; }
  else { matmul(p, H, f, scale); zc* tp = p; p = f; f = tp; }
  // q <- q - p; solve q X = p (partial pivoting), p <- 2 X + I
  for (int e = tid; e < mm; e += nt) q[e] = zsub(q[e], p[e]);
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    if (tid < 64) {
      // the first row of largest |q(i, k)| (i >= k) by a wave reduction of
      // keys |q|^2 (non-negative, so their bits order as integers) with the
      // low six bits replaced by 63 - i (the earlier row wins ties)
      const int i = tid;
      const zc v = q[(i < m ? i : k) + m * k];
      const double a2 = v.x * v.x + v.y * v.y;
      unsigned long long key = (i >= k && i < m)
                                   ? (((unsigned long long)__double_as_longlong(a2)) & ~63ull) | (unsigned)(63 - i)
                                   : 0ull;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(key, o);
        key = other > key ? other : key;
      }
      if (tid == 0) piv_s = 63 - (int)(key & 63ull);
    }
    __syncthreads();
    const int pv = piv_s;
    if (pv != k) {
      for (int j = tid; j < m; j += nt) {
        zc t1 = q[k + m * j]; q[k + m * j] = q[pv + m * j]; q[pv + m * j] = t1;
        zc t2 = p[k + m * j]; p[k + m * j] = p[pv + m * j]; p[pv + m * j] = t2;
      }
      __syncthreads();
    }
    const zc d = q[k + m * k];
    const double dn = d.x * d.x + d.y * d.y;
    const zc dinv = make_hipDoubleComplex(d.x / dn, -d.y / dn);
    // eliminate below row k: q columns k+1.. and all columns of p; the
    // multipliers q(i, k) / q(k, k) come from column k, which is not written
    for (int e = tid; e < (m - k - 1) * (2 * m - k - 1); e += nt) {
      const int ri = e % (m - k - 1), cj = e / (m - k - 1);
      const int i = k + 1 + ri;
      const zc f1 = zmul(q[i + m * k], dinv);
      if (cj < m - k - 1) {
        const int j = k + 1 + cj;
        q[i + m * j] = zsub(q[i + m * j], zmul(f1, q[k + m * j]));
      } else {
        const int j = cj - (m - k - 1);
        p[i + m * j] = zsub(p[i + m * j], zmul(f1, p[k + m * j]));
      }
    }
    __syncthreads();
  }
  for (int k = m - 1; k >= 0; --k) {
    const zc d = q[k + m * k];
    const double dn = d.x * d.x + d.y * d.y;
    const zc dinv = make_hipDoubleComplex(d.x / dn, -d.y / dn);
    for (int j = tid; j < m; j += nt) {
      // two partial sums, two l per step each, loads together
      zc s0 = p[k + m * j], s1 = make_hipDoubleComplex(0.0, 0.0);
      int l = k + 1;
      for (; l + 2 <= m; l += 2) {
        const zc q0 = q[k + m * l], p0 = p[l + m * j], q1 = q[k + m * (l + 1)], p1 = p[l + 1 + m * j];
        s0 = zsub(s0, zmul(q0, p0));
        s1 = zsub(s1, zmul(q1, p1));
      }
      if (l < m) s0 = zsub(s0, zmul(q[k + m * l], p[l + m * j]));
      p[k + m * j] = zmul(zadd(s0, s1), dinv);
    }
    __syncthreads();
  }
  for (int e = tid; e < mm; e += nt) {
    zc v = p[e];
    v.x = 2.0 * v.x + ((e % m) == (e / m) ? 1.0 : 0.0);
    v.y = 2.0 * v.y;
    if (ns == 0 && odd) { v.x = -v.x; v.y = -v.y; }
    p[e] = v;
  }
  __syncthreads();
  zc* cur = p;
  zc* nxt = q;
  for (int s = 0; s < ns; ++s) {
    matmul(cur, cur, nxt, 1.0);
    zc* tt = cur; cur = nxt; nxt = tt;
  }
  for (int e = tid; e < mm; e += nt) Eg[e] = cur[e];
  if (tid == 0 && nsout) *nsout = ns;
}

// ---------------------------------------------------------------- Hager 1-norm
__global__ void zsign_abs_kernel(int64_t n, const zc* __restrict__ y, zc* __restrict__ s, double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double a = hypot(y[e].x, y[e].y);
    acc += a;
    s[e] = a > 0.0 ? make_hipDoubleComplex(y[e].x / a, y[e].y / a) : make_hipDoubleComplex(1.0, 0.0);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// part[2b] = max |z| over the block, part[2b+1] = its first index
__global__ void zargmax_kernel(int64_t n, const zc* __restrict__ z, double* __restrict__ part) {
  __shared__ double bv[256];
  __shared__ int64_t bi[256];
  double best = -1.0;
  int64_t arg = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double a = hypot(z[e].x, z[e].y);
    if (a > best) { best = a; arg = e; }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = arg;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      if (bv[threadIdx.x + o] > bv[threadIdx.x] ||
          (bv[threadIdx.x + o] == bv[threadIdx.x] && bi[threadIdx.x + o] < bi[threadIdx.x])) {
        bv[threadIdx.x] = bv[threadIdx.x + o];
        bi[threadIdx.x] = bi[threadIdx.x + o];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = bv[0]; part[2 * blockIdx.x + 1] = (double)bi[0]; }
}

__global__ void zfill_kernel(int64_t n, zc v, zc* x) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) x[e] = v;
}

// B (cols x rows) = A^T for the column-major rows x cols A (no conjugation)
// complex128 <-> complex64 (round to nearest / exact widening)
__global__ void z2c_kernel(int64_t n, const zc* __restrict__ x, hipFloatComplex* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = make_hipFloatComplex((float)x[e].x, (float)x[e].y);
}
__global__ void c2z_kernel(int64_t n, const hipFloatComplex* __restrict__ x, zc* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = make_hipDoubleComplex((double)x[e].x, (double)x[e].y);
}
void to_c64(int64_t n, const zc* x, hipFloatComplex* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(z2c_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  TTF_HIP(hipGetLastError());
}
void to_c128(int64_t n, const hipFloatComplex* x, zc* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(c2z_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  TTF_HIP(hipGetLastError());
}

__global__ void ztranspose_kernel(int rows, int cols, const zc* __restrict__ A, zc* __restrict__ B) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % rows, c = e / rows;
    B[c + (int64_t)cols * r] = A[e];
  }
}

using ZOp = std::function<void(const zc*, zc*)>;

// Hager's estimate of ||B||_1 from B and B^H (Higham's algorithm with one
// column, at most 5 iterations).
double hager_norm1(int64_t n, const ZOp& op, const ZOp& adj, hipStream_t st) {
  zc* x = (zc*)arena(kArenaSolver2, sizeof(zc) * (size_t)n * 3 + 4096);
  zc* y = x + n;
  zc* z = y + n;
  double* part = (double*)(z + n);
  const int blocks = std::min<int>(grid_for(n), 128);
  hipLaunchKernelGGL(zfill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, make_hipDoubleComplex(1.0 / n, 0.0), x);
  double est = 0.0;
  int64_t jprev = -1;
  std::vector<double> hp(2 * blocks);
  for (int it = 0; it < 5; ++it) {
    op(x, y);
    hipLaunchKernelGGL(zsign_abs_kernel, dim3(blocks), dim3(256), 0, st, n, y, y, part);
    read_back(hp.data(), part, sizeof(double) * blocks, st);
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += hp[b];
    if (it > 0 && s <= est) break;
    est = s;
    adj(y, z);
    hipLaunchKernelGGL(zargmax_kernel, dim3(blocks), dim3(256), 0, st, n, z, part);
    TTF_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, st));
    // Re(z^H x): x is either uniform 1/n or e_jprev
    zc zx{0.0, 0.0};
    if (jprev < 0) {
      zc* dot = (zc*)(part + 2 * blocks + 8);
      zmdot(n, 1, x, n, z, dot, st);   // x^H z = conj(z^H x); the real parts agree
      TTF_HIP(hipMemcpyAsync(&zx, dot, sizeof(zc), hipMemcpyDeviceToHost, st));
    } else {
      TTF_HIP(hipMemcpyAsync(&zx, z + jprev, sizeof(zc), hipMemcpyDeviceToHost, st));
    }
    TTF_HIP(hipStreamSynchronize(st));
    double zmax = -1.0;
    int64_t jmax = 0;
    for (int b = 0; b < blocks; ++b)
      if (hp[2 * b] > zmax || (hp[2 * b] == zmax && (int64_t)hp[2 * b + 1] < jmax)) {
        zmax = hp[2 * b];
        jmax = (int64_t)hp[2 * b + 1];
      }
    if (it > 0 && zmax <= zx.x) break;
    if (jmax == jprev) break;
    hipLaunchKernelGGL(zfill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, make_hipDoubleComplex(0.0, 0.0), x);
    const zc one{1.0, 0.0};
    TTF_HIP(hipMemcpyAsync(x + jmax, &one, sizeof(zc), hipMemcpyHostToDevice, st));
    jprev = jmax;
  }
  return est;
}

}  // namespace

// ---------------------------------------------------------------- zexp_mv
struct ExpInfo {
  int nmult = 0, nstep = 0, nexph = 0;
  double hump = 0.0;
};

void zexpm_small(int m, double t, const zc* Hdev, zc* Edev, hipStream_t st) {
  TTF_REQUIRE(m >= 1 && m <= kExpmMax, "zexpm: order %d outside 1..%d", m, kExpmMax);
  hipLaunchKernelGGL(zexpm_kernel, dim3(1), dim3(256), 0, st, m, t, Hdev, Edev, (int*)nullptr);
  TTF_HIP(hipGetLastError());
}

namespace {

int nint(double x) { return (int)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

// w = exp(t A) v for the operator `op` of size n, v normalised (ZGEXPV,
// expokit.f:2403-2791); v and w are device vectors, w may alias nothing.
ExpInfo zgexpv(int64_t n, int m, double t, const zc* v, zc* w, double tol, double anorm, const ZOp& op,
               hipStream_t st) {
  ExpInfo info;
  const double eps = 2.220446049250313e-16;
  if (tol <= eps) tol = std::sqrt(eps);
  const double rndoff = eps * anorm;
  (void)rndoff;
  const double break_tol = 1.0e-7, delta = 1.2, gamma = 0.9;
  const int mxstep = 500;
  const int mh = m + 2;
  TTF_REQUIRE(mh <= kExpmMax, "zexp_mv: Krylov order %d exceeds %d", m, kExpmMax - 2);
  const double sgn = t >= 0 ? 1.0 : -1.0, t_out = std::fabs(t);
  // Krylov basis (m + 2 vectors), then per Arnoldi step its dot products
  // (two passes of 64) and squared norm, read back every kArnoldiSync steps
  constexpr int kArnoldiSync = 8;
  zc* V = (zc*)arena(kArenaSolver, sizeof(zc) * ((size_t)n * (m + 2) + (size_t)129 * m));
  zc* Hs = V + (size_t)n * (m + 2);
  zc* sm = (zc*)arena(kArenaSmall, sizeof(zc) * (size_t)(3 * kExpmMax * kExpmMax + 256));
  zc* Hd = sm;
  zc* Ed = Hd + kExpmMax * kExpmMax;
  zc* hb = Ed + kExpmMax * kExpmMax;   // dot products and coefficients
  TTF_HIP(hipMemcpyAsync(w, v, sizeof(zc) * n, hipMemcpyDeviceToDevice, st));
  double beta = znorm_dev(n, w, hb, st);
  const double vnorm = beta;
  double hump = beta;
  const double SQR1 = std::sqrt(0.1);
  double xm = 1.0 / m;
  double p2 = tol * std::pow((m + 1) / 2.72, m + 1) * std::sqrt(2.0 * 3.14 * (m + 1));
  double t_new = (1.0 / anorm) * std::pow(p2 / (4.0 * beta * anorm), xm);
  double p1 = std::pow(10.0, nint(std::log10(t_new) - SQR1) - 1);
  t_new = std::trunc(t_new / p1 + 0.55) * p1;
  double t_now = 0.0;
  std::vector<zc> H((size_t)mh * mh), hh(2 * 64 + 2), E((size_t)mh * mh);
  while (t_now < t_out) {
    ++info.nstep;
    double t_step = std::min(t_out - t_now, t_new);
    zscale(n, make_hipDoubleComplex(1.0 / beta, 0.0), w, V, st);
    std::fill(H.begin(), H.end(), make_hipDoubleComplex(0.0, 0.0));
    int k1 = 2, mbrkdwn = m;
    double avnorm = 0.0;
    bool brk = false;
    // Arnoldi steps queued without waiting: each step's coefficients stay on
    // the device and the normalisation is done there (a vector below the
    // breakdown tolerance becomes zero, so steps queued past a breakdown
    // only produce zeros); the host reads the columns every kArnoldiSync
    // steps and takes them in order, stopping at the first breakdown as the
    // step-by-step loop does.
    int j_done = 0;
    for (int j = 1; j <= m && !brk; ++j) {
      zc* u = V + (size_t)n * j;
      op(V + (size_t)n * (j - 1), u);
      ++info.nmult;
      zc* hs = Hs + (size_t)129 * (j - 1);
      // Gram-Schmidt twice against V(:, 0..j-1), then u / |u|
      if (gs::fits(n, j)) {
        gs::step<zc>(n, j, V, n, u, hs, break_tol, u, st);
      } else {
        for (int pass = 0; pass < 2; ++pass) {
          zc* h = hs + 64 * pass;
          zmdot(n, j, V, n, u, h, st);
          zgemm_cm('N', 'N', (int)n, 1, j, make_hipDoubleComplex(-1.0, 0.0), V, n, h, j,
                   make_hipDoubleComplex(1.0, 0.0), u, n, st);
        }
        zmdot(n, 1, u, n, u, hs + 128, st);
        hipLaunchKernelGGL(zscale_by_norm_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 4096)),
                           dim3(256), 0, st, n, hs + 128, break_tol, u);
        TTF_HIP(hipGetLastError());
      }
      if (j % kArnoldiSync != 0 && j != m) continue;
      const int cnt = j - j_done;
      hh.resize((size_t)129 * cnt);
      read_back(hh.data(), Hs + (size_t)129 * j_done, sizeof(zc) * 129 * cnt, st);
      for (int jj = j_done + 1; jj <= j; ++jj) {
        const zc* c = hh.data() + (size_t)129 * (jj - 1 - j_done);
        for (int i = 0; i < jj; ++i)
          H[(size_t)(jj - 1) * mh + i] = make_hipDoubleComplex(c[i].x + c[64 + i].x, c[i].y + c[64 + i].y);
        const double hj1j = std::sqrt(std::max(c[128].x, 0.0));
        if (hj1j <= break_tol) {   // happy breakdown
          k1 = 0;
          mbrkdwn = jj;
          t_step = t_out - t_now;
          brk = true;
          break;
        }
        H[(size_t)(jj - 1) * mh + jj] = make_hipDoubleComplex(hj1j, 0.0);
      }
      j_done = j;
    }
    if (!brk) {
      zc* u = V + (size_t)n * (m + 1);
      op(V + (size_t)n * m, u);
      ++info.nmult;
      avnorm = znorm_dev(n, u, hb + 128, st);
    }
    H[(size_t)m * mh + m + 1] = make_hipDoubleComplex(1.0, 0.0);
    double err_loc = 0.0;
    int mx = 0;
    while (true) {
      mx = mbrkdwn + k1;
      std::vector<zc> Hc((size_t)mx * mx);
      for (int c = 0; c < mx; ++c)
        for (int r = 0; r < mx; ++r) Hc[(size_t)c * mx + r] = H[(size_t)c * mh + r];
      TTF_HIP(hipMemcpyAsync(Hd, Hc.data(), sizeof(zc) * mx * mx, hipMemcpyHostToDevice, st));
      zexpm_small(mx, sgn * t_step, Hd, Ed, st);
      ++info.nexph;
      read_back(E.data(), Ed, sizeof(zc) * mx, st);   // first column
      if (k1 == 0) {
        err_loc = tol;
      } else {
        const double q1 = std::hypot(E[m].x, E[m].y) * beta;
        const double q2 = std::hypot(E[m + 1].x, E[m + 1].y) * beta * avnorm;
        if (q1 > 10.0 * q2) { err_loc = q2; xm = 1.0 / m; }
        else if (q1 > q2) { err_loc = (q1 * q2) / (q1 - q2); xm = 1.0 / m; }
        else { err_loc = q1; xm = 1.0 / (m - 1); }
      }
      if (k1 != 0 && err_loc > delta * t_step * tol) {
        t_step = gamma * t_step * std::pow(t_step * tol / err_loc, xm);
        p1 = std::pow(10.0, nint(std::log10(t_step) - SQR1) - 1);
        t_step = std::trunc(t_step / p1 + 0.55) * p1;
        continue;
      }
      break;
    }
    mx = mbrkdwn + std::max(0, k1 - 1);
    std::vector<zc> coefs(mx);
    for (int i = 0; i < mx; ++i) coefs[i] = make_hipDoubleComplex(beta * E[i].x, beta * E[i].y);
    TTF_HIP(hipMemcpyAsync(hb, coefs.data(), sizeof(zc) * mx, hipMemcpyHostToDevice, st));
    zgemm_cm('N', 'N', (int)n, 1, mx, make_hipDoubleComplex(1.0, 0.0), V, n, hb, mx, make_hipDoubleComplex(0.0, 0.0),
             w, n, st);
    beta = znorm_dev(n, w, hb + 128, st);
    hump = std::max(hump, beta);
    t_new = gamma * t_step * std::pow(t_step * tol / err_loc, xm);
    p1 = std::pow(10.0, nint(std::log10(t_new) - SQR1) - 1);
    t_new = std::trunc(t_new / p1 + 0.55) * p1;
    t_now += t_step;
    if (info.nstep >= mxstep) break;
  }
  info.hump = vnorm > 0 ? hump / vnorm : 0.0;
  return info;
}

}  // namespace

// w = exp(tau A) v (explib.f90:53-125): v normalised, ZGEXPV, scaled back;
// the 1 x 1 problem directly.  v and w are distinct device vectors.
ExpInfo zexp_mv(int64_t n, int order, double tau, const zc* v, zc* w, double tol, double anorm, const ZOp& op,
                hipStream_t st) {
  const int m1 = (int)std::min<int64_t>(order, n);
  zc* tmp = (zc*)arena(kArenaSolver2, sizeof(zc) * (size_t)(n + 4));
  if (m1 == 1) {
    const zc one{1.0, 0.0};
    TTF_HIP(hipMemcpyAsync(tmp, &one, sizeof(zc), hipMemcpyHostToDevice, st));
    op(tmp, tmp + 1);
    zc a, v0;
    TTF_HIP(hipMemcpyAsync(&a, tmp + 1, sizeof(zc), hipMemcpyDeviceToHost, st));
    read_back(&v0, v, sizeof(zc), st);
    const double ex = std::exp(a.x * tau);
    const zc e = make_hipDoubleComplex(ex * std::cos(a.y * tau), ex * std::sin(a.y * tau));
    const zc r = make_hipDoubleComplex(v0.x * e.x - v0.y * e.y, v0.x * e.y + v0.y * e.x);
    TTF_HIP(hipMemcpyAsync(w, &r, sizeof(zc), hipMemcpyHostToDevice, st));
    ExpInfo info;
    info.nmult = 1;
    return info;
  }
  const double nrm = znorm_dev(n, v, tmp, st);
  if (!(nrm > 0.0)) {
    TTF_HIP(hipMemsetAsync(w, 0, sizeof(zc) * n, st));
    return ExpInfo{};
  }
  zscale(n, make_hipDoubleComplex(1.0 / nrm, 0.0), v, tmp, st);
  ExpInfo info = zgexpv(n, m1, tau, tmp, w, tol, anorm, op, st);
  zscale(n, make_hipDoubleComplex(nrm, 0.0), w, w, st);
  return info;
}

// The panel-blocked complex QR (TTF_ZQR_PANELS=0 selects the column-by-column
// kernel).
static bool zqr_panels() {
  static const bool on = [] {
    const char* e = std::getenv("TTF_ZQR_PANELS");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

void zqr(int rows, int cols, const zc* A, zc* Q, zc* R, hipStream_t st) {
  const int rnew = std::min(rows, cols);
  TTF_REQUIRE((int64_t)rows * cols <= kQrLds, "zqr: %d x %d block exceeds %d entries", rows, cols, kQrLds);
  TTF_REQUIRE(rnew <= 64, "zqr: %d columns exceed 64", rnew);
  static std::atomic<bool> attr{false};   // concurrent first calls only repeat the call
  if (!attr) {
    TTF_HIP(hipFuncSetAttribute((const void*)zqr_gs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(zc) * kQrLds)));
    TTF_HIP(hipFuncSetAttribute((const void*)zqr_panel_gs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(zc) * kQrLds)));
    attr = true;
  }
  if (zqr_panels())
    hipLaunchKernelGGL(zqr_panel_gs_kernel, dim3(1), dim3(256), sizeof(zc) * (size_t)rows * cols, st, rows, cols,
                       rnew, A, Q);
  else
    hipLaunchKernelGGL(zqr_gs_kernel, dim3(1), dim3(256), sizeof(zc) * (size_t)rows * cols, st, rows, cols, rnew, A,
                       Q);
  TTF_HIP(hipGetLastError());
  // R = Q^H A (rnew x cols)
  zgemm_cm('C', 'N', rnew, cols, rows, make_hipDoubleComplex(1.0, 0.0), Q, rows, A, rows,
           make_hipDoubleComplex(0.0, 0.0), R, rnew, st);
}

// The S-step operator of bond k (zsfun_matvec, tt_ksl.f90:130-136) and its
// adjoint: Sy (ry1 x ry2) = phi1 (ry1 x rx1 ra) [Sx (rx1 x rx2) phi2 (rx2 x ra ry2)].
// Local tolerance of the exponentials whose core products are complex64.
constexpr double kC64Tol = 1e-6;

// complex64 copy of one local operator (the KSL local exponential's matvec
// in reduced precision): interfaces and core rounded once per local step,
// each product rounds x, runs the chain on the fp32 matrix cores and widens
// y back; the Krylov basis, Hessenberg matrix and exponential stay
// complex128.
struct CMixedOp {
  Shape s;
  hipFloatComplex *phi1 = nullptr, *A = nullptr, *phi2 = nullptr, *x = nullptr, *y = nullptr;
  void apply(const zc* xin, zc* yout, hipStream_t st) const {
    const int64_t nx = (int64_t)s.rx1 * s.m * s.rx2, ny = (int64_t)s.ry1 * s.n * s.ry2;
    to_c64(nx, xin, x, st);
    cbfun3_chain(s, 1, phi1, A, phi2, x, nx, y, ny, st);
    to_c128(ny, y, yout, st);
  }
};

CMixedOp make_cmixed(const Shape& s, const zc* phi1, const zc* A, const zc* phi2, hipStream_t st) {
  const int64_t n1 = (int64_t)s.ry1 * s.rx1 * s.ra1, na = (int64_t)s.ra1 * s.n * s.m * s.ra2;
  const int64_t n2 = (int64_t)s.rx2 * s.ra2 * s.ry2;
  const int64_t nx = (int64_t)s.rx1 * s.m * s.rx2, ny = (int64_t)s.ry1 * s.n * s.ry2;
  hipFloatComplex* base =
      (hipFloatComplex*)arena(kArenaMixed, sizeof(hipFloatComplex) * (size_t)(n1 + na + n2 + nx + ny + 64));
  CMixedOp mo;
  mo.s = s;
  mo.phi1 = base;
  mo.A = mo.phi1 + n1;
  mo.phi2 = mo.A + na;
  mo.x = mo.phi2 + n2;
  mo.y = mo.x + nx;
  to_c64(n1, phi1, mo.phi1, st);
  to_c64(na, A, mo.A, st);
  to_c64(n2, phi2, mo.phi2, st);
  return mo;
}

struct SOp {
  int rx1, ry1, ra, rx2, ry2;
  const zc* phi1;
  const zc* phi2;
  zc* work;   // rx1 * ra * ry2
  void apply(const zc* Sx, zc* Sy, hipStream_t st) const {
    const zc one{1.0, 0.0}, zero{0.0, 0.0};
    zgemm_cm('N', 'N', rx1, ra * ry2, rx2, one, Sx, rx1, phi2, rx2, zero, work, rx1, st);
    zgemm_cm('N', 'N', ry1, ry2, rx1 * ra, one, phi1, ry1, work, rx1 * ra, zero, Sy, ry1, st);
  }
  void adjoint(const zc* Sy, zc* Sx, hipStream_t st) const {
    const zc one{1.0, 0.0}, zero{0.0, 0.0};
    // T (rx1 ra x ry2) = phi1^H Sy;  Sx (rx1 x rx2) = T (rx1 x ra ry2) conj(phi2)^T
    zgemm_cm('C', 'N', rx1 * ra, ry2, ry1, one, phi1, ry1, Sy, ry1, zero, work, rx1 * ra, st);
    zgemm_cm('N', 'C', rx1, rx2, ra * ry2, one, work, rx1, phi2, rx2, zero, Sx, rx1, st);
  }
};

}  // namespace ttf

using namespace ttf;

namespace {

struct ZBuf {
  zc* p = nullptr;
  size_t cap = 0;
  ZBuf() = default;
  ZBuf(const ZBuf&) = delete;
  ZBuf& operator=(const ZBuf&) = delete;
  zc* get(size_t n) {
    if (n > cap) {
      if (p) TTF_HIP(hipFree(p));
      p = nullptr;
      TTF_HIP(hipMalloc(&p, sizeof(zc) * std::max<size_t>(n, 1)));
      cap = n;
    }
    return p;
  }
  ~ZBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" int ttf_ztt_ksl(int d, const int* n, const int* m, const int* ra, const ttf_zdouble* crA,
                           const ttf_zdouble* crY0, int* ry, double tau, int rmax, int kickrank, int nswp, int verb,
                           int typ, int order, int usenrm, ttf_zdouble* result, int64_t result_cap,
                           int64_t* result_size, ttf_ksl_stats* stats) {
  return guarded([&] {
    (void)rmax; (void)kickrank; (void)nswp;   // fixed-rank scheme: accepted and unused, as in the reference
    TTF_REQUIRE(d >= 1, "ttf_ztt_ksl: d=%d", d);
    for (int k = 0; k < d; ++k) TTF_REQUIRE(n[k] == m[k], "ttf_ztt_ksl: operator must be square");
    TTF_REQUIRE(ry[0] == 1 && ry[d] == 1 && ra[0] == 1 && ra[d] == 1, "ttf_ztt_ksl: boundary ranks must be 1");
    TTF_REQUIRE(typ == 1 || typ == 2, "ttf_ztt_ksl: typ=%d (1 KSL, 2 symmetrised)", typ);
    TTF_REQUIRE(order >= 1 && order <= kExpmMax - 2, "ttf_ztt_ksl: order=%d outside 1..%d", order, kExpmMax - 2);
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const double tol = 1e-8;   // tt_ksl.f90:474
    int R = 1, nmax = 1, ramax = 1;
    for (int k = 0; k <= d; ++k) { R = std::max(R, ry[k]); ramax = std::max(ramax, ra[k]); }
    for (int k = 0; k < d; ++k) nmax = std::max(nmax, n[k]);
    std::vector<int64_t> pa(d + 1, 0);
    for (int k = 0; k < d; ++k) pa[k + 1] = pa[k] + (int64_t)ra[k] * n[k] * m[k] * ra[k + 1];
    ZBuf bA, bcur, bQ, bR, bS, bphit, bwork;
    zc* dA = bA.get(pa[d]);
    TTF_HIP(hipMemcpyAsync(dA, crA, sizeof(zc) * pa[d], hipMemcpyHostToDevice, st));
    std::vector<ZBuf> cr(d), phi(d + 1);
    const size_t core_cap = (size_t)R * nmax * R;
    for (int k = 0; k < d; ++k) cr[k].get(core_cap);
    for (int k = 0; k <= d; ++k) phi[k].get((size_t)R * R * ramax);
    {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)ry[k] * n[k] * ry[k + 1];
        TTF_HIP(hipMemcpyAsync(cr[k].p, crY0 + off, sizeof(zc) * sz, hipMemcpyHostToDevice, st));
        off += sz;
      }
    }
    const zc one{1.0, 0.0};
    TTF_HIP(hipMemcpyAsync(phi[0].p, &one, sizeof(zc), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phi[d].p, &one, sizeof(zc), hipMemcpyHostToDevice, st));
    zc* cur = bcur.get(core_cap);
    zc* Q = bQ.get(core_cap);
    zc* Rm = bR.get(core_cap);
    zc* St = bS.get((size_t)R * R);
    zc* phit = bphit.get((size_t)R * R * ramax);
    zc* work = bwork.get(core_cap + (size_t)R * R * ramax);
    long nmult = 0, nsteps = 0, nsolves = 0, nmult32 = 0;
    double flops = 0.0;   // algorithmic real flops of the Krylov operator applications
    // complex64 local matvecs in the local exponentials (the S-steps, the
    // norm estimates and everything but the core products stay complex128)
    const bool c64 = local_matvec_bits() == 32;
    // initial left-to-right QR and left interfaces (tt_ksl.f90:538-556)
    for (int i = 0; i < d - 1; ++i) {
      const int rows = ry[i] * n[i], cols = ry[i + 1];
      const int rnew = std::min(rows, cols);
      zqr(rows, cols, cr[i].p, Q, Rm, st);
      TTF_HIP(hipMemcpyAsync(cr[i].p, Q, sizeof(zc) * (size_t)rows * rnew, hipMemcpyDeviceToDevice, st));
      const int ncol2 = n[i + 1] * ry[i + 2];
      zgemm_cm('N', 'N', rnew, ncol2, cols, one, Rm, rnew, cr[i + 1].p, cols, zc{0.0, 0.0}, work, rnew, st);
      TTF_HIP(hipMemcpyAsync(cr[i + 1].p, work, sizeof(zc) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice, st));
      ry[i + 1] = rnew;
      Shape s{ry[i], m[i], ry[i + 1], ry[i], n[i], ry[i + 1], ra[i], ra[i + 1]};
      zphi_left_chain(s, phi[i].p, dA + pa[i], cr[i].p, cr[i].p, phi[i + 1].p, st);
    }
    const double tau0 = typ == 1 ? tau : tau / 2.0;
    auto local_exp = [&](int i, const zc* v, zc* w, double t) {
      Shape s{ry[i], m[i], ry[i + 1], ry[i], n[i], ry[i + 1], ra[i], ra[i + 1]};
      const int64_t N = (int64_t)ry[i] * n[i] * ry[i + 1];
      const zc* p1 = phi[i].p;
      const zc* Ai = dA + pa[i];
      const zc* p2 = phi[i + 1].p;
      ZOp op = [&, s, p1, Ai, p2, N](const zc* x, zc* y) { zbfun3_chain(s, 1, p1, Ai, p2, x, N, y, N, st); };
      ZOp adj = [&, s, p1, Ai, p2, N](const zc* x, zc* y) { zbfun3_transp_chain(s, 1, p1, Ai, p2, x, N, y, N, st); };
      const double anorm = usenrm == 2 ? 1.0 : std::max(hager_norm1(N, op, adj, st), 1e-300);
      ExpInfo info;
      if (c64) {
        // the Krylov error estimate cannot fall below the complex64 rounding
        // of the products, so these exponentials run at tolerance kC64Tol
        const CMixedOp mo = make_cmixed(s, p1, Ai, p2, st);
        ZOp op32 = [&, mo](const zc* x, zc* y) { mo.apply(x, y, st); };
        info = zexp_mv(N, order, t, v, w, std::max(tol, kC64Tol), anorm, op32, st);
        nmult32 += info.nmult;
      } else {
        info = zexp_mv(N, order, t, v, w, tol, anorm, op, st);
      }
      nmult += info.nmult;
      nsteps += info.nstep;
      ++nsolves;
      flops += 4.0 * info.nmult * bfun3_flops(s);
    };
    auto s_exp = [&](const SOp& so, const zc* v, zc* w, double t) {
      const int64_t N = (int64_t)so.ry1 * so.ry2;
      ZOp op = [&](const zc* x, zc* y) { so.apply(x, y, st); };
      ZOp adj = [&](const zc* x, zc* y) { so.adjoint(x, y, st); };
      const double anorm = usenrm == 2 ? 1.0 : std::max(hager_norm1(N, op, adj, st), 1e-300);
      ExpInfo info = zexp_mv(N, order, t, v, w, tol, anorm, op, st);
      nmult += info.nmult;
      nsteps += info.nstep;
      ++nsolves;
      flops += 4.0 * info.nmult *
               (2.0 * so.rx1 * so.rx2 * so.ra * so.ry2 + 2.0 * so.ry1 * so.ry2 * so.rx1 * so.ra);
    };
    auto ztrans = [&](int rows, int cols, const zc* A, zc* B) {   // B (cols x rows) = A^T
      hipLaunchKernelGGL(ztranspose_kernel, dim3((unsigned)ceil_div((int64_t)rows * cols, 256)), dim3(256), 0, st, rows,
                         cols, A, B);
      TTF_HIP(hipGetLastError());
    };
    int i = d - 1, dir = -1;
    while (true) {
      local_exp(i, cr[i].p, cur, tau0);
      if (dir < 0 && i > 0) {
        // transposed QR of the evolved core: Q (n ry_{i+1} x rnew), R (rnew x ry_i)
        const int rows = n[i] * ry[i + 1], cols = ry[i];
        const int rnew = std::min(rows, cols);
        ztrans(cols, rows, cur, work);
        zqr(rows, cols, work, Q, Rm, st);
        ztrans(rows, rnew, Q, cr[i].p);                      // core i = Q^T (rnew x n ry_{i+1})
        ztrans(rnew, cols, Rm, work);                        // R^T (ry_i x rnew)
        Shape s{rnew, m[i], ry[i + 1], rnew, n[i], ry[i + 1], ra[i], ra[i + 1]};
        zphi_right_chain(s, phi[i + 1].p, dA + pa[i], cr[i].p, cr[i].p, phit, st);
        SOp so{ry[i], ry[i], ra[i], rnew, rnew, phi[i].p, phit, work + (size_t)cols * rnew};
        s_exp(so, work, St, -tau0);
        const int mp = ry[i - 1] * n[i - 1];
        zgemm_cm('N', 'N', mp, rnew, cols, one, cr[i - 1].p, mp, St, cols, zc{0.0, 0.0}, Q, mp, st);
        TTF_HIP(hipMemcpyAsync(cr[i - 1].p, Q, sizeof(zc) * (size_t)mp * rnew, hipMemcpyDeviceToDevice, st));
        ry[i] = rnew;
        TTF_HIP(hipMemcpyAsync(phi[i].p, phit, sizeof(zc) * (size_t)rnew * ra[i] * rnew, hipMemcpyDeviceToDevice, st));
      } else if (dir > 0 && i < d - 1) {
        const int rows = ry[i] * n[i], cols = ry[i + 1];
        const int rnew = std::min(rows, cols);
        zqr(rows, cols, cur, Q, Rm, st);
        TTF_HIP(hipMemcpyAsync(cr[i].p, Q, sizeof(zc) * (size_t)rows * rnew, hipMemcpyDeviceToDevice, st));
        Shape s{ry[i], m[i], rnew, ry[i], n[i], rnew, ra[i], ra[i + 1]};
        zphi_left_chain(s, phi[i].p, dA + pa[i], cr[i].p, cr[i].p, phit, st);
        SOp so{rnew, rnew, ra[i + 1], cols, cols, phit, phi[i + 1].p, work};
        s_exp(so, Rm, St, -tau0);
        const int ncol2 = n[i + 1] * ry[i + 2];
        zgemm_cm('N', 'N', rnew, ncol2, cols, one, St, rnew, cr[i + 1].p, cols, zc{0.0, 0.0}, Q, rnew, st);
        TTF_HIP(hipMemcpyAsync(cr[i + 1].p, Q, sizeof(zc) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice, st));
        ry[i + 1] = rnew;
        TTF_HIP(hipMemcpyAsync(phi[i + 1].p, phit, sizeof(zc) * (size_t)rnew * ra[i + 1] * rnew,
                               hipMemcpyDeviceToDevice, st));
      } else {
        TTF_HIP(hipMemcpyAsync(cr[i].p, cur, sizeof(zc) * (size_t)ry[i] * n[i] * ry[i + 1], hipMemcpyDeviceToDevice,
                               st));
      }
      if (verb >= 2) std::fprintf(stderr, "ksl: core %d dir %d ry %d\n", i + 1, dir, ry[i]);
      if (dir > 0 && i == d - 1) break;
      if (dir < 0 && i == 0) {
        if (typ == 1) break;
        dir = 1;
        continue;
      }
      i += dir;
    }
    // result cores (the reference's dyn_tt::zresult_core)
    int64_t total_sz = 0;
    for (int k = 0; k < d; ++k) total_sz += (int64_t)ry[k] * n[k] * ry[k + 1];
    if (result_size) *result_size = total_sz;
    if (result) {
      TTF_REQUIRE(result_cap >= total_sz, "ttf_ztt_ksl: result buffer holds %lld entries, need %lld",
                  (long long)result_cap, (long long)total_sz);
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)ry[k] * n[k] * ry[k + 1];
        TTF_HIP(hipMemcpyAsync(result + off, cr[k].p, sizeof(zc) * sz, hipMemcpyDeviceToHost, st));
        off += sz;
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
    if (stats) {
      stats->nsolves = nsolves;
      stats->matvecs = nmult;
      stats->krylov_steps = nsteps;
      stats->flops = flops;
      stats->matvecs32 = nmult32;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

// ---------------------------------------------------------------- C ABI of the building blocks
extern "C" {

int ttf_zexpm(int m, double t, const ttf_zdouble* H, ttf_zdouble* E) {
  return guarded([&] {
    zexpm_small(m, t, (const zc*)H, (zc*)E, stream());
    TTF_HIP(hipStreamSynchronize(stream()));
  });
}

int ttf_zexp_bfun3(int rx1, int n, int rx2, int ra1, int ra2, const ttf_zdouble* phi1, const ttf_zdouble* A,
                   const ttf_zdouble* phi2, const ttf_zdouble* v, ttf_zdouble* w, double tau, int order, double tol,
                   double* anorm_out, int* nmult) {
  return guarded([&] {
    hipStream_t st = stream();
    Shape s{rx1, n, rx2, rx1, n, rx2, ra1, ra2};
    const int64_t N = (int64_t)rx1 * n * rx2;
    const zc* p1 = (const zc*)phi1;
    const zc* Ad = (const zc*)A;
    const zc* p2 = (const zc*)phi2;
    ZOp op = [&](const zc* x, zc* y) { zbfun3_chain(s, 1, p1, Ad, p2, x, N, y, N, st); };
    ZOp adj = [&](const zc* x, zc* y) { zbfun3_transp_chain(s, 1, p1, Ad, p2, x, N, y, N, st); };
    const double an = std::max(hager_norm1(N, op, adj, st), 1e-300);
    ExpInfo info = zexp_mv(N, order, tau, (const zc*)v, (zc*)w, tol, an, op, st);
    TTF_HIP(hipStreamSynchronize(st));
    if (anorm_out) *anorm_out = an;
    if (nmult) *nmult = info.nmult;
  });
}

int ttf_zqr(int rows, int cols, const ttf_zdouble* A, ttf_zdouble* Q, ttf_zdouble* R) {
  return guarded([&] {
    zqr(rows, cols, (const zc*)A, (zc*)Q, (zc*)R, stream());
    TTF_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"

extern "C" int ttf_cbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
                          const ttf_zfloat* phi1, const ttf_zfloat* A, const ttf_zfloat* phi2, const ttf_zfloat* x,
                          int64_t ldx, ttf_zfloat* y, int64_t ldy) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && m > 0 && rx2 > 0 && ry1 > 0 && n > 0 && ry2 > 0 && ra1 > 0 && ra2 > 0,
                "ttf_cbfun3: invalid sizes");
    TTF_REQUIRE(ldx >= (int64_t)rx1 * m * rx2 || nvec <= 1, "ttf_cbfun3: ldx too small");
    TTF_REQUIRE(ldy >= (int64_t)ry1 * n * ry2 || nvec <= 1, "ttf_cbfun3: ldy too small");
    cbfun3_chain(Shape{rx1, m, rx2, ry1, n, ry2, ra1, ra2}, nvec, (const hipFloatComplex*)phi1,
                 (const hipFloatComplex*)A, (const hipFloatComplex*)phi2, (const hipFloatComplex*)x, ldx,
                 (hipFloatComplex*)y, ldy, stream());
  });
}

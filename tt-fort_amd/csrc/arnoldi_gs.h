// One Arnoldi step's orthogonalisation in four launches: classical
// Gram-Schmidt, twice, of a new Krylov vector u against the basis V(:, 0..j-1),
// then its normalisation, with every coefficient left on the device.
//
//   1. partial dot products conj(V(:, i)) . u over row chunks;
//   2. each work-group sums the partials (the same fixed order in every
//      work-group, so all of them hold bit-identical coefficients h), updates
//      its rows u -= V h, and forms the second pass's partial dot products
//      from the updated rows it still holds in LDS;
//   3. the same with the second pass's partials, ending with the partial
//      squared norms of u;
//   4. every work-group sums the norm partials and scales its rows of u.
//
// The coefficients land in one 129-entry column: hs[0 .. j) from the first
// pass, hs[64 .. 64 + j) from the second, hs[128] = |u|^2 after both passes
// (real part).  In place of the 12 launches of separate dot-product,
// reduction, GEMM-update and norm kernels, and without the one- and
// two-block reductions whose latency dominated small Krylov vectors.
#pragma once

#include "ttf_common.h"
#include "wave.h"

#include <algorithm>

namespace ttf {
namespace gs {

constexpr int kMaxCols = 64;     // basis columns per pass (one Hessenberg column)
constexpr int kMaxParts = 256;   // row chunks, one per work-group
constexpr int kMaxChunk = 1024;  // rows per chunk (LDS copy of u)
constexpr int kThreads = 256;

using zc = hipDoubleComplex;

__device__ __forceinline__ void set_zero(double& x) { x = 0.0; }
__device__ __forceinline__ void set_zero(zc& x) { x = make_hipDoubleComplex(0.0, 0.0); }
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ zc add(zc a, zc b) { return make_hipDoubleComplex(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ zc sub(zc a, zc b) { return make_hipDoubleComplex(a.x - b.x, a.y - b.y); }
// acc + conj(a) b
__device__ __forceinline__ double dotc_acc(double acc, double a, double b) { return fma(a, b, acc); }
__device__ __forceinline__ zc dotc_acc(zc acc, zc a, zc b) {
  return make_hipDoubleComplex(fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.x, b.y, fma(-a.y, b.x, acc.y)));
}
// acc + a b
__device__ __forceinline__ double mul_acc(double acc, double a, double b) { return fma(a, b, acc); }
__device__ __forceinline__ zc mul_acc(zc acc, zc a, zc b) {
  return make_hipDoubleComplex(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
__device__ __forceinline__ double abs2(double a) { return a * a; }
__device__ __forceinline__ double abs2(zc a) { return fma(a.x, a.x, a.y * a.y); }
__device__ __forceinline__ double scaled(double a, double s) { return a * s; }
__device__ __forceinline__ zc scaled(zc a, double s) { return make_hipDoubleComplex(a.x * s, a.y * s); }
__device__ __forceinline__ double wave_sum(double a) { return wave_sum64(a); }
__device__ __forceinline__ zc wave_sum(zc a) { return make_hipDoubleComplex(wave_sum64(a.x), wave_sum64(a.y)); }
__device__ __forceinline__ void store_real(double* p, double v) { *p = v; }
__device__ __forceinline__ void store_real(zc* p, double v) { *p = make_hipDoubleComplex(v, 0.0); }

// MODE 0: part_out[b j + i] = sum over chunk b of conj(V(r, i)) u(r).
// MODE 1: h = sum over b of part_in[b j + i] (work-group 0 stores it in
//         h_out), u -= V h over the chunk, then MODE 0's partials of the new u.
// MODE 2: h and the update as MODE 1, then norm_out[b] = sum over the chunk
//         of |u(r)|^2.
template <class T, int MODE>
__global__ __launch_bounds__(kThreads) void gs_pass_kernel(int64_t N, int j, int chunk, const T* __restrict__ V,
                                                          int64_t ldv, T* __restrict__ u,
                                                          const T* __restrict__ part_in, int nparts,
                                                          T* __restrict__ h_out, T* __restrict__ part_out,
                                                          double* __restrict__ norm_out) {
  __shared__ T us[kMaxChunk];
  __shared__ T hsh[kMaxCols];
  __shared__ T red[4][kMaxCols];
  __shared__ double nred[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int rows = (int)std::min<int64_t>(chunk, N - r0);
  if (MODE == 0) {
    for (int r = tid; r < rows; r += kThreads) us[r] = u[r0 + r];
  } else {
    {
      const int col = tid & 63, grp = tid >> 6;
      T s;
      set_zero(s);
      if (col < j) {
        // eight partials in flight per step, added in the same order
        int b = grp;
        for (; b + 28 < nparts; b += 32) {
          T pv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) pv[u] = part_in[(int64_t)(b + 4 * u) * j + col];
#pragma unroll
          for (int u = 0; u < 8; ++u) s = add(s, pv[u]);
        }
        for (; b < nparts; b += 4) s = add(s, part_in[(int64_t)b * j + col]);
      }
      red[grp][col] = s;
    }
    __syncthreads();
    // coefficients j .. j8 - 1 are zero, so that the update runs in whole
    // steps of eight columns (their loads in flight together, no remainder
    // loop); the padded columns read column j - 1 and add exact zeros
    const int j8 = (j + 7) & ~7;
    if (tid < j) {
      const T h = add(add(red[0][tid], red[1][tid]), add(red[2][tid], red[3][tid]));
      hsh[tid] = h;
      if (blockIdx.x == 0) h_out[tid] = h;
    } else if (tid < j8) {
      set_zero(hsh[tid]);
    }
    __syncthreads();
    for (int r = tid; r < rows; r += kThreads) {
      const T* v = V + r0 + r;
      T s;
      set_zero(s);
      for (int i0 = 0; i0 < j8; i0 += 8) {
        T vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) vv[u] = v[ldv * min(i0 + u, j - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) s = mul_acc(s, vv[u], hsh[i0 + u]);
      }
      const T x = sub(u[r0 + r], s);
      u[r0 + r] = x;
      us[r] = x;
    }
  }
  __syncthreads();
  if (MODE == 2) {
    double s = 0.0;
    for (int r = tid; r < rows; r += kThreads) s += abs2(us[r]);
    s = wave_sum64(s);
    if (lane == 0) nred[wv] = s;
    __syncthreads();
    if (tid == 0) norm_out[blockIdx.x] = (nred[0] + nred[1]) + (nred[2] + nred[3]);
    return;
  }
  // wave wv takes columns wv, wv + 4, ..., four of them per pass over the
  // rows so that their loads are in flight together
  for (int i0 = wv; i0 < j; i0 += 16) {
    T acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) set_zero(acc[q]);
    for (int r = lane; r < rows; r += 64) {
      const T x = us[r];
      T a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + 4 * q;
        if (i < j) a[q] = V[ldv * i + r0 + r];
        else set_zero(a[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = dotc_acc(acc[q], a[q], x);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 4 * q;
      if (i < j) {   // uniform over the wave
        const T s = wave_sum(acc[q]);
        if (lane == 0) part_out[(int64_t)blockIdx.x * j + i] = s;
      }
    }
  }
}

// Sums the norm partials (the same order in every work-group), stores the
// total in hs_norm (work-group 0) and writes out = u / |u|, or zeros when
// |u| <= tol; with out == nullptr only the total is stored.
template <class T>
__global__ __launch_bounds__(kThreads) void gs_normalise_kernel(int64_t N, const double* __restrict__ norm_part,
                                                               int nparts, T* __restrict__ hs_norm, double tol,
                                                               const T* u, T* out) {
  __shared__ double nred[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double s = tid < nparts ? norm_part[tid] : 0.0;
  s = wave_sum64(s);
  if (lane == 0) nred[wv] = s;
  __syncthreads();
  const double tot = (nred[0] + nred[1]) + (nred[2] + nred[3]);
  if (blockIdx.x == 0 && tid == 0) store_real(hs_norm, tot);
  if (out == nullptr) return;
  const double h = sqrt(fmax(tot, 0.0));
  const double inv = h > tol ? 1.0 / h : 0.0;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + tid; e < N; e += (int64_t)gridDim.x * kThreads)
    out[e] = scaled(u[e], inv);
}

// Rows per chunk for N rows: a multiple of 128 that keeps the chunk count
// within kMaxParts; 0 when N is beyond the fused kernels' reach.
inline int chunk_rows(int64_t N) {
  const int64_t c = 128 * ceil_div(N, (int64_t)128 * kMaxParts);
  return c <= kMaxChunk ? (int)c : 0;
}

// Whether step() takes (N, j); the callers keep their separate-kernel path
// for the rest.
inline bool fits(int64_t N, int j) { return N >= 1 && j >= 1 && j <= kMaxCols && chunk_rows(N) > 0; }

// Orthogonalises u (N) against V(:, 0..j-1) (leading dimension ldv) twice,
// stores the column hs as described above, and writes the normalised vector
// to out (may be u; nullptr to skip).  A vector whose norm after both passes
// is at most tol is written as zeros.  Requires fits(N, j).
template <class T>
void step(int64_t N, int j, const T* V, int64_t ldv, T* u, T* hs, double tol, T* out, hipStream_t st) {
  TTF_REQUIRE(fits(N, j), "gram-schmidt: %lld rows x %d columns outside the fused kernels' limits", (long long)N, j);
  const int chunk = chunk_rows(N);
  const int nparts = (int)ceil_div(N, chunk);
  T* pa = (T*)arena(kArenaGram, sizeof(T) * (size_t)2 * kMaxParts * kMaxCols + sizeof(double) * kMaxParts);
  T* pb = pa + (size_t)kMaxParts * kMaxCols;
  double* pn = (double*)(pb + (size_t)kMaxParts * kMaxCols);
  const dim3 grid((unsigned)nparts), blk(kThreads);
  hipLaunchKernelGGL((gs_pass_kernel<T, 0>), grid, blk, 0, st, N, j, chunk, V, ldv, u, (const T*)nullptr, nparts,
                     (T*)nullptr, pa, (double*)nullptr);
  hipLaunchKernelGGL((gs_pass_kernel<T, 1>), grid, blk, 0, st, N, j, chunk, V, ldv, u, (const T*)pa, nparts, hs,
                     pb, (double*)nullptr);
  hipLaunchKernelGGL((gs_pass_kernel<T, 2>), grid, blk, 0, st, N, j, chunk, V, ldv, u, (const T*)pb, nparts,
                     hs + 64, (T*)nullptr, pn);
  const unsigned gn = out ? (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(N, kThreads), 1), 1024) : 1u;
  hipLaunchKernelGGL((gs_normalise_kernel<T>), dim3(gn), blk, 0, st, N, (const double*)pn, nparts, hs + 128, tol,
                     (const T*)u, out);
  TTF_HIP(hipGetLastError());
}

}  // namespace gs
}  // namespace ttf

// Strided, batched GEMM on the CDNA4 matrix cores (gfx950).
//
//   C[b](i,j) = alpha * sum_k A[b](i,k) * B[b](k,j) + beta * C[b](i,j)
//
// Every operand is addressed through explicit element strides, so the
// transposes and index permutations of the reference's GEMM chains
// (dgemm + dtransp, ttals.f90 / tt_linalg.f90) become stride choices and are
// never materialised.  Batch index b is split in two levels,
// b -> (b % nb0, b / nb0), each with its own stride, which covers "one GEMM
// per vector of a block" and "one GEMM per bond index" at once.
//
// Tiling: 256 threads = 4 wave64s arranged 2x2; each wave owns a
// (BM/2)x(BN/2) sub-tile made of 16x16 MFMA fragments
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  K advances in chunks of
// 16 staged through LDS; the next chunk is prefetched into registers while
// the current one is consumed.  The global->LDS loader picks the coalesced
// index order from the operand strides.  Work-group ids are remapped so that
// neighbouring output tiles land on the same XCD (L2) of the 8-XCD chip.
//
// Split-K: with ksplit > 1 each work-group reduces a K range into a
// workspace slice; gemm_splitk_reduce then sums the slices in a fixed order
// (deterministic) and applies alpha/beta.  Every output element is summed
// over K in increasing order, four products per MFMA, whatever the tile
// shape and K chunk, so only ksplit changes the rounding of a product.
#pragma once

#include "ttf_common.h"

namespace ttf {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Acc4;
template <> struct Acc4<double> { typedef d4_t type; };
template <> struct Acc4<float> { typedef f4_t type; };

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 result block held by (lane, accumulator register r).
// Measured on gfx950: the f64 16x16x4 result interleaves the four registers
// across row groups (row = lane/16 + 4r); the f32 16x16x4 result keeps them
// contiguous (row = 4*(lane/16) + r).  The column is lane%16 in both.
template <typename T> __device__ __forceinline__ int mfma_row(int lane, int r);
template <> __device__ __forceinline__ int mfma_row<double>(int lane, int r) {
  return (lane >> 4) + 4 * r;
}
template <> __device__ __forceinline__ int mfma_row<float>(int lane, int r) {
  return 4 * (lane >> 4) + r;
}

template <typename T>
struct GemmOp {
  int M, N, K;
  const T* A; int64_t sAi, sAk, sAb0, sAb1;
  const T* B; int64_t sBk, sBj, sBb0, sBb1;
  T* C;       int64_t sCi, sCj, sCb0, sCb1;
  int nb0;    // inner batch extent (>= 1)
  int nbatch; // total batch count
  T alpha, beta;
  int ksplit;     // number of K ranges (>= 1)
  T* work;        // ksplit * nbatch * M * N workspace when ksplit > 1
};

// Map the hardware work-group id so that consecutive logical tiles share an
// XCD: the dispatcher hands work-groups to the 8 XCDs round-robin.
__device__ __forceinline__ int xcd_remap(int wg, int total) {
  const int nx = 8;
  int full = (total / nx) * nx;
  if (wg >= full) return wg;
  int per = total / nx;
  return (wg % nx) * per + (wg / nx);
}

// NW waves per work-group: NW = 4 arranges them 2x2 over the tile, NW = 1
// lets a single wave own the whole tile.  Every wave keeps at least four
// independent 16x16 accumulator chains (a 32x32 wave tile), since dependent
// f64 MFMAs issue back to back only when several chains interleave.
// CT (column-contiguous output, op.sCi == 1): the fragments are multiplied
// in swapped order, B^T A^T, so that each accumulator holds a tile of C^T:
// lane l then owns row l%16 of the 16x16 block and consecutive lanes store
// consecutive rows (contiguous in memory) instead of consecutive columns.
template <typename T, int BM, int BN, int BK, int NW = 4, bool CT = false, int WROWS = 0>
__global__ __launch_bounds__(64 * NW) void gemm_mfma_kernel(GemmOp<T> op) {
  constexpr int NT = 64 * NW;
  constexpr int PAD = 2;
  // wave rows: 2 x 2 waves by default, WROWS x (NW / WROWS) when given
  constexpr int WR = WROWS > 0 ? WROWS : ((NW == 4) ? 2 : 1);
  constexpr int WC = NW / WR;             // wave columns
  constexpr int WM = BM / WR, WN = BN / WC;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int LA = BM * BK / NT;   // A elements per thread
  constexpr int LB = BN * BK / NT;   // B elements per thread
  typedef typename Acc4<T>::type acc_t;

  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WC, wn = wave % WC;

  const int tilesM = (op.M + BM - 1) / BM;
  const int tilesN = (op.N + BN - 1) / BN;
  const int ntiles = tilesM * tilesN;
  const int tile = xcd_remap(blockIdx.x, ntiles);
  const int tm = tile / tilesN, tn = tile % tilesN;

  const int b = blockIdx.y;
  const int b0 = b % op.nb0, b1 = b / op.nb0;
  const T* __restrict__ A = op.A + b0 * op.sAb0 + b1 * op.sAb1;
  const T* __restrict__ B = op.B + b0 * op.sBb0 + b1 * op.sBb1;

  // K range of this work-group.  Split-K ranges are multiples of 64 for every
  // BK, so the slices (and the order of the sums) depend on ksplit alone.
  constexpr int KG = 64;
  static_assert(KG % BK == 0, "split-K granularity must be a multiple of BK");
  const int kchunk = op.ksplit > 1 ? ((op.K + op.ksplit - 1) / op.ksplit + KG - 1) / KG * KG
                                   : ((op.K + BK - 1) / BK) * BK;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(op.K, kbeg + kchunk);

  const bool a_ifast = (op.sAi == 1);
  const bool b_jfast = (op.sBj == 1);

  T ra[LA], rb[LB];
  int ai[LA], ak[LA], bk[LB], bj[LB];
#pragma unroll
  for (int q = 0; q < LA; ++q) {
    int idx = tid + q * NT;
    if (a_ifast) { ai[q] = idx % BM; ak[q] = idx / BM; }
    else         { ak[q] = idx % BK; ai[q] = idx / BK; }
  }
#pragma unroll
  for (int q = 0; q < LB; ++q) {
    int idx = tid + q * NT;
    if (b_jfast) { bj[q] = idx % BN; bk[q] = idx / BN; }
    else         { bk[q] = idx % BK; bj[q] = idx / BK; }
  }

  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      int gi = tm * BM + ai[q], gk = k0 + ak[q];
      ra[q] = (gi < op.M && gk < kend) ? A[gi * op.sAi + gk * op.sAk] : T(0);
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      int gj = tn * BN + bj[q], gk = k0 + bk[q];
      rb[q] = (gj < op.N && gk < kend) ? B[gk * op.sBk + gj * op.sBj] : T(0);
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < LA; ++q) As[ak[q]][ai[q]] = ra[q];
#pragma unroll
    for (int q = 0; q < LB; ++q) Bs[bk[q]][bj[q]] = rb[q];
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = CT ? mfma16(bf[j], af[i], acc[i][j]) : mfma16(af[i], bf[j], acc[i][j]);
    }
  }

  // Epilogue: lane l, register r holds row mfma_row<T>(l, r), column l%16
  // of each 16x16 block (CT: row l%16, column mfma_row<T>(l, r)).
  if (op.ksplit == 1) {
    T* __restrict__ C = op.C + b0 * op.sCb0 + b1 * op.sCb1;
    // with beta != 0 the tile's C elements are all loaded (from clamped
    // in-range addresses) before any is used, so the loads stay in flight
    // together instead of one load and wait per element
    T cv[FM][FN][4];
    if (op.beta != T(0)) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int fr = CT ? (lane & 15) : mfma_row<T>(lane, r);
            const int fc = CT ? mfma_row<T>(lane, r) : (lane & 15);
            const int gi = min(tm * BM + wm * WM + i * 16 + fr, op.M - 1);
            const int gj = min(tn * BN + wn * WN + j * 16 + fc, op.N - 1);
            cv[i][j][r] = C[gi * op.sCi + gj * op.sCj];
          }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int fr = CT ? (lane & 15) : mfma_row<T>(lane, r);
          const int fc = CT ? mfma_row<T>(lane, r) : (lane & 15);
          int gi = tm * BM + wm * WM + i * 16 + fr;
          int gj = tn * BN + wn * WN + j * 16 + fc;
          if (gi < op.M && gj < op.N) {
            T v = op.alpha * acc[i][j][r];
            if (op.beta != T(0)) v += op.beta * cv[i][j][r];
            C[gi * op.sCi + gj * op.sCj] = v;
          }
        }
  } else {
    T* __restrict__ W = op.work + ((int64_t)blockIdx.z * op.nbatch + b) * op.M * op.N;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int fr = CT ? (lane & 15) : mfma_row<T>(lane, r);
          const int fc = CT ? mfma_row<T>(lane, r) : (lane & 15);
          int gi = tm * BM + wm * WM + i * 16 + fr;
          int gj = tn * BN + wn * WN + j * 16 + fc;
          if (gi < op.M && gj < op.N) W[gi + (int64_t)gj * op.M] = acc[i][j][r];
        }
  }
}

template <typename T>
__global__ void gemm_splitk_reduce(GemmOp<T> op) {
  int64_t mn = (int64_t)op.M * op.N;
  int64_t total = mn * op.nbatch;
  // 32-bit unsigned index arithmetic whenever it fits (a 64-bit division is
  // a long instruction sequence on the vector ALU)
  const bool narrow = total <= 0x7fffffff;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int b, gi, gj;
    int64_t e;
    if (narrow) {
      const unsigned tt = (unsigned)t, m = (unsigned)op.M, umn = (unsigned)mn;
      b = (int)(tt / umn);
      const unsigned ee = tt - (unsigned)b * umn;
      e = ee;
      gj = (int)(ee / m);
      gi = (int)(ee - (unsigned)gj * m);
    } else {
      b = (int)(t / mn);
      e = t % mn;
      gi = (int)(e % op.M);
      gj = (int)(e / op.M);
    }
    // load all partial sums first (latency-bound), then add in fixed order
    T part[16];
#pragma unroll
    for (int z = 0; z < 16; ++z)
      part[z] = (z < op.ksplit) ? op.work[((int64_t)z * op.nbatch + b) * mn + e] : T(0);
    T s = 0;
#pragma unroll
    for (int z = 0; z < 16; ++z) s += part[z];
    int b0 = 0, b1 = b;
    if (op.nb0 != 1) {
      b0 = (int)((unsigned)b % (unsigned)op.nb0);
      b1 = (int)((unsigned)b / (unsigned)op.nb0);
    }
    T* c = op.C + b0 * op.sCb0 + b1 * op.sCb1 + gi * op.sCi + gj * op.sCj;
    T v = op.alpha * s;
    if (op.beta != T(0)) v += op.beta * (*c);
    *c = v;
  }
}

// ---------------------------------------------------------------- complex
// Complex GEMM (complex128 on the FP64 matrix cores, complex64 on the FP32
// ones): real and imaginary planes are staged separately in LDS and each
// complex multiply-add is four real MFMAs.  conjA / conjB conjugate the
// operand on load.
template <typename R> struct CplxT;
template <> struct CplxT<double> {
  typedef hipDoubleComplex type;
  typedef double2 vec;
  __device__ __host__ static type make(double re, double im) { return make_hipDoubleComplex(re, im); }
};
template <> struct CplxT<float> {
  typedef hipFloatComplex type;
  typedef float2 vec;
  __device__ __host__ static type make(float re, float im) { return make_hipFloatComplex(re, im); }
};

template <typename R>
struct ZGemmOpT {
  typedef typename CplxT<R>::type C_t;
  int M, N, K;
  const C_t* A; int64_t sAi, sAk, sAb0, sAb1;
  const C_t* B; int64_t sBk, sBj, sBb0, sBb1;
  C_t* C;       int64_t sCi, sCj, sCb0, sCb1;
  int nb0, nbatch;
  int conjA, conjB;
  C_t alpha, beta;
  // split-K (ksplit > 1): slice z of K goes to work + (z nbatch + b) M N,
  // column-major and without alpha / beta; zgemm_splitk_reduce finishes C
  int ksplit;
  C_t* work;
};
typedef ZGemmOpT<double> ZGemmOp;   // complex128
typedef ZGemmOpT<float> CGemmOp;    // complex64

template <typename R, int BM, int BN, int BK = 16>
__global__ __launch_bounds__(256) void zgemm_mfma_kernel(ZGemmOpT<R> op) {
  static_assert(BM * BK % 256 == 0 && BN * BK % 256 == 0 && BK % 4 == 0, "zgemm tile");
  typedef typename CplxT<R>::type C_t;
  typedef typename CplxT<R>::vec V2;
  typedef typename Acc4<R>::type acc_t;
  constexpr int PAD = 2;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int LA = BM * BK / 256;
  constexpr int LB = BN * BK / 256;

  __shared__ R Ar[BK][BM + PAD], Ai[BK][BM + PAD];
  __shared__ R Br[BK][BN + PAD], Bi[BK][BN + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tilesM = (op.M + BM - 1) / BM, tilesN = (op.N + BN - 1) / BN;
  const int tile = xcd_remap(blockIdx.x, tilesM * tilesN);
  const int tm = tile / tilesN, tn = tile % tilesN;
  const int b = blockIdx.y, b0 = b % op.nb0, b1 = b / op.nb0;
  const C_t* __restrict__ A = op.A + b0 * op.sAb0 + b1 * op.sAb1;
  const C_t* __restrict__ B = op.B + b0 * op.sBb0 + b1 * op.sBb1;
  const bool a_ifast = (op.sAi == 1), b_jfast = (op.sBj == 1);
  const int ks = op.ksplit > 1 ? op.ksplit : 1;
  const int kchunk = ((op.K + ks - 1) / ks + BK - 1) / BK * BK;
  const int kbeg = blockIdx.z * kchunk, kend = min(op.K, kbeg + kchunk);
  const R sa = op.conjA ? R(-1) : R(1), sb = op.conjB ? R(-1) : R(1);

  V2 ra[LA], rb[LB];
  int ai[LA], ak[LA], bk[LB], bj[LB];
#pragma unroll
  for (int q = 0; q < LA; ++q) {
    int idx = tid + q * 256;
    if (a_ifast) { ai[q] = idx % BM; ak[q] = idx / BM; } else { ak[q] = idx % BK; ai[q] = idx / BK; }
  }
#pragma unroll
  for (int q = 0; q < LB; ++q) {
    int idx = tid + q * 256;
    if (b_jfast) { bj[q] = idx % BN; bk[q] = idx / BN; } else { bk[q] = idx % BK; bj[q] = idx / BK; }
  }
  // Every element is loaded from a clamped in-range address and zeroed by a
  // select afterwards: a load under a per-lane condition whose value is used
  // inside the condition compiles to a branch with a full wait, which would
  // serialise the tile's loads.
  auto load = [&](int k0) {
    C_t va[LA], vb[LB];
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int gi = min(tm * BM + ai[q], op.M - 1), gk = min(k0 + ak[q], kend - 1);
      va[q] = A[gi * op.sAi + gk * op.sAk];
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int gj = min(tn * BN + bj[q], op.N - 1), gk = min(k0 + bk[q], kend - 1);
      vb[q] = B[gk * op.sBk + gj * op.sBj];
    }
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const bool ok = tm * BM + ai[q] < op.M && k0 + ak[q] < kend;
      ra[q].x = ok ? va[q].x : R(0);
      ra[q].y = ok ? sa * va[q].y : R(0);
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const bool ok = tn * BN + bj[q] < op.N && k0 + bk[q] < kend;
      rb[q].x = ok ? vb[q].x : R(0);
      rb[q].y = ok ? sb * vb[q].y : R(0);
    }
  };

  acc_t cr[FM][FN], ci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) { cr[i][j] = acc_t{0, 0, 0, 0}; ci[i][j] = acc_t{0, 0, 0, 0}; }

  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < LA; ++q) { Ar[ak[q]][ai[q]] = ra[q].x; Ai[ak[q]][ai[q]] = ra[q].y; }
#pragma unroll
    for (int q = 0; q < LB; ++q) { Br[bk[q]][bj[q]] = rb[q].x; Bi[bk[q]][bj[q]] = rb[q].y; }
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      R ar[FM], ai_[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        int r = kk + (lane >> 4), c = wm * WM + i * 16 + (lane & 15);
        ar[i] = Ar[r][c]; ai_[i] = Ai[r][c]; nai[i] = -ai_[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int r = kk + (lane >> 4), c = wn * WN + j * 16 + (lane & 15);
        br[j] = Br[r][c]; bi[j] = Bi[r][c];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          cr[i][j] = mfma16(ar[i], br[j], cr[i][j]);
          cr[i][j] = mfma16(nai[i], bi[j], cr[i][j]);
          ci[i][j] = mfma16(ar[i], bi[j], ci[i][j]);
          ci[i][j] = mfma16(ai_[i], br[j], ci[i][j]);
        }
    }
  }
  if (ks > 1) {   // raw partial sums of this K slice
    C_t* __restrict__ Wz = op.work + ((int64_t)blockIdx.z * op.nbatch + b) * op.M * op.N;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = tm * BM + wm * WM + i * 16 + mfma_row<R>(lane, r);
          const int gj = tn * BN + wn * WN + j * 16 + (lane & 15);
          if (gi < op.M && gj < op.N) Wz[gi + (int64_t)gj * op.M] = CplxT<R>::make(cr[i][j][r], ci[i][j][r]);
        }
    return;
  }
  C_t* __restrict__ C = op.C + b0 * op.sCb0 + b1 * op.sCb1;
  const bool beta0 = (op.beta.x == R(0) && op.beta.y == R(0));
  // with beta != 0 the tile's C elements are all loaded (from clamped
  // in-range addresses) before any is used
  C_t cv[FM][FN][4];
  if (!beta0) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = min(tm * BM + wm * WM + i * 16 + mfma_row<R>(lane, r), op.M - 1);
          const int gj = min(tn * BN + wn * WN + j * 16 + (lane & 15), op.N - 1);
          cv[i][j][r] = C[gi * op.sCi + gj * op.sCj];
        }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gi = tm * BM + wm * WM + i * 16 + mfma_row<R>(lane, r);
        int gj = tn * BN + wn * WN + j * 16 + (lane & 15);
        if (gi < op.M && gj < op.N) {
          R xr = cr[i][j][r], xi = ci[i][j][r];
          R vr = op.alpha.x * xr - op.alpha.y * xi;
          R vi = op.alpha.x * xi + op.alpha.y * xr;
          if (!beta0) {
            const C_t o = cv[i][j][r];
            vr += op.beta.x * o.x - op.beta.y * o.y;
            vi += op.beta.x * o.y + op.beta.y * o.x;
          }
          C[gi * op.sCi + gj * op.sCj] = CplxT<R>::make(vr, vi);
        }
      }
}

// C = alpha (sum of the ksplit partial slices, in slice order) + beta C.
template <typename R>
__global__ void zgemm_splitk_reduce(ZGemmOpT<R> op) {
  typedef typename CplxT<R>::type C_t;
  const int64_t mn = (int64_t)op.M * op.N, total = mn * op.nbatch;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / mn);
    const int64_t e = t - (int64_t)b * mn;
    const int gj = (int)(e / op.M), gi = (int)(e - (int64_t)gj * op.M);
    // all slices (at most 8) and the old C loaded before the sums, added in
    // slice order
    const int b0 = b % op.nb0, b1 = b / op.nb0;
    C_t* c = op.C + b0 * op.sCb0 + b1 * op.sCb1 + gi * op.sCi + gj * op.sCj;
    const bool beta0 = op.beta.x == R(0) && op.beta.y == R(0);
    C_t part[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) part[z] = op.work[((int64_t)(z < op.ksplit ? z : 0) * op.nbatch + b) * mn + e];
    const C_t o = beta0 ? CplxT<R>::make(R(0), R(0)) : *c;
    R xr = R(0), xi = R(0);
#pragma unroll
    for (int z = 0; z < 8; ++z)
      if (z < op.ksplit) {
        xr += part[z].x;
        xi += part[z].y;
      }
    R vr = op.alpha.x * xr - op.alpha.y * xi;
    R vi = op.alpha.x * xi + op.alpha.y * xr;
    if (!beta0) {
      vr += op.beta.x * o.x - op.beta.y * o.y;
      vi += op.beta.x * o.y + op.beta.y * o.x;
    }
    *c = CplxT<R>::make(vr, vi);
  }
}

// ---------------------------------------------------------------- launchers
// Implemented in gemm.hip.  The tile shape (64x64 or 32x32, K chunk 16 or
// 32) and the split-K factor of each GEMM are chosen by a runtime autotuner
// keyed on (type, M, N, K, batch, operand layout): the first launch of a new
// shape times the candidates with events on the stream and caches the
// fastest.  Small-panel products (K <= 64, N <= 32, B shared by all batches)
// go to a dedicated vector kernel instead of the matrix cores.
template <typename T>
void gemm_launch(GemmOp<T> op, hipStream_t st);
// Elements of workspace gemm_launch needs for op's split-K slices (0 when
// the product runs without them); a function of the shape alone.
template <typename T>
size_t gemm_workspace(const GemmOp<T>& op);
// gemm_launch without any allocation and without tuning runs: split-K
// slices go to ws (ws_cap elements; an error if too small), the tiles are
// the ones tuned earlier for this shape if any, else the fixed heuristic.
// The split-K factor, and so every bit of C, is the same as gemm_launch's.
template <typename T>
void gemm_launch_ws(GemmOp<T> op, T* ws, size_t ws_cap, hipStream_t st);
void zgemm_launch(ZGemmOp op, hipStream_t st);
void cgemm_launch(CGemmOp op, hipStream_t st);

// While on, gemm_launch on this thread runs no tuning timings (for code
// that records a stream capture); untuned shapes use the heuristic tiles.
void gemm_no_tuning(bool on);

// Convenience constructor for a column-major GEMM with optional transposes:
// C (M x N, leading dim ldc) = alpha * op(A) op(B) + beta * C.
template <typename T>
GemmOp<T> colmajor(char ta, char tb, int M, int N, int K, T alpha, const T* A, int64_t lda,
                   const T* B, int64_t ldb, T beta, T* C, int64_t ldc) {
  GemmOp<T> op{};
  op.M = M; op.N = N; op.K = K;
  op.A = A; op.B = B; op.C = C;
  if (ta == 'N' || ta == 'n') { op.sAi = 1; op.sAk = lda; } else { op.sAi = lda; op.sAk = 1; }
  if (tb == 'N' || tb == 'n') { op.sBk = 1; op.sBj = ldb; } else { op.sBk = ldb; op.sBj = 1; }
  op.sCi = 1; op.sCj = ldc;
  op.nb0 = 1; op.nbatch = 1;
  op.alpha = alpha; op.beta = beta;
  op.ksplit = 1; op.work = nullptr;
  return op;
}

}  // namespace ttf

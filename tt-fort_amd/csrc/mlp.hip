// Learned block preconditioner (see mlp.h): feature extraction, the network
// (bf16 weights, fp32 activations and accumulation on the fp32 matrix
// cores), its backward pass and Adam training, inference inside the local
// GMRES, and the collection of training pairs from block-Jacobi solves.
//
// Network: y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 + jac(x) on
// columns x of the feature matrix X (kMlpIn x M, column-major, one column
// per block), where jac(x) = residual features / diagonal features is the
// point-Jacobi solution of the block that the layers correct.
// Features of block b = (b1, b2) of the local vector v (rx1, n, rx2):
//   [diag(B_b), superdiag(B_b), subdiag(B_b)] / s_b and r_b / |r_b|,
// with B_b(i,j) = sum_{a1,a2} Phi1(b1,b1,a1) A(a1,i,j,a2) Phi2(b2,a2,b2)
// (the block of djac_gen 'c', tt_linalg.f90:412-416), r_b = v(b1, :, b2)
// and s_b = mean |diag(B_b)|; the block correction is y_b |r_b| / s_b.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ttf {

namespace {

struct MlpModel {
  bool ready = false;
  float* master = nullptr;    // fp32 parameters (what Adam updates)
  uint16_t* wbf = nullptr;    // bf16 copy read by the forward kernels
  float* wr = nullptr;        // fp32 values of the bf16 weights (backward GEMMs)
  float* am = nullptr;        // Adam moments
  float* av = nullptr;
  long step = 0;
};
MlpModel g_model;

struct MlpSink {
  bool active = false;
  int64_t cap = 0, count = 0;
  int per_apply = 0;
  uint64_t counter = 0;
  float* X = nullptr;   // kMlpIn x cap
  float* T = nullptr;   // kMlpOut x cap
  float* inv = nullptr; // per-sample s_b / |r_b| staging
  int64_t inv_cap = 0;
};
MlpSink g_sink;

void ensure_model() {
  if (g_model.master) return;
  const size_t nb = sizeof(float) * (size_t)kMlpParams;
  TTF_HIP(hipMalloc(&g_model.master, nb));
  TTF_HIP(hipMalloc(&g_model.wbf, sizeof(uint16_t) * (size_t)kMlpParams));
  TTF_HIP(hipMalloc(&g_model.wr, nb));
  TTF_HIP(hipMalloc(&g_model.am, nb));
  TTF_HIP(hipMalloc(&g_model.av, nb));
}

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);   // inf / nan kept
  u += 0x7fffu + ((u >> 16) & 1u);                                      // nearest, ties to even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// master -> bf16 copy and its fp32 value
__global__ void round_params_kernel(int64_t np, const float* __restrict__ master, uint16_t* __restrict__ wbf,
                                    float* __restrict__ wr) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np; e += (int64_t)gridDim.x * blockDim.x) {
    const uint16_t h = to_bf16(master[e]);
    wbf[e] = h;
    wr[e] = from_bf16(h);
  }
}

// One wave per output column c; block b = (b_start + c * b_stride) mod nb.
// Writes X(:, c), scale(c) = |r_b| / s_b and, if inv is given,
// inv(c) = s_b / |r_b| (0 for a zero slice).
__global__ __launch_bounds__(256) void features_kernel(int rx1, int n, int rx2, int ra1, int ra2,
                                                       const double* __restrict__ phi1, const double* __restrict__ A,
                                                       const double* __restrict__ phi2, const double* __restrict__ v,
                                                       int ncols, int64_t b_start, int64_t b_stride,
                                                       float* __restrict__ X, float* __restrict__ scale,
                                                       float* __restrict__ inv) {
  const int lane = threadIdx.x & 63;
  const int64_t col = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (col >= ncols) return;
  const int64_t nb = (int64_t)rx1 * rx2;
  const int64_t b = (b_start + col * b_stride) % nb;
  const int b1 = (int)(b % rx1), b2 = (int)(b / rx1);
  const int i = lane;   // n == 64
  double dg = 0.0, up = 0.0, lo = 0.0;
  for (int a1 = 0; a1 < ra1; ++a1) {
    const double d1 = phi1[b1 + (int64_t)rx1 * (b1 + (int64_t)rx1 * a1)];
    for (int a2 = 0; a2 < ra2; ++a2) {
      const double c = d1 * phi2[b2 + (int64_t)rx2 * (a2 + (int64_t)ra2 * b2)];
      const int64_t base = a1 + (int64_t)ra1 * n * n * a2;
      dg += c * A[base + (int64_t)ra1 * (i + (int64_t)n * i)];
      if (i < n - 1) {
        up += c * A[base + (int64_t)ra1 * (i + (int64_t)n * (i + 1))];
        lo += c * A[base + (int64_t)ra1 * ((i + 1) + (int64_t)n * i)];
      }
    }
  }
  const double r = v[b1 + (int64_t)rx1 * (i + (int64_t)n * b2)];
  const double s = wave_sum(fabs(dg)) / n;
  const double rn = sqrt(wave_sum(r * r));
  const double sinv = s > 0.0 ? 1.0 / s : 0.0;
  const double rinv = rn > 0.0 ? 1.0 / rn : 0.0;
  float* x = X + (int64_t)kMlpIn * col;
  x[i] = (float)(dg * sinv);
  x[n + i] = (float)(up * sinv);
  x[2 * n + i] = (float)(lo * sinv);
  x[3 * n + i] = (float)(r * rinv);
  if (lane == 0) {
    scale[col] = (float)(rn * sinv);
    if (inv) inv[col] = (float)(rn > 0.0 ? s * rinv : 0.0);
  }
}

// Y (N x M, ld N) = act(W^T X + b): W bf16 (K x N, ld K), X fp32 (K x M,
// ld K).  Work-group tile 64 (N) x 64 (M), 4 waves of 32 x 32, K in chunks
// of 16 through LDS; v_mfma_f32_16x16x4f32 with the weights widened on load.
// With SKIP (the output layer), X0 are the network's input features and
// Y(n, m) += X0(3 kMlpBlock + n, m) / X0(n, m): the point-Jacobi solution of
// the block, which the layers correct.
template <bool RELU, bool SKIP>
__global__ __launch_bounds__(256) void layer_kernel(int M, int K, int N, const float* __restrict__ X,
                                                    const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                                    float* __restrict__ Y, const float* __restrict__ X0) {
  constexpr int BK = 16, PAD = 4;
  __shared__ float Ws[BK][64 + PAD];
  __shared__ float Xs[BK][64 + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wn = wave & 1, wm = wave >> 1;
  const int n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
  f4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q;
      const int kk = idx & (BK - 1), cc = idx >> 4;
      const int gn = n0 + cc, gm = m0 + cc;
      Ws[kk][cc] = gn < N ? from_bf16(W[(k0 + kk) + (int64_t)K * gn]) : 0.f;
      Xs[kk][cc] = gm < M ? X[(k0 + kk) + (int64_t)K * gm] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = Ws[kk + (lane >> 4)][wn * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Xs[kk + (lane >> 4)][wm * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  // the epilogue's operands for the 16 outputs (bias, and with SKIP the
  // block diagonal and the residual), loaded from clamped in-range indices
  // before any is used: loaded under each output's own condition they
  // become a branch with a full wait apiece
  float eb[16], ed[16], ex[16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = i * 8 + j * 4 + r;
        const int gn = min(n0 + wn * 32 + i * 16 + mfma_row<float>(lane, r), N - 1);
        const int gm = min(m0 + wm * 32 + j * 16 + (lane & 15), M - 1);
        eb[q] = bias[gn];
        if (SKIP) {
          ed[q] = X0[gn + (int64_t)kMlpIn * gm];
          ex[q] = X0[3 * kMlpBlock + gn + (int64_t)kMlpIn * gm];
        }
      }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = i * 8 + j * 4 + r;
        const int gn = n0 + wn * 32 + i * 16 + mfma_row<float>(lane, r);
        const int gm = m0 + wm * 32 + j * 16 + (lane & 15);
        if (gn < N && gm < M) {
          float val = acc[i][j][r] + eb[q];
          if (RELU) val = fmaxf(val, 0.f);
          if (SKIP) val += ed[q] != 0.f ? ex[q] / ed[q] : 0.f;
          Y[gn + (int64_t)N * gm] = val;
        }
      }
}

// The inference form of layer_kernel on gfx950's bf16 matrix cores: the
// same Y = act(W^T X + b) (+ the point-Jacobi term with SKIP), with the
// activations X rounded to bf16 (round to nearest even) and the products
// accumulated in fp32 by v_mfma_f32_16x16x32_bf16.  K = 256 for every layer
// of the network.  Lane l feeds row (or column) l % 16 and the eight
// consecutive k = 8 (l / 16) .. +7 of each 32-deep step to both operands,
// so the A and B elements it supplies meet at the same k; fragments are
// read straight from global memory (weights 16 bytes, activations 32 bytes
// per lane), all eight k-steps' loads issued before the products.
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
constexpr int kBfK = 256;
static_assert(kMlpIn == kBfK && kMlpHid == kBfK, "the bf16 layers assume K = 256");

template <bool RELU, bool SKIP>
__global__ __launch_bounds__(256) void layer_bf16_kernel(int M, int N, const float* __restrict__ X,
                                                         const uint16_t* __restrict__ W,
                                                         const float* __restrict__ bias, float* __restrict__ Y,
                                                         const float* __restrict__ X0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wn = wave & 1, wm = wave >> 1;
  const int n0 = blockIdx.x * 64 + wn * 32, m0 = blockIdx.y * 64 + wm * 32;
  const int g = lane >> 4, c = lane & 15;
  f4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  const uint16_t* wa[2];
  const float* xb[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int n = min(n0 + 16 * i + c, N - 1);   // rows past N are computed and not stored
    wa[i] = W + 8 * g + (int64_t)kBfK * n;
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int m = min(m0 + 16 * j + c, M - 1);
    xb[j] = X + 8 * g + (int64_t)kBfK * m;
  }
  uint4 wr[kBfK / 32][2];
  float4 xr[kBfK / 32][2][2];
#pragma unroll
  for (int s = 0; s < kBfK / 32; ++s) {
#pragma unroll
    for (int i = 0; i < 2; ++i) wr[s][i] = *reinterpret_cast<const uint4*>(wa[i] + 32 * s);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      xr[s][j][0] = *reinterpret_cast<const float4*>(xb[j] + 32 * s);
      xr[s][j][1] = *reinterpret_cast<const float4*>(xb[j] + 32 * s + 4);
    }
  }
#pragma unroll
  for (int s = 0; s < kBfK / 32; ++s) {
    bf8_t a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = __builtin_bit_cast(bf8_t, wr[s][i]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float4 lo = xr[s][j][0], hi = xr[s][j][1];
      b[j] = bf8_t{(__bf16)lo.x, (__bf16)lo.y, (__bf16)lo.z, (__bf16)lo.w,
                   (__bf16)hi.x, (__bf16)hi.y, (__bf16)hi.z, (__bf16)hi.w};
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  // the epilogue's operands for the 16 outputs (bias, and with SKIP the
  // block diagonal and the residual), loaded from clamped in-range indices
  // before any is used: loaded under each output's own condition they
  // become a branch with a full wait apiece
  float eb[16], ed[16], ex[16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = i * 8 + j * 4 + r;
        const int gn = min(n0 + i * 16 + mfma_row<float>(lane, r), N - 1);
        const int gm = min(m0 + j * 16 + c, M - 1);
        eb[q] = bias[gn];
        if (SKIP) {
          ed[q] = X0[gn + (int64_t)kMlpIn * gm];
          ex[q] = X0[3 * kMlpBlock + gn + (int64_t)kMlpIn * gm];
        }
      }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = i * 8 + j * 4 + r;
        const int gn = n0 + i * 16 + mfma_row<float>(lane, r);
        const int gm = m0 + j * 16 + c;
        if (gn < N && gm < M) {
          float val = acc[i][j][r] + eb[q];
          if (RELU) val = fmaxf(val, 0.f);
          if (SKIP) val += ed[q] != 0.f ? ex[q] / ed[q] : 0.f;
          Y[gn + (int64_t)N * gm] = val;
        }
      }
}

// z(b1, i, b2) = Y(i, b) * scale(b), one wave per block
__global__ __launch_bounds__(256) void scatter_kernel(int rx1, int n, int rx2, const float* __restrict__ Y,
                                                      const float* __restrict__ scale, double* __restrict__ z) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (b >= (int64_t)rx1 * rx2) return;
  const int b1 = (int)(b % rx1), b2 = (int)(b / rx1);
  z[b1 + (int64_t)rx1 * (lane + (int64_t)n * b2)] = (double)Y[lane + (int64_t)kMlpOut * b] * (double)scale[b];
}

// T(i, c) = z(b1, i, b2) * inv(c) for the sampled blocks
__global__ __launch_bounds__(256) void targets_kernel(int rx1, int n, int rx2, const double* __restrict__ z,
                                                      int ncols, int64_t b_start, int64_t b_stride,
                                                      const float* __restrict__ inv, float* __restrict__ T) {
  const int lane = threadIdx.x & 63;
  const int64_t col = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (col >= ncols) return;
  const int64_t nb = (int64_t)rx1 * rx2;
  const int64_t b = (b_start + col * b_stride) % nb;
  const int b1 = (int)(b % rx1), b2 = (int)(b / rx1);
  T[lane + (int64_t)kMlpOut * col] = (float)(z[b1 + (int64_t)rx1 * (lane + (int64_t)n * b2)] * (double)inv[col]);
}

// dY = 2 (Y - T) / (M * Out) and per-work-group partial sums of (Y - T)^2
__global__ __launch_bounds__(256) void loss_grad_kernel(int64_t count, int M, const float* __restrict__ Y,
                                                        const float* __restrict__ T, float* __restrict__ dY,
                                                        double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  const float f = 2.0f / ((float)M * (float)kMlpOut);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x) {
    const float d = Y[e] - T[e];
    dY[e] = f * d;
    s += (double)d * d;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_partials_kernel(int np, const double* __restrict__ part, double scale, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += part[i];
    *out = s * scale;
  }
}

// dH *= (H > 0) (ReLU derivative from the activation itself)
__global__ void relu_mask_kernel(int64_t count, const float* __restrict__ H, float* __restrict__ dH) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    if (!(H[e] > 0.f)) dH[e] = 0.f;
}

// out(r) = sum_c D(r, c), D (rows x cols, ld rows); one thread per row
__global__ void rowsum_kernel(int rows, int cols, const float* __restrict__ D, float* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float s = 0.f;
  for (int c = 0; c < cols; ++c) s += D[r + (int64_t)rows * c];
  out[r] = s;
}

__global__ void gather_cols_kernel(int rows, int ncols, const int* __restrict__ idx, const float* __restrict__ src,
                                   float* __restrict__ dst) {
  const int64_t total = (int64_t)rows * ncols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % rows);
    const int64_t c = e / rows;
    dst[e] = src[r + (int64_t)rows * idx[c]];
  }
}

__global__ void adam_kernel(int64_t np, float* __restrict__ master, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1, float c2,
                            uint16_t* __restrict__ wbf, float* __restrict__ wr) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np; e += (int64_t)gridDim.x * blockDim.x) {
    const float ge = g[e];
    const float me = beta1 * m[e] + (1.f - beta1) * ge;
    const float ve = beta2 * v[e] + (1.f - beta2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    const float th = master[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
    master[e] = th;
    const uint16_t h = to_bf16(th);
    wbf[e] = h;
    wr[e] = from_bf16(h);
  }
}

int grid_of(int64_t n, int per = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, 4096)); }

void features(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A, const double* phi2,
              const double* v, int ncols, int64_t b_start, int64_t b_stride, float* X, float* scale, float* inv,
              hipStream_t st) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)((ncols + 3) / 4)), dim3(256), 0, st, rx1, n, rx2, ra1, ra2, phi1,
                     A, phi2, v, ncols, b_start, b_stride, X, scale, inv);
  TTF_HIP(hipGetLastError());
}

template <bool RELU, bool SKIP>
void layer(int M, int K, int N, const float* X, const uint16_t* W, const float* b, float* Y, const float* X0,
           hipStream_t st) {
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  hipLaunchKernelGGL((layer_kernel<RELU, SKIP>), grid, dim3(256), 0, st, M, K, N, X, W, b, Y, X0);
  TTF_HIP(hipGetLastError());
}

// H1, H2 (Hid x M), Y (Out x M)
void forward(int M, const float* X, float* H1, float* H2, float* Y, hipStream_t st) {
  const uint16_t* w = g_model.wbf;
  const float* p = g_model.master;
  layer<true, false>(M, kMlpIn, kMlpHid, X, w + kMlpOffW1, p + kMlpOffB1, H1, nullptr, st);
  layer<true, false>(M, kMlpHid, kMlpHid, H1, w + kMlpOffW2, p + kMlpOffB2, H2, nullptr, st);
  layer<false, true>(M, kMlpHid, kMlpOut, H2, w + kMlpOffW3, p + kMlpOffB3, Y, X, st);
}

template <bool RELU, bool SKIP>
void layer_bf16(int M, int N, const float* X, const uint16_t* W, const float* b, float* Y, const float* X0,
                hipStream_t st) {
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  hipLaunchKernelGGL((layer_bf16_kernel<RELU, SKIP>), grid, dim3(256), 0, st, M, N, X, W, b, Y, X0);
  TTF_HIP(hipGetLastError());
}

// Inference: the same network with bf16 activations on the bf16 matrix
// cores (training keeps the fp32 forward above, whose activations the
// backward pass differentiates)
void forward_infer(int M, const float* X, float* H1, float* H2, float* Y, hipStream_t st) {
  const uint16_t* w = g_model.wbf;
  const float* p = g_model.master;
  layer_bf16<true, false>(M, kMlpHid, X, w + kMlpOffW1, p + kMlpOffB1, H1, nullptr, st);
  layer_bf16<true, false>(M, kMlpHid, H1, w + kMlpOffW2, p + kMlpOffB2, H2, nullptr, st);
  layer_bf16<false, true>(M, kMlpOut, H2, w + kMlpOffW3, p + kMlpOffB3, Y, X, st);
}

// g <- gradient of the batch loss; dY must already hold dL/dY; dH2, dH1 work
void backward(int M, const float* X, const float* H1, const float* H2, const float* dY, float* dH2, float* dH1,
              float* g, hipStream_t st) {
  const float* wr = g_model.wr;
  gemm_launch(colmajor<float>('N', 'T', kMlpHid, kMlpOut, M, 1.f, H2, kMlpHid, dY, kMlpOut, 0.f, g + kMlpOffW3,
                              kMlpHid), st);
  hipLaunchKernelGGL(rowsum_kernel, dim3(1), dim3(256), 0, st, kMlpOut, M, dY, g + kMlpOffB3);
  gemm_launch(colmajor<float>('N', 'N', kMlpHid, M, kMlpOut, 1.f, wr + kMlpOffW3, kMlpHid, dY, kMlpOut, 0.f, dH2,
                              kMlpHid), st);
  hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of((int64_t)kMlpHid * M)), dim3(256), 0, st, (int64_t)kMlpHid * M, H2,
                     dH2);
  gemm_launch(colmajor<float>('N', 'T', kMlpHid, kMlpHid, M, 1.f, H1, kMlpHid, dH2, kMlpHid, 0.f, g + kMlpOffW2,
                              kMlpHid), st);
  hipLaunchKernelGGL(rowsum_kernel, dim3(1), dim3(256), 0, st, kMlpHid, M, dH2, g + kMlpOffB2);
  gemm_launch(colmajor<float>('N', 'N', kMlpHid, M, kMlpHid, 1.f, wr + kMlpOffW2, kMlpHid, dH2, kMlpHid, 0.f, dH1,
                              kMlpHid), st);
  hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of((int64_t)kMlpHid * M)), dim3(256), 0, st, (int64_t)kMlpHid * M, H1,
                     dH1);
  gemm_launch(colmajor<float>('N', 'T', kMlpIn, kMlpHid, M, 1.f, X, kMlpIn, dH1, kMlpHid, 0.f, g + kMlpOffW1, kMlpIn),
              st);
  hipLaunchKernelGGL(rowsum_kernel, dim3(1), dim3(256), 0, st, kMlpHid, M, dH1, g + kMlpOffB1);
  TTF_HIP(hipGetLastError());
}

// Workspace for M columns: X | H1 | H2 | Y | dY | dH2 | dH1 | scale (floats)
struct MlpWork {
  float *X, *H1, *H2, *Y, *dY, *dH2, *dH1, *scale;
};
MlpWork work_for(int M, bool train) {
  const size_t per = (size_t)kMlpIn + 2 * kMlpHid + kMlpOut + (train ? (size_t)kMlpOut + 2 * kMlpHid : 0) + 1;
  float* base = (float*)arena(kArenaMlp, sizeof(float) * per * (size_t)M + 256);
  MlpWork w{};
  w.X = base;
  w.H1 = w.X + (size_t)kMlpIn * M;
  w.H2 = w.H1 + (size_t)kMlpHid * M;
  w.Y = w.H2 + (size_t)kMlpHid * M;
  float* p = w.Y + (size_t)kMlpOut * M;
  if (train) {
    w.dY = p;
    w.dH2 = w.dY + (size_t)kMlpOut * M;
    w.dH1 = w.dH2 + (size_t)kMlpHid * M;
    p = w.dH1 + (size_t)kMlpHid * M;
  }
  w.scale = p;
  return w;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

void set_params_device(const float* host) {
  ensure_model();
  hipStream_t st = stream();
  TTF_HIP(hipMemcpyAsync(g_model.master, host, sizeof(float) * (size_t)kMlpParams, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(round_params_kernel, dim3(grid_of(kMlpParams)), dim3(256), 0, st, (int64_t)kMlpParams,
                     g_model.master, g_model.wbf, g_model.wr);
  TTF_HIP(hipGetLastError());
  TTF_HIP(hipMemsetAsync(g_model.am, 0, sizeof(float) * (size_t)kMlpParams, st));
  TTF_HIP(hipMemsetAsync(g_model.av, 0, sizeof(float) * (size_t)kMlpParams, st));
  TTF_HIP(hipStreamSynchronize(st));
  g_model.step = 0;
  g_model.ready = true;
}

double batch_loss_grad(int M, const float* X, const float* T, MlpWork& w, float* g, double* part, double* out,
                       hipStream_t st) {
  forward(M, X, w.H1, w.H2, w.Y, st);
  const int64_t cnt = (int64_t)kMlpOut * M;
  const int ng = std::min(grid_of(cnt), 256);
  hipLaunchKernelGGL(loss_grad_kernel, dim3(ng), dim3(256), 0, st, cnt, M, w.Y, T, w.dY, part);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, st, ng, part, 1.0 / (double)cnt, out);
  backward(M, X, w.H1, w.H2, w.dY, w.dH2, w.dH1, g, st);
  return 0.0;
}

}  // namespace

bool mlp_ready() { return g_model.ready; }

void mlp_precond(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A, const double* phi2,
                 const double* v, double* z, hipStream_t st) {
  TTF_REQUIRE(g_model.ready, "learned preconditioner: no parameters set (ttf_mlp_init / ttf_mlp_set_params)");
  TTF_REQUIRE(n == kMlpBlock, "learned preconditioner: mode size %d, the network is built for %d", n, kMlpBlock);
  const int64_t nb64 = (int64_t)rx1 * rx2;
  TTF_REQUIRE(nb64 <= (1 << 24), "learned preconditioner: %lld blocks", (long long)nb64);
  const int nb = (int)nb64;
  MlpWork w = work_for(nb, false);
  features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, nb, 0, 1, w.X, w.scale, nullptr, st);
  forward_infer(nb, w.X, w.H1, w.H2, w.Y, st);
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, st, rx1, n, rx2, w.Y, w.scale, z);
  TTF_HIP(hipGetLastError());
}

bool mlp_collecting() { return g_sink.active; }

void mlp_collect_pair(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                      const double* phi2, const double* v, const double* z, hipStream_t st) {
  if (!g_sink.active || n != kMlpBlock || g_sink.count >= g_sink.cap) return;
  const int64_t nb = (int64_t)rx1 * rx2;
  const int64_t take = std::min<int64_t>({(int64_t)g_sink.per_apply, nb, g_sink.cap - g_sink.count});
  if (take <= 0) return;
  const int64_t stride = std::max<int64_t>(1, nb / take);
  const int64_t start = (int64_t)(splitmix(g_sink.counter) % (uint64_t)nb);
  if (g_sink.inv_cap < take) {
    if (g_sink.inv) TTF_HIP(hipFree(g_sink.inv));
    TTF_HIP(hipMalloc(&g_sink.inv, sizeof(float) * 2 * (size_t)take));
    g_sink.inv_cap = take;
  }
  float* X = g_sink.X + (size_t)kMlpIn * g_sink.count;
  float* T = g_sink.T + (size_t)kMlpOut * g_sink.count;
  features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, (int)take, start, stride, X, g_sink.inv + take, g_sink.inv, st);
  hipLaunchKernelGGL(targets_kernel, dim3((unsigned)((take + 3) / 4)), dim3(256), 0, st, rx1, n, rx2, z, (int)take,
                     start, stride, g_sink.inv, T);
  TTF_HIP(hipGetLastError());
  g_sink.count += take;
}

}  // namespace ttf

using namespace ttf;

extern "C" {

long ttf_mlp_param_count(void) { return (long)kMlpParams; }

int ttf_mlp_init(unsigned seed) {
  return guarded([&] {
    std::vector<float> p((size_t)kMlpParams, 0.f);
    uint64_t s = 0x1234abcdull ^ ((uint64_t)seed << 17);
    auto fill = [&](int64_t off, int fan_in, int64_t cnt) {
      const double lim = 1.0 / std::sqrt((double)fan_in);
      for (int64_t e = 0; e < cnt; ++e) {
        const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
        p[(size_t)(off + e)] = (float)((2.0 * u - 1.0) * lim);
      }
    };
    fill(kMlpOffW1, kMlpIn, (int64_t)kMlpIn * kMlpHid);
    fill(kMlpOffW2, kMlpHid, (int64_t)kMlpHid * kMlpHid);
    // W3 and all biases start at zero: the untrained network is the
    // point-Jacobi solve of each block
    set_params_device(p.data());
  });
}

int ttf_mlp_set_params(const float* params) {
  return guarded([&] {
    TTF_REQUIRE(params != nullptr, "ttf_mlp_set_params: null parameters");
    set_params_device(params);
  });
}

int ttf_mlp_get_params(float* params) {
  return guarded([&] {
    TTF_REQUIRE(g_model.ready, "ttf_mlp_get_params: no parameters set");
    hipStream_t st = stream();
    read_back(params, g_model.master, sizeof(float) * (size_t)kMlpParams, st);
  });
}

int ttf_mlp_forward(int M, const float* X, float* Y) {
  return guarded([&] {
    TTF_REQUIRE(g_model.ready, "ttf_mlp_forward: no parameters set");
    TTF_REQUIRE(M >= 0, "ttf_mlp_forward: M=%d", M);
    if (M == 0) return;
    MlpWork w = work_for(M, false);
    forward_infer(M, X, w.H1, w.H2, Y, stream());
  });
}

int ttf_mlp_loss_grad(int M, const float* X, const float* T, double* loss, float* grad) {
  return guarded([&] {
    TTF_REQUIRE(g_model.ready, "ttf_mlp_loss_grad: no parameters set");
    TTF_REQUIRE(M > 0, "ttf_mlp_loss_grad: M=%d", M);
    hipStream_t st = stream();
    MlpWork w = work_for(M, true);
    float* g = (float*)arena(kArenaMlp2, sizeof(float) * (size_t)kMlpParams + sizeof(double) * 512);
    double* part = (double*)(g + kMlpParams + (kMlpParams & 1));
    double* out = part + 256;
    batch_loss_grad(M, X, T, w, g, part, out, st);
    TTF_HIP(hipMemcpyAsync(loss, out, sizeof(double), hipMemcpyDeviceToHost, st));
    if (grad) read_back(grad, g, sizeof(float) * (size_t)kMlpParams, st);
  });
}

int ttf_mlp_train(long nsamples, const float* X, const float* T, int epochs, int batch, float lr, unsigned seed,
                  double* loss_hist, int hist_cap, int* hist_len) {
  return guarded([&] {
    TTF_REQUIRE(g_model.ready, "ttf_mlp_train: no parameters set (ttf_mlp_init first)");
    TTF_REQUIRE(nsamples > 0 && epochs >= 0 && batch > 0 && lr > 0.f, "ttf_mlp_train: bad arguments");
    TTF_REQUIRE(nsamples <= (1L << 30), "ttf_mlp_train: %ld samples", nsamples);
    hipStream_t st = stream();
    const int B = (int)std::min<long>(batch, nsamples);
    const long nbat = nsamples / B;
    MlpWork w = work_for(B, true);
    float* g = (float*)arena(kArenaMlp2, sizeof(float) * (size_t)kMlpParams + sizeof(double) * (512 + nbat + 1) +
                                             sizeof(int) * (size_t)nsamples + 64);
    double* part = (double*)(g + kMlpParams + (kMlpParams & 1));
    double* losses = part + 256;
    int* idx = (int*)(losses + nbat + 1);
    float* Tb = (float*)arena(kArenaMlp3, sizeof(float) * (size_t)kMlpOut * B);   // batch targets
    std::vector<int> perm((size_t)nsamples);
    for (long i = 0; i < nsamples; ++i) perm[(size_t)i] = (int)i;
    uint64_t s = 0x5eedull ^ ((uint64_t)seed << 21);
    std::vector<double> hl((size_t)nbat);
    if (hist_len) *hist_len = 0;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
      for (long i = nsamples - 1; i > 0; --i) std::swap(perm[(size_t)i], perm[(size_t)(splitmix(s) % (uint64_t)(i + 1))]);
      TTF_HIP(hipMemcpyAsync(idx, perm.data(), sizeof(int) * (size_t)nsamples, hipMemcpyHostToDevice, st));
      for (long bi = 0; bi < nbat; ++bi) {
        const int* ib = idx + bi * B;
        hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_of((int64_t)kMlpIn * B)), dim3(256), 0, st, kMlpIn, B, ib, X, w.X);
        hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_of((int64_t)kMlpOut * B)), dim3(256), 0, st, kMlpOut, B, ib, T,
                           Tb);
        batch_loss_grad(B, w.X, Tb, w, g, part, losses + bi, st);
        ++g_model.step;
        const float c1 = 1.f - std::pow(beta1, (float)g_model.step);
        const float c2 = 1.f - std::pow(beta2, (float)g_model.step);
        hipLaunchKernelGGL(adam_kernel, dim3(grid_of(kMlpParams)), dim3(256), 0, st, (int64_t)kMlpParams, g_model.master,
                           g, g_model.am, g_model.av, lr, beta1, beta2, eps, c1, c2, g_model.wbf, g_model.wr);
        TTF_HIP(hipGetLastError());
      }
      read_back(hl.data(), losses, sizeof(double) * (size_t)nbat, st);   // idx is rewritten by the next epoch
      if (loss_hist && hist_len && *hist_len < hist_cap) {
        double m = 0.0;
        for (double v : hl) m += v;
        loss_hist[(*hist_len)++] = m / (double)std::max<long>(nbat, 1);
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
  });
}

int ttf_mlp_features(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                     const double* phi2, const double* v, float* X, float* scale) {
  return guarded([&] {
    TTF_REQUIRE(n == kMlpBlock, "ttf_mlp_features: mode size %d, features are built for %d", n, kMlpBlock);
    TTF_REQUIRE(rx1 > 0 && rx2 > 0 && ra1 > 0 && ra2 > 0, "ttf_mlp_features: invalid sizes");
    features(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, rx1 * rx2, 0, 1, X, scale, nullptr, stream());
  });
}

int ttf_mlp_precond(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                    const double* phi2, const double* v, double* z) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && rx2 > 0 && ra1 > 0 && ra2 > 0, "ttf_mlp_precond: invalid sizes");
    mlp_precond(rx1, n, rx2, ra1, ra2, phi1, A, phi2, v, z, stream());
  });
}

int ttf_mlp_collect_begin(long capacity, int per_apply) {
  return guarded([&] {
    TTF_REQUIRE(capacity > 0 && per_apply > 0, "ttf_mlp_collect_begin: bad arguments");
    if (g_sink.X) TTF_HIP(hipFree(g_sink.X));
    if (g_sink.T) TTF_HIP(hipFree(g_sink.T));
    g_sink.X = g_sink.T = nullptr;
    TTF_HIP(hipMalloc(&g_sink.X, sizeof(float) * (size_t)kMlpIn * capacity));
    TTF_HIP(hipMalloc(&g_sink.T, sizeof(float) * (size_t)kMlpOut * capacity));
    g_sink.cap = capacity;
    g_sink.count = 0;
    g_sink.per_apply = per_apply;
    g_sink.counter = 0x9e37ull;
    g_sink.active = true;
  });
}

int ttf_mlp_collect_end(long* count) {
  return guarded([&] {
    g_sink.active = false;
    TTF_HIP(hipStreamSynchronize(stream()));
    if (count) *count = (long)g_sink.count;
  });
}

int ttf_mlp_collected(float** X, float** T, long* count) {
  return guarded([&] {
    TTF_REQUIRE(X && T && count, "ttf_mlp_collected: null arguments");
    *X = g_sink.X;
    *T = g_sink.T;
    *count = (long)g_sink.count;
  });
}

}  // extern "C"

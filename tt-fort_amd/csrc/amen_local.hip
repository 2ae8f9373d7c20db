// Local building blocks of the AMEn linear solver on the device:
//   * block-Jacobi preconditioner blocks and their application
//     (tt_linalg.f90:363-533, djac_gen / djac_apply);
//   * right-preconditioned restarted GMRES for the local system
//     (tt_linalg.f90:651-845, dgmresr_hh_fort);
//   * the pivoted incomplete Cholesky that picks the enrichment ("kick")
//     directions (tt_linalg.f90:540-643, duchol_fort);
//   * the half contractions of the local operator used to build the kick
//     block (tt_linalg.f90:210-256, dbfun3_right / dbfun3_left).
#include "amen.h"
#include "arnoldi_gs.h"
#include "mlp.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

namespace ttf {

namespace {

int grid_for(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 16384); }

__global__ void scale_copy_kernel(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = alpha * x[e];
}

// out = u / sqrt(nrm2[0]), or 0 when that norm is 0: the normalisation of
// an Arnoldi step without a host round trip.
__global__ void scale_by_norm_kernel(int64_t N, const double* __restrict__ nrm2, const double* __restrict__ u,
                                     double* __restrict__ out) {
  const double h = sqrt(fmax(nrm2[0], 0.0));
  const double inv = h > 0.0 ? 1.0 / h : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = u[i] * inv;
}

// y = a x + b y
__global__ void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b, double* y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = a * x[e] + b * y[e];
}

// ---------------------------------------------------------------- djac_gen
// Diagonal blocks of the local matrix restricted as the reference does it
// (tt_linalg.f90:397-491).  d1(b1, a1) = Phi1(b1, b1, a1), d2(b2, a2) =
// Phi2(b2, a2, b2); A(a1, i, j, a2).
//  'c': block (b1, b2) of size n: M(i, j) = sum d1(b1,a1) A(a1,i,j,a2) d2(b2,a2),
//       stored at (b2 + rx2 b1) n^2 (layout (n, n, rx2, rx1));
//  'l': block b2 of size rx1 n: M((c1,i),(b1,j)) = sum Phi1(c1,b1,a1) A d2(b2,a2);
//  'r': block b1 of size n rx2: M((i,c2),(j,b2)) = sum d1(b1,a1) A Phi2(b2,a2,c2).
__global__ void jac_blocks_kernel(char type, int rx1, int n, int rx2, int ra1, int ra2,
                                  const double* __restrict__ Phi1, const double* __restrict__ A,
                                  const double* __restrict__ Phi2, double* __restrict__ M) {
  int64_t bs, nblk;
  if (type == 'c') { bs = n; nblk = (int64_t)rx1 * rx2; }
  else if (type == 'l') { bs = (int64_t)rx1 * n; nblk = rx2; }
  else { bs = (int64_t)n * rx2; nblk = rx1; }
  const int64_t total = bs * bs * nblk;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e % bs, col = (e / bs) % bs, blk = e / (bs * bs);
    double s = 0.0;
    if (type == 'c') {
      const int i = (int)row, j = (int)col, b2 = (int)(blk % rx2), b1 = (int)(blk / rx2);
      for (int a2 = 0; a2 < ra2; ++a2) {
        const double d2 = Phi2[b2 + (int64_t)rx2 * (a2 + (int64_t)ra2 * b2)];
        double t = 0.0;
        for (int a1 = 0; a1 < ra1; ++a1)
          t += Phi1[b1 + (int64_t)rx1 * (b1 + (int64_t)rx1 * a1)] *
               A[a1 + (int64_t)ra1 * (i + (int64_t)n * (j + (int64_t)n * a2))];
        s += t * d2;
      }
    } else if (type == 'l') {
      const int c1 = (int)(row % rx1), i = (int)(row / rx1), b1 = (int)(col % rx1), j = (int)(col / rx1);
      const int b2 = (int)blk;
      for (int a2 = 0; a2 < ra2; ++a2) {
        const double d2 = Phi2[b2 + (int64_t)rx2 * (a2 + (int64_t)ra2 * b2)];
        double t = 0.0;
        for (int a1 = 0; a1 < ra1; ++a1)
          t += Phi1[c1 + (int64_t)rx1 * (b1 + (int64_t)rx1 * a1)] *
               A[a1 + (int64_t)ra1 * (i + (int64_t)n * (j + (int64_t)n * a2))];
        s += t * d2;
      }
    } else {
      const int i = (int)(row % n), c2 = (int)(row / n), j = (int)(col % n), b2 = (int)(col / n);
      const int b1 = (int)blk;
      for (int a2 = 0; a2 < ra2; ++a2) {
        const double p2 = Phi2[b2 + (int64_t)rx2 * (a2 + (int64_t)ra2 * c2)];
        double t = 0.0;
        for (int a1 = 0; a1 < ra1; ++a1)
          t += Phi1[b1 + (int64_t)rx1 * (b1 + (int64_t)rx1 * a1)] *
               A[a1 + (int64_t)ra1 * (i + (int64_t)n * (j + (int64_t)n * a2))];
        s += t * p2;
      }
    }
    M[e] = s;
  }
}

// In-place inverse of `count` n x n column-major matrices (n <= 128; matrix
// b at M + b stride, leading dimension ldm) by Gauss-Jordan elimination with
// partial (row) pivoting in LDS, one work-group per matrix; the row
// interchanges are undone as column interchanges at the end.  A zero pivot
// leaves the block's info flag set (the first such step) and the block's
// entries non-finite.  With todo, only the blocks whose todo flag is set.
constexpr int kInvMaxN = 128;

__global__ __launch_bounds__(256) void batched_inverse_kernel(int n, double* __restrict__ M, int64_t ldm,
                                                              int64_t stride, int* __restrict__ info,
                                                              const int* __restrict__ todo) {
  extern __shared__ double a[];   // column-major, ld = n + 1
  __shared__ int piv[kInvMaxN];
  __shared__ double rmax[256];
  __shared__ int rarg[256];
  if (todo && todo[blockIdx.x] == 0) return;   // uniform: a block another kernel inverted
  const int ld = n + 1, tid = threadIdx.x, nt = blockDim.x;
  double* g = M + (int64_t)blockIdx.x * stride;
  for (int e = tid; e < n * n; e += nt) a[(e % n) + ld * (e / n)] = g[(e % n) + ldm * (e / n)];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // pivot row: largest |a(i, k)|, i >= k (first index among ties)
    double best = -1.0;
    int arg = k;
    for (int i = k + tid; i < n; i += nt) {
      const double v = fabs(a[i + ld * k]);
      if (v > best) { best = v; arg = i; }
    }
    rmax[tid] = best;
    rarg[tid] = arg;
    __syncthreads();
    for (int o = nt / 2; o > 0; o >>= 1) {
      if (tid < o) {
        const double v2 = rmax[tid + o];
        const int a2 = rarg[tid + o];
        if (v2 > rmax[tid] || (v2 == rmax[tid] && a2 < rarg[tid])) { rmax[tid] = v2; rarg[tid] = a2; }
      }
      __syncthreads();
    }
    const int p = rarg[0];
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int j = tid; j < n; j += nt) {
        const double t = a[k + ld * j];
        a[k + ld * j] = a[p + ld * j];
        a[p + ld * j] = t;
      }
    __syncthreads();
    const double pv = a[k + ld * k];
    if (tid == 0 && pv == 0.0 && info[blockIdx.x] == 0) info[blockIdx.x] = k + 1;
    const double inv = 1.0 / pv;
    __syncthreads();
    // row k scaled, with its pivot position becoming 1/pivot
    for (int j = tid; j < n; j += nt) a[k + ld * j] = (j == k) ? inv : a[k + ld * j] * inv;
    __syncthreads();
    // eliminate column k from the other rows; column k becomes -f / pivot
    for (int e = tid; e < n * n; e += nt) {
      const int i = e % n, j = e / n;
      if (i == k) continue;
      const double f = a[i + ld * k];
      if (j == k) continue;
      a[i + ld * j] -= f * a[k + ld * j];
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt)
      if (i != k) a[i + ld * k] = -a[i + ld * k] * inv;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int i = tid; i < n; i += nt) {
        const double t = a[i + ld * k];
        a[i + ld * k] = a[i + ld * p];
        a[i + ld * p] = t;
      }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += nt) g[(e % n) + ldm * (e / n)] = a[(e % n) + ld * (e / n)];
}

// In-place inverse of symmetric positive definite n x n blocks (n <= 128;
// block b at M + b stride, leading dimension ldm) by Gauss-Jordan
// elimination without pivoting, one work-group of 256 per block with the
// block in registers: thread t owns the 8 x 8 tile at rows 8 (t % 16) and
// columns 8 (t / 16) of the block padded to order 128 with the identity.
// Step k broadcasts column k and row k through LDS, double-buffered so that
// each step has one barrier.  The pivots are those of the LDL^T
// factorisation, all positive exactly when the block is positive definite;
// a pivot that is not positive sets info[b] and leaves the block as it
// was.  A block whose info[b] is already set is skipped.
__global__ __launch_bounds__(256) void spd_inverse_tiled_kernel(int n, double* __restrict__ M, int64_t ldm,
                                                                int64_t stride, int* __restrict__ info) {
  __shared__ double colb[2][128];
  __shared__ double rowb[2][128];
  if (info[blockIdx.x] != 0) return;   // uniform
  const int t = threadIdx.x, ti = t & 15, tj = t >> 4;
  double* g = M + (int64_t)blockIdx.x * stride;
  double a[8][8];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int i = 8 * ti + r, j = 8 * tj + s;
      a[r][s] = (i < n && j < n) ? g[i + ldm * j] : (i == j ? 1.0 : 0.0);
    }
  bool bad = false;
  const int nkb = (n + 7) / 8;
  for (int kb = 0; kb < nkb; ++kb) {
    const bool rowk = ti == kb, colk = tj == kb;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = 8 * kb + c;
      if (k < n) {   // uniform
        const int buf = k & 1;
        if (colk) {
#pragma unroll
          for (int r = 0; r < 8; ++r) colb[buf][8 * ti + r] = a[r][c];
        }
        if (rowk) {
#pragma unroll
          for (int s = 0; s < 8; ++s) rowb[buf][8 * tj + s] = a[c][s];
        }
        __syncthreads();
        const double pv = rowb[buf][k];
        bad = bad || !(pv > 0.0);
        const double inv = 1.0 / pv;
        double f[8], h[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) f[r] = colb[buf][8 * ti + r];
#pragma unroll
        for (int s = 0; s < 8; ++s) h[s] = rowb[buf][8 * tj + s] * inv;
        // a(i, j) - a(i, k) a(k, j) / pv off row and column k; row k scaled
        // by 1 / pv; column k -a(i, k) / pv; the pivot 1 / pv
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            double v = a[r][s] - f[r] * h[s];
            if (s == c && colk) v = -f[r] * inv;
            if (r == c && rowk) v = (s == c && colk) ? inv : h[s];
            a[r][s] = v;
          }
      }
    }
  }
  // every thread read the same pivots, so `bad` is the same everywhere
  if (bad) {
    if (t == 0) info[blockIdx.x] = 1;
    return;
  }
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int i = 8 * ti + r, j = 8 * tj + s;
      if (i < n && j < n) g[i + ldm * j] = a[r][s];
    }
}

// flag[b] = 1 unless block b (n x n, consecutive blocks of n^2) is exactly
// symmetric with a positive diagonal; one work-group per block.
__global__ __launch_bounds__(256) void sym_check_kernel(int n, const double* __restrict__ M, int* __restrict__ flag) {
  const double* g = M + (int64_t)blockIdx.x * n * n;
  bool bad = false;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    if (i < j) bad = bad || g[i + (int64_t)n * j] != g[j + (int64_t)n * i];
    else if (i == j) bad = bad || !(g[i + (int64_t)n * i] > 0.0);
  }
  if (bad) flag[blockIdx.x] = 1;   // every writer stores the same value
}

// y(blk) = J(blk) x(blk) for all blocks: block b has size bs, matrix at
// J + b bs^2 (column-major), vector entries at off(b) + s k with
// off(b) = (b % nb0) s0 + (b / nb0) s1.
__global__ void jac_apply_kernel(int bs, int64_t nblk, int nb0, int64_t s0, int64_t s1, int64_t s,
                                 const double* __restrict__ J, const double* __restrict__ x,
                                 double* __restrict__ y) {
  const int64_t total = (int64_t)bs * nblk;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % bs);
    const int64_t b = e / bs;
    const int64_t off = (b % nb0) * s0 + (b / nb0) * s1;
    const double* Jb = J + b * (int64_t)bs * bs + i;
    double acc = 0.0;
    for (int j = 0; j < bs; ++j) acc += Jb[(int64_t)bs * j] * x[off + s * j];
    y[off + s * i] = acc;
  }
}

void jac_layout(char type, int rx1, int n, int rx2, int* bs, int64_t* nblk, int* nb0, int64_t* s0,
                int64_t* s1, int64_t* s) {
  if (type == 'c') {   // block (b1, b2) = x(b1, :, b2): b = b2 + rx2 b1
    *bs = n; *nblk = (int64_t)rx1 * rx2; *nb0 = rx2; *s0 = (int64_t)rx1 * n; *s1 = 1; *s = rx1;
  } else if (type == 'l') {   // block b2 = x(:, :, b2)
    *bs = rx1 * n; *nblk = rx2; *nb0 = rx2; *s0 = (int64_t)rx1 * n; *s1 = 0; *s = 1;
  } else {   // 'r': block b1 = x(b1, :, :)
    *bs = n * rx2; *nblk = rx1; *nb0 = rx1; *s0 = 1; *s1 = 0; *s = rx1;
  }
}

}  // namespace

char norm_ptype(char p) { return (char)((p >= 'A' && p <= 'Z') ? p - 'A' + 'a' : p); }

size_t jac_size(char type, int rx1, int n, int rx2) {
  type = norm_ptype(type);
  if (type == 'n') return 0;
  int bs, nb0;
  int64_t nblk, s0, s1, s;
  jac_layout(type, rx1, n, rx2, &bs, &nblk, &nb0, &s0, &s1, &s);
  return (size_t)bs * bs * nblk;
}

void jac_gen(char type, int rx1, int n, int rx2, int ra1, int ra2, const double* Phi1, const double* A,
             const double* Phi2, double* jacs, hipStream_t st) {
  type = norm_ptype(type);
  if (type == 'n') return;
  TTF_REQUIRE(type == 'c' || type == 'l' || type == 'r', "djac_gen: unknown ptype '%c'", type);
  int bs, nb0;
  int64_t nblk, s0, s1, s;
  jac_layout(type, rx1, n, rx2, &bs, &nblk, &nb0, &s0, &s1, &s);
  TTF_REQUIRE(bs <= kInvMaxN, "djac_gen: block size %d exceeds %d for ptype '%c'", bs, kInvMaxN, type);
  TTF_REQUIRE(nblk <= 0x7fffffff, "djac_gen: too many blocks");
  const int64_t total = (int64_t)bs * bs * nblk;
  hipLaunchKernelGGL(jac_blocks_kernel, dim3(grid_for(total)), dim3(256), 0, st, type, rx1, n, rx2, ra1, ra2,
                     Phi1, A, Phi2, jacs);
  TTF_HIP(hipGetLastError());
  int* info = (int*)arena(kArenaSmall, sizeof(int) * (size_t)nblk * 2);
  int* other = info + nblk;   // blocks left to the pivoting kernel
  TTF_HIP(hipMemsetAsync(info, 0, sizeof(int) * (size_t)nblk * 2, st));
  const size_t lds = sizeof(double) * (size_t)bs * (bs + 1);
  static std::atomic<bool> attr{false};   // concurrent first calls only repeat the call
  if (!attr) {
    TTF_HIP(hipFuncSetAttribute((const void*)batched_inverse_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kInvMaxN * (kInvMaxN + 1))));
    attr = true;
  }
  // Exactly symmetric blocks with positive pivots (the blocks of symmetric
  // operators: identical products in identical order make B(i, j) and
  // B(j, i) the same double) go through the register-tiled inverse without
  // pivoting; the others, and any that turns out not positive definite,
  // through Gauss-Jordan with partial pivoting.  Both are the inverse to
  // rounding, which is all dgesv's solve promises.
  hipLaunchKernelGGL(sym_check_kernel, dim3((unsigned)nblk), dim3(256), 0, st, bs, jacs, other);
  hipLaunchKernelGGL(spd_inverse_tiled_kernel, dim3((unsigned)nblk), dim3(256), 0, st, bs, jacs, (int64_t)bs,
                     (int64_t)bs * bs, other);
  hipLaunchKernelGGL(batched_inverse_kernel, dim3((unsigned)nblk), dim3(256), lds, st, bs, jacs, (int64_t)bs,
                     (int64_t)bs * bs, info, (const int*)other);
  TTF_HIP(hipGetLastError());
  std::vector<int> hinfo(nblk);
  read_back(hinfo.data(), info, sizeof(int) * (size_t)nblk, st);
  for (int64_t b = 0; b < nblk; ++b)
    TTF_REQUIRE(hinfo[b] == 0, "djac_gen: block %lld is singular (zero pivot %d)", (long long)b, hinfo[b]);
}

void dense_inverse(int n, double* M, hipStream_t st) {
  TTF_REQUIRE(n >= 1 && n <= kInvMaxN, "dense_inverse: order %d outside 1..%d", n, kInvMaxN);
  int* info = (int*)arena(kArenaSmall, sizeof(int) * 4);
  TTF_HIP(hipMemsetAsync(info, 0, sizeof(int), st));
  static std::atomic<bool> attr{false};
  if (!attr) {
    TTF_HIP(hipFuncSetAttribute((const void*)batched_inverse_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kInvMaxN * (kInvMaxN + 1))));
    attr = true;
  }
  hipLaunchKernelGGL(batched_inverse_kernel, dim3(1), dim3(256), sizeof(double) * (size_t)n * (n + 1), st, n, M,
                     (int64_t)n, (int64_t)n * n, info, nullptr);
  TTF_HIP(hipGetLastError());
  int h = 0;
  read_back(&h, info, sizeof(int), st);
  TTF_REQUIRE(h == 0, "dense_inverse: singular matrix (zero pivot %d)", h);
}

namespace {
// A21 = A12^T in each nb x nb block (h1 leading rows and columns).
__global__ void mirror_upper_kernel(int nb, int h1, int64_t nblk, double* __restrict__ M) {
  const int h2 = nb - h1;
  const int64_t per = (int64_t)h2 * h1, total = per * nblk;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, r = e % per;
    const int i = h1 + (int)(r % h2), j = (int)(r / h2);
    double* Mb = M + b * (int64_t)nb * nb;
    Mb[i + (int64_t)nb * j] = Mb[j + (int64_t)nb * i];
  }
}
}  // namespace

bool spd_block_inverse(int nb, int64_t nblk, double* M, hipStream_t st) {
  TTF_REQUIRE(nb >= 1 && nb <= 2 * kInvMaxN, "spd_block_inverse: order %d outside 1..%d", nb, 2 * kInvMaxN);
  TTF_REQUIRE(nblk >= 1 && nblk <= 0x7fffffff, "spd_block_inverse: %lld blocks", (long long)nblk);
  const int h1 = nb <= kInvMaxN ? nb : nb / 2, h2 = nb - h1;
  const int64_t nb2 = (int64_t)nb * nb, hh = (int64_t)h1 * h2;
  char* ws = (char*)arena(kArenaEigPrec2, sizeof(double) * (size_t)(hh * nblk) + sizeof(int) * (size_t)nblk);
  double* T = (double*)ws;
  int* info = (int*)(ws + sizeof(double) * (size_t)(hh * nblk));
  TTF_HIP(hipMemsetAsync(info, 0, sizeof(int) * (size_t)nblk, st));
  auto inv = [&](int n, double* base) {
    hipLaunchKernelGGL(spd_inverse_tiled_kernel, dim3((unsigned)nblk), dim3(256), 0, st, n, base, (int64_t)nb, nb2,
                       info);
    TTF_HIP(hipGetLastError());
  };
  auto batched = [&](GemmOp<double> g, int64_t sa, int64_t sb, int64_t sc) {
    g.nb0 = (int)nblk;
    g.nbatch = (int)nblk;
    g.sAb0 = sa;
    g.sBb0 = sb;
    g.sCb0 = sc;
    gemm_launch(g, st);
  };
  if (h2 == 0) {
    inv(nb, M);
  } else {
    // [A11 A12; A21 A22] with A21 = A12^T: S = A22 - A21 A11^-1 A12 is
    // positive definite exactly when the block is (given A11 is), and
    // A^-1 = [A11^-1 + T S^-1 T^T, -T S^-1; (-T S^-1)^T, S^-1], T = A11^-1 A12
    double* A11 = M;
    double* A12 = M + (int64_t)nb * h1;
    double* A21 = M + h1;
    double* A22 = M + h1 + (int64_t)nb * h1;
    inv(h1, A11);
    batched(colmajor<double>('N', 'N', h1, h2, h1, 1.0, A11, nb, A12, nb, 0.0, T, h1), nb2, nb2, hh);
    batched(colmajor<double>('N', 'N', h2, h2, h1, -1.0, A21, nb, T, h1, 1.0, A22, nb), nb2, hh, nb2);
    inv(h2, A22);
    batched(colmajor<double>('N', 'N', h1, h2, h2, -1.0, T, h1, A22, nb, 0.0, A12, nb), hh, nb2, nb2);
    batched(colmajor<double>('N', 'T', h1, h1, h2, -1.0, A12, nb, T, h1, 1.0, A11, nb), nb2, hh, nb2);
    hipLaunchKernelGGL(mirror_upper_kernel, dim3(grid_for(hh * nblk)), dim3(256), 0, st, nb, h1, nblk, M);
    TTF_HIP(hipGetLastError());
  }
  std::vector<int> hinfo(nblk);
  read_back(hinfo.data(), info, sizeof(int) * (size_t)nblk, st);
  for (int64_t b = 0; b < nblk; ++b)
    if (hinfo[b] != 0) return false;
  return true;
}

void jac_apply(char type, int rx1, int n, int rx2, const double* jacs, const double* x, double* y,
               hipStream_t st) {
  type = norm_ptype(type);
  const int64_t N = (int64_t)rx1 * n * rx2;
  if (type == 'n') {
    if (x != y) TTF_HIP(hipMemcpyAsync(y, x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    return;
  }
  TTF_REQUIRE(type == 'c' || type == 'l' || type == 'r', "djac_apply: unknown ptype '%c'", type);
  int bs, nb0;
  int64_t nblk, s0, s1, s;
  jac_layout(type, rx1, n, rx2, &bs, &nblk, &nb0, &s0, &s1, &s);
  const double* xin = x;
  if (x == y) {   // in place: stage the input
    double* t = (double*)arena(kArenaSolver2, sizeof(double) * N);
    TTF_HIP(hipMemcpyAsync(t, x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    xin = t;
  }
  hipLaunchKernelGGL(jac_apply_kernel, dim3(grid_for((int64_t)bs * nblk)), dim3(256), 0, st, bs, nblk, nb0, s0,
                     s1, s, jacs, xin, y);
  TTF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- GMRES
namespace {
__global__ void d2f_kernel(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = (float)x[e];
}
__global__ void f2d_kernel(int64_t n, const float* __restrict__ x, double* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = (double)x[e];
}
void to_float(int64_t n, const double* x, float* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(d2f_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  TTF_HIP(hipGetLastError());
}
// columns of a block with leading dimensions: y(:, c) = x(:, c), c < k
__global__ void d2f_cols_kernel(int64_t rows, int k, const double* __restrict__ x, int64_t ldx,
                                float* __restrict__ y, int64_t ldy) {
  const int64_t total = rows * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, c = e / rows;
    y[i + ldy * c] = (float)x[i + ldx * c];
  }
}
__global__ void f2d_cols_kernel(int64_t rows, int k, const float* __restrict__ x, int64_t ldx,
                                double* __restrict__ y, int64_t ldy) {
  const int64_t total = rows * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, c = e / rows;
    y[i + ldy * c] = (double)x[i + ldx * c];
  }
}
}  // namespace

MixedOp make_mixed(const LocalOp& op, hipStream_t st, int nvec) {
  const Shape& s = op.s;
  TTF_REQUIRE(nvec >= 1, "make_mixed: nvec=%d", nvec);
  const int64_t n1 = (int64_t)s.ry1 * s.rx1 * s.ra1, na = (int64_t)s.ra1 * s.n * s.m * s.ra2;
  const int64_t n2 = (int64_t)s.rx2 * s.ra2 * s.ry2;
  const int64_t nx = (int64_t)s.rx1 * s.m * s.rx2 * nvec, ny = (int64_t)s.ry1 * s.n * s.ry2 * nvec;
  float* base = (float*)arena(kArenaMixed, sizeof(float) * (size_t)(n1 + na + n2 + nx + ny + 64));
  MixedOp mo;
  mo.s = s;
  mo.nvec = nvec;
  mo.phi1 = base;
  mo.A = mo.phi1 + n1;
  mo.phi2 = mo.A + na;
  mo.x = mo.phi2 + n2;
  mo.y = mo.x + nx;
  to_float(n1, op.phi1, mo.phi1, st);
  to_float(na, op.A, mo.A, st);
  to_float(n2, op.phi2, mo.phi2, st);
  return mo;
}

void MixedOp::apply(const double* xin, double* yout, hipStream_t st) const {
  const int64_t nx = (int64_t)s.rx1 * s.m * s.rx2, ny = (int64_t)s.ry1 * s.n * s.ry2;
  to_float(nx, xin, x, st);
  bfun3_chain<float>(s, 1, phi1, A, phi2, x, nx, y, ny, st);
  hipLaunchKernelGGL(f2d_kernel, dim3(grid_for(ny)), dim3(256), 0, st, ny, y, yout);
  TTF_HIP(hipGetLastError());
}

void MixedOp::apply_block(int k, const double* xin, int64_t ldx, double* yout, int64_t ldy, hipStream_t st) const {
  TTF_REQUIRE(k >= 1 && k <= nvec, "MixedOp: %d vectors for staging of %d", k, nvec);
  const int64_t nx = (int64_t)s.rx1 * s.m * s.rx2, ny = (int64_t)s.ry1 * s.n * s.ry2;
  hipLaunchKernelGGL(d2f_cols_kernel, dim3(grid_for(nx * k)), dim3(256), 0, st, nx, k, xin, ldx, x, nx);
  TTF_HIP(hipGetLastError());
  bfun3_chain<float>(s, k, phi1, A, phi2, x, nx, y, ny, st);
  hipLaunchKernelGGL(f2d_cols_kernel, dim3(grid_for(ny * k)), dim3(256), 0, st, ny, k, y, ny, yout, ldy);
  TTF_HIP(hipGetLastError());
}

GmresStats gmres_solve(const LocalOp& op, const JacPrec& P, const double* rhs, double* sol, int nrestart,
                       double tol, int niters, hipStream_t st, const MixedOp* mix) {
  GmresStats gs;
  const int64_t N = op.N();
  const int m = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)nrestart, N, (int64_t)63}));
  // device workspace: V (N x (m+1)) | u | t | per Arnoldi step its dot
  // products (two passes of 64) and squared norm, read every kGmresSync steps
  constexpr int kGmresSync = 4;
  double* V = (double*)arena(kArenaSolver, sizeof(double) * ((size_t)N * (m + 3) + (size_t)129 * m));
  double* u = V + (size_t)N * (m + 1);
  double* t = u + N;
  double* Hs = t + N;
  double* hb = (double*)arena(kArenaSmall, sizeof(double) * (2 * 64 + 8));
  TTF_HIP(hipMemsetAsync(sol, 0, sizeof(double) * N, st));
  const int gN = grid_for(N);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), hh(2 * 64 + 8);
  double nrmrhs = 0.0, curres = 1.0;
  // B P v; Krylov directions use the fp32 operator when given, restart
  // residuals (exact) always the FP64 one
  auto apply_BP = [&](const double* v, double* out, bool exact) {
    const double* in = v;
    if (P.type != 'n') {
      jac_apply(P.type, P.rx1, P.n, P.rx2, P.jacs, v, t, st);
      if (P.type == 'c' && P.phi1 && mlp_collecting())
        mlp_collect_pair(P.rx1, P.n, P.rx2, P.ra1, P.ra2, P.phi1, P.A, P.phi2, v, t, st);
      in = t;
    }
    if (mix && !exact) mix->apply(in, out, st);
    else op.apply(1, in, N, out, N, st);
    ++gs.matvecs;
  };
  for (int it = 0; it < niters; ++it) {
    gs.it = it;
    // r = rhs - B P sol into V(:, 0)
    if (it == 0) {
      TTF_HIP(hipMemcpyAsync(V, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    } else {
      apply_BP(sol, V, true);
      hipLaunchKernelGGL(axpby_kernel, dim3(gN), dim3(256), 0, st, N, 1.0, rhs, -1.0, V);
      TTF_HIP(hipGetLastError());
    }
    tsgram((int)N, 1, 1, V, N, V, N, hb, 1, st);
    read_back(hh.data(), hb, sizeof(double), st);
    const double beta = std::sqrt(std::max(hh[0], 0.0));
    if (it == 0) nrmrhs = beta;
    if (beta == 0.0) { curres = 0.0; break; }
    hipLaunchKernelGGL(scale_copy_kernel, dim3(gN), dim3(256), 0, st, N, 1.0 / beta, V, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    // Arnoldi steps are queued without waiting: each step's coefficients
    // stay on the device and the next basis vector is normalised there; the
    // host reads the coefficients every kGmresSync steps and applies the
    // Givens rotations in order, stopping at the first step that meets the
    // tolerance (steps queued past it are not used).
    int jj = 0, j_done = 0;
    bool stop = false;
    for (int j = 0; j < m && !stop; ++j) {
      apply_BP(V + (size_t)N * j, u, false);
      const int k = j + 1;
      double* hs = Hs + (size_t)129 * j;
      // classical Gram-Schmidt, twice, against V(:, 0..j), and the next
      // basis vector u / |u|
      if (gs::fits(N, k)) {
        gs::step<double>(N, k, V, N, u, hs, 0.0, j + 1 < m ? V + (size_t)N * (j + 1) : nullptr, st);
      } else {
        for (int pass = 0; pass < 2; ++pass) {
          double* h = hs + 64 * pass;
          tsgram((int)N, k, 1, V, N, u, N, h, k, st);
          GemmOp<double> gm = colmajor<double>('N', 'N', (int)N, 1, k, -1.0, V, N, h, k, 1.0, u, N);
          gemm_launch(gm, st);
        }
        tsgram((int)N, 1, 1, u, N, u, N, hs + 128, 1, st);
        if (j + 1 < m) {
          hipLaunchKernelGGL(scale_by_norm_kernel, dim3(gN), dim3(256), 0, st, N, hs + 128, u,
                             V + (size_t)N * (j + 1));
          TTF_HIP(hipGetLastError());
        }
      }
      if ((j + 1) % kGmresSync != 0 && j + 1 < m) continue;
      const int cnt = j + 1 - j_done;
      hh.resize((size_t)129 * cnt);
      read_back(hh.data(), Hs + (size_t)129 * j_done, sizeof(double) * 129 * cnt, st);
      for (int jc = j_done; jc <= j; ++jc) {
        const double* c = hh.data() + (size_t)129 * (jc - j_done);
        const int kc = jc + 1;
        double* col = &H[(size_t)(m + 1) * jc];
        for (int i = 0; i < kc; ++i) col[i] = c[i] + c[64 + i];
        const double hn = std::sqrt(std::max(c[128], 0.0));
        col[kc] = hn;
        for (int i = 0; i < jc; ++i) {   // previous rotations
          const double a = col[i], b = col[i + 1];
          col[i] = cs[i] * a + sn[i] * b;
          col[i + 1] = -sn[i] * a + cs[i] * b;
        }
        const double r = std::hypot(col[jc], col[kc]);
        cs[jc] = r > 0.0 ? col[jc] / r : 1.0;
        sn[jc] = r > 0.0 ? col[kc] / r : 0.0;
        col[jc] = r;
        col[kc] = 0.0;
        g[kc] = -sn[jc] * g[jc];
        g[jc] = cs[jc] * g[jc];
        curres = std::fabs(g[kc]) / nrmrhs;
        jj = kc;
        if (curres < tol || hn == 0.0) {
          stop = true;
          break;
        }
      }
      j_done = j + 1;
    }
    gs.j = jj;
    // y = R^-1 g (upper triangular jj x jj), sol += V(:, 0..jj-1) y
    for (int i = jj - 1; i >= 0; --i) {
      double s = g[i];
      for (int l = i + 1; l < jj; ++l) s -= H[(size_t)(m + 1) * l + i] * y[l];
      y[i] = s / H[(size_t)(m + 1) * i + i];
    }
    TTF_HIP(hipMemcpyAsync(hb, y.data(), sizeof(double) * jj, hipMemcpyHostToDevice, st));
    GemmOp<double> gm = colmajor<double>('N', 'N', (int)N, 1, jj, 1.0, V, N, hb, jj, 1.0, sol, N);
    gemm_launch(gm, st);
    TTF_HIP(hipStreamSynchronize(st));   // hb is reused by the next restart
    if (curres < tol) break;
  }
  gs.relres = curres;
  return gs;
}

GmresStats gmres_flex(const LocalOp& op, const PrecFn& M, const double* rhs, double* x, int nrestart, double tol,
                      int niters, hipStream_t st, const MixedOp* mix) {
  GmresStats gs;
  const int64_t N = op.N();
  const int m = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)nrestart, N, (int64_t)63}));
  // V (N x (m+1)) | u ; Z (N x m) holds the preconditioned directions
  double* V = (double*)arena(kArenaSolver, sizeof(double) * (size_t)N * (m + 2));
  double* u = V + (size_t)N * (m + 1);
  double* Z = (double*)arena(kArenaZ, sizeof(double) * (size_t)N * m);
  double* hb = (double*)arena(kArenaSmall, sizeof(double) * (2 * 64 + 8));
  TTF_HIP(hipMemsetAsync(x, 0, sizeof(double) * N, st));
  const int gN = grid_for(N);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), hh(2 * 64 + 8);
  double nrmrhs = 0.0, curres = 1.0;
  for (int it = 0; it < niters; ++it) {
    gs.it = it;
    if (it == 0) {
      TTF_HIP(hipMemcpyAsync(V, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    } else {
      op.apply(1, x, N, V, N, st);
      ++gs.matvecs;
      hipLaunchKernelGGL(axpby_kernel, dim3(gN), dim3(256), 0, st, N, 1.0, rhs, -1.0, V);
      TTF_HIP(hipGetLastError());
    }
    tsgram((int)N, 1, 1, V, N, V, N, hb, 1, st);
    read_back(hh.data(), hb, sizeof(double), st);
    const double beta = std::sqrt(std::max(hh[0], 0.0));
    if (it == 0) nrmrhs = beta;
    if (!(beta > 0.0)) { curres = beta == 0.0 ? 0.0 : curres; break; }
    hipLaunchKernelGGL(scale_copy_kernel, dim3(gN), dim3(256), 0, st, N, 1.0 / beta, V, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int jj = 0;
    for (int j = 0; j < m; ++j) {
      double* zj = Z + (size_t)N * j;
      M(V + (size_t)N * j, zj);
      if (mix) mix->apply(zj, u, st);
      else op.apply(1, zj, N, u, N, st);
      ++gs.matvecs;
      const int k = j + 1;
      // the fused path also writes the next basis vector u / |u|
      const bool fused = gs::fits(N, k);
      if (fused) {
        gs::step<double>(N, k, V, N, u, hb, 0.0, j + 1 < m ? V + (size_t)N * (j + 1) : nullptr, st);
      } else {
        for (int pass = 0; pass < 2; ++pass) {
          double* h = hb + 64 * pass;
          tsgram((int)N, k, 1, V, N, u, N, h, k, st);
          GemmOp<double> gm = colmajor<double>('N', 'N', (int)N, 1, k, -1.0, V, N, h, k, 1.0, u, N);
          gemm_launch(gm, st);
        }
        tsgram((int)N, 1, 1, u, N, u, N, hb + 128, 1, st);
      }
      read_back(hh.data(), hb, sizeof(double) * 129, st);
      double* col = &H[(size_t)(m + 1) * j];
      for (int i = 0; i < k; ++i) col[i] = hh[i] + hh[64 + i];
      const double hn = std::sqrt(std::max(hh[128], 0.0));
      col[k] = hn;
      for (int i = 0; i < j; ++i) {
        const double a = col[i], b = col[i + 1];
        col[i] = cs[i] * a + sn[i] * b;
        col[i + 1] = -sn[i] * a + cs[i] * b;
      }
      const double r = std::hypot(col[j], col[k]);
      cs[j] = r > 0.0 ? col[j] / r : 1.0;
      sn[j] = r > 0.0 ? col[k] / r : 0.0;
      col[j] = r;
      col[k] = 0.0;
      g[k] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      curres = std::fabs(g[k]) / nrmrhs;
      jj = k;
      if (!(curres >= tol) || hn == 0.0) break;   // also stops on a NaN from the preconditioner
      if (j + 1 < m && !fused)
        hipLaunchKernelGGL(scale_copy_kernel, dim3(gN), dim3(256), 0, st, N, 1.0 / hn, u, V + (size_t)N * (j + 1));
    }
    gs.j = jj;
    for (int i = jj - 1; i >= 0; --i) {
      double s = g[i];
      for (int l = i + 1; l < jj; ++l) s -= H[(size_t)(m + 1) * l + i] * y[l];
      y[i] = s / H[(size_t)(m + 1) * i + i];
    }
    TTF_HIP(hipMemcpyAsync(hb, y.data(), sizeof(double) * jj, hipMemcpyHostToDevice, st));
    GemmOp<double> gm = colmajor<double>('N', 'N', (int)N, 1, jj, 1.0, Z, N, hb, jj, 1.0, x, N);
    gemm_launch(gm, st);
    TTF_HIP(hipStreamSynchronize(st));
    if (!(curres >= tol)) break;
  }
  gs.relres = curres;
  return gs;
}

// ---------------------------------------------------------------- duchol
namespace {

// squared column norms of X (m x n, ld m) ('N') or row norms of X (n x m,
// ld n) ('T') into d (n)
__global__ void gram_diag_kernel(char trans, int m, int n, const double* __restrict__ X, double* __restrict__ d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    if (trans == 'n')
      for (int k = 0; k < m; ++k) { const double v = X[k + (int64_t)m * i]; s += v * v; }
    else
      for (int k = 0; k < m; ++k) { const double v = X[i + (int64_t)n * k]; s += v * v; }
    d[i] = s;
  }
}

__global__ void sub_square_kernel(int n, const double* __restrict__ u, double* d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] -= u[i] * u[i];
}

}  // namespace

int duchol(char trans, int m, int n, const double* dA, int numvec, double* du, double* lam_out,
           hipStream_t st) {
  trans = norm_ptype(trans);
  TTF_REQUIRE(trans == 'n' || trans == 't', "duchol: trans '%c'", trans);
  TTF_REQUIRE(numvec >= 1 && numvec <= 32, "duchol: numvec=%d outside 1..32", numvec);
  TTF_REQUIRE(m >= 1 && n >= 1, "duchol: sizes %d x %d", m, n);
  double* ddiag = (double*)arena(kArenaDriver, sizeof(double) * ((size_t)n * (numvec + 2) + 64 * 64 * 3 + 256));
  double* tmp = ddiag + n;                      // n x numvec
  double* small = tmp + (size_t)n * numvec;     // coefficient vectors, eig workspace
  double* hcoef = small + 64 * 64 * 3;
  hipLaunchKernelGGL(gram_diag_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, trans, m, n, dA, ddiag);
  TTF_HIP(hipGetLastError());
  std::vector<double> hd(n), lam(numvec + 1, 0.0), hv(64 * 64), hw(64);
  int cnt = 0, real = 0;
  for (cnt = 0; cnt < numvec; ++cnt) {
    read_back(hd.data(), ddiag, sizeof(double) * n, st);
    int maxi = 0;
    double best = -1.0;
    for (int i = 0; i < n; ++i)
      if (std::fabs(hd[i]) > best) { best = std::fabs(hd[i]); maxi = i; }
    double* uc = du + (size_t)n * cnt;
    // uc = G(:, maxi) with G the Gram matrix of the items
    GemmOp<double> g{};
    g.M = n; g.N = 1; g.K = m; g.nb0 = 1; g.nbatch = 1; g.alpha = 1.0; g.beta = 0.0; g.ksplit = 1;
    g.C = uc; g.sCi = 1; g.sCj = n;
    if (trans == 'n') {   // G = A^T A, A m x n
      g.A = dA; g.sAi = m; g.sAk = 1;
      g.B = dA + (int64_t)m * maxi; g.sBk = 1; g.sBj = m;
    } else {              // G = A A^T, A n x m
      g.A = dA; g.sAi = 1; g.sAk = n;
      g.B = dA + maxi; g.sBk = n; g.sBj = 1;
    }
    gemm_launch(g, st);
    if (cnt > 0) {
      // uc -= sum_j lam_j du(maxi, j) du(:, j)
      std::vector<double> hrow(cnt);
      TTF_HIP(hipMemcpy2DAsync(hrow.data(), sizeof(double), du + maxi, sizeof(double) * n, sizeof(double), cnt,
                               hipMemcpyDeviceToHost, st));
      TTF_HIP(hipStreamSynchronize(st));
      for (int j = 0; j < cnt; ++j) hrow[j] *= lam[j];
      TTF_HIP(hipMemcpyAsync(hcoef, hrow.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, st));
      GemmOp<double> s = colmajor<double>('N', 'N', n, 1, cnt, -1.0, du, n, hcoef, cnt, 1.0, uc, n);
      gemm_launch(s, st);
    }
    double piv = 0.0;
    read_back(&piv, uc + maxi, sizeof(double), st);
    const double maxa = std::sqrt(std::fabs(piv));
    if (maxa < 1e-300) {   // exact zero remainder (tt_linalg.f90:604-607)
      real = cnt + 1;
      break;
    }
    hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for(n)), dim3(256), 0, st, (int64_t)n, 1.0 / maxa, uc, uc);
    hipLaunchKernelGGL(sub_square_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n, uc, ddiag);
    // reorthogonalisation of uc against the orthonormal du(:, 0..cnt-1): the
    // R column dx of the QR of du(:, 0..cnt) (tt_linalg.f90:613-617)
    const int k = cnt + 1;
    std::vector<double> dx(k, 0.0);
    if (cnt > 0) {
      for (int pass = 0; pass < 2; ++pass) {
        tsgram(n, cnt, 1, du, n, uc, n, small + 64 * pass, cnt, st);
        GemmOp<double> s = colmajor<double>('N', 'N', n, 1, cnt, -1.0, du, n, small + 64 * pass, cnt, 1.0, uc, n);
        gemm_launch(s, st);
      }
    }
    tsgram(n, 1, 1, uc, n, uc, n, small + 128, 1, st);
    std::vector<double> hh(129);
    read_back(hh.data(), small, sizeof(double) * 129, st);
    for (int j = 0; j < cnt; ++j) dx[j] = hh[j] + hh[64 + j];
    dx[cnt] = std::sqrt(std::max(hh[128], 0.0));
    if (dx[cnt] > 0.0)
      hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for(n)), dim3(256), 0, st, (int64_t)n, 1.0 / dx[cnt], uc, uc);
    // dv = dx dx^T + diag(lam), eigen-decomposition (ascending)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) hv[i + k * j] = dx[i] * dx[j] + (i == j ? lam[i] : 0.0);
    double* dv = small + 256;
    double* w = dv + 64 * 64;
    double* Q = w + 64;
    TTF_HIP(hipMemcpyAsync(dv, hv.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, st));
    syev_small(k, 1, dv, w, Q, 0, st);
    TTF_HIP(hipMemcpyAsync(hw.data(), w, sizeof(double) * k, hipMemcpyDeviceToHost, st));
    // du(:, 0..cnt) <- du(:, 0..cnt) V
    GemmOp<double> r = colmajor<double>('N', 'N', n, k, k, 1.0, du, n, Q, k, 0.0, tmp, n);
    gemm_launch(r, st);
    TTF_HIP(hipMemcpyAsync(du, tmp, sizeof(double) * (size_t)n * k, hipMemcpyDeviceToDevice, st));
    TTF_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < k; ++i) lam[i] = hw[i];
    real = k;
  }
  if (lam_out)
    for (int i = 0; i < numvec; ++i) lam_out[i] = i < (int)lam.size() ? lam[i] : 0.0;
  return real;
}

// ---------------------------------------------------------------- half contractions
void bfun3_right_half(const Shape& s, const double* A, const double* phi2, const double* x, double* y,
                      hipStream_t st) {
  const size_t sz1 = (size_t)s.rx1 * s.m * s.ra2 * s.ry2;
  double* T1 = (double*)arena(kArenaSolver2, sizeof(double) * sz1);
  bfun3_stage12<double>(s, 1, A, phi2, x, (int64_t)s.rx1 * s.m * s.rx2, T1, y, st);
}

void bfun3_left_half(const Shape& s, const double* phi1, const double* A, const double* x, double* y,
                     hipStream_t st) {
  // T(c1, j, b2, a1) = sum_b1 Phi1(c1, b1, a1) x(b1, j, b2): one GEMM per a1
  const size_t szT = (size_t)s.ry1 * s.m * s.rx2 * s.ra1;
  double* T = (double*)arena(kArenaSolver2, sizeof(double) * szT);
  {
    GemmOp<double> g{};
    g.M = s.ry1; g.N = s.m * s.rx2; g.K = s.rx1;
    g.A = phi1; g.sAi = 1; g.sAk = s.ry1; g.sAb0 = (int64_t)s.ry1 * s.rx1;
    g.B = x; g.sBk = 1; g.sBj = s.rx1; g.sBb0 = 0;
    g.C = T; g.sCi = 1; g.sCj = s.ry1; g.sCb0 = (int64_t)s.ry1 * s.m * s.rx2;
    g.nb0 = s.ra1; g.nbatch = s.ra1; g.alpha = 1.0; g.beta = 0.0; g.ksplit = 1;
    gemm_launch(g, st);
  }
  // y(c1, i, b2, a2) = sum_{a1, j} T(c1, j, b2, a1) A(a1, i, j, a2): batched
  // over (b2, a2), accumulated over a1
  for (int a1 = 0; a1 < s.ra1; ++a1) {
    GemmOp<double> g{};
    g.M = s.ry1; g.N = s.n; g.K = s.m;
    g.A = T + (size_t)a1 * s.ry1 * s.m * s.rx2; g.sAi = 1; g.sAk = s.ry1;
    g.sAb0 = (int64_t)s.ry1 * s.m; g.sAb1 = 0;
    g.B = A + a1; g.sBk = (int64_t)s.ra1 * s.n; g.sBj = s.ra1;
    g.sBb0 = 0; g.sBb1 = (int64_t)s.ra1 * s.n * s.m;
    g.C = y; g.sCi = 1; g.sCj = s.ry1; g.sCb0 = (int64_t)s.ry1 * s.n; g.sCb1 = (int64_t)s.ry1 * s.n * s.rx2;
    g.nb0 = s.rx2; g.nbatch = s.rx2 * s.ra2; g.alpha = 1.0; g.beta = a1 == 0 ? 0.0 : 1.0; g.ksplit = 1;
    gemm_launch(g, st);
  }
}

}  // namespace ttf

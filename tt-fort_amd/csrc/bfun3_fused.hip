// The three stages of the local matvec (ttals.f90:9-32) in one launch, FP64:
//   T1(b1,j,a2,c2) = sum_b2 x(b1,j,b2) phi2(b2,a2,c2)              (MFMA)
//   T2(b1,a1,i,c2) = sum_{j,a2} A(a1,i,j,a2) T1(b1,j,a2,c2)        (vector)
//   y(c1,i,c2)     = sum_{b1,a1} phi1(c1,b1,a1) T2(b1,a1,i,c2)      (MFMA)
// A work-group owns one vector, a tile of WC right-interface columns c2 and
// a range of BB left-interface rows b1; T1 and T2 stay in LDS, so the
// kernel reads x, phi1, phi2 and A and writes y, and nothing else.  With
// the b1 range split in two (to give the launch enough work-groups), the
// two halves of each output are added onto a zero-filled y by two atomic
// additions: 0 + a + b = 0 + b + a exactly, so the result does not depend
// on their order.
#include "contract.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>

namespace ttf {

namespace {

// Eight waves, two per SIMD, so that one wave's matrix-core work overlaps
// the other's load latency.  Each wave owns one 16-row tile of the stage-1
// and stage-3 products (R1 = BB m <= 128, ry1 <= 128) and a quarter of the
// stage-2 tiles (BB WC <= 512).
constexpr int kNW = 8;
constexpr int kNT = 64 * kNW;
constexpr int kT2 = 4;     // stage-2 column tiles per wave
constexpr int kPF1 = 8;    // stage-1 k-steps of x loads kept in flight
constexpr int kPF3 = 16;   // stage-3 k-steps of phi1 loads kept in flight
constexpr int kIn = 12;    // phi2 loads per thread issued together

struct FusedArgs {
  int rx1, m, rx2, ry1, n, ry2, ra1, ra2;
  int nvec, WC, BB, nsplit, ntc2;
  int K1p, K3p;         // stage-1 / stage-3 depths padded to whole prefetch rounds
  int ldt1;             // leading dimension of T1 in LDS (R1 rounded up to 32)
  int off_t1;           // LDS offset (doubles) of T1; phi2 / T2 at 0
  const double* phi1;
  const double* A;
  const double* phi2;
  const double* x;
  int64_t ldx;
  double* y;
  int64_t ldy;
};

__device__ __forceinline__ int sw(int c) { return (c & 7) << 2; }

// TN1 = stage-1 column tiles (N1 = ra2 WC <= 16 TN1).  The inner loops have
// no branches and no masked loads: a global load outside the valid ranges
// reads a clamped in-range element, and the matching entries of the LDS
// operand it meets are zero (x and phi1 entries are finite, so the product
// is an exact zero).
template <int TN1>
__global__ __launch_bounds__(kNT) void bfun3_fused_kernel(FusedArgs a) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kq = lane >> 4, lc = lane & 15;
  const int total = a.nvec * a.ntc2 * a.nsplit;
  // consecutive logical work-groups share an XCD: one vector's tiles (and
  // both halves of each tile) meet in one L2
  const int L = xcd_remap(blockIdx.x, total);
  const int sp = L % a.nsplit;
  const int tc = (L / a.nsplit) % a.ntc2;
  const int v = L / (a.nsplit * a.ntc2);
  const int c0 = tc * a.WC, b0 = sp * a.BB;
  const int wc = min(a.WC, a.ry2 - c0);
  const int bb = min(a.BB, a.rx1 - b0);
  const int R1 = a.BB * a.m, N1 = a.ra2 * a.WC, K1 = a.rx2;
  const int K3 = a.BB * a.ra1;
  // LDS operands are stored with the row index XOR-ed by 4 (col % 8)
  // (rows stay inside their aligned block of 32), so that the 16 columns a
  // matrix-core fragment reads at one row fall in different banks:
  double* P2s = lds;              // P2s[(k ^ sw(c)) + K1p c], c < N1
  double* T2s = lds;              // T2s[(q ^ sw(p)) + K3p p], p < 16
  double* T1s = lds + a.off_t1;   // T1s[(r ^ sw(c)) + ldt1 c], c < N1

  // the first x loads of stage 1 go out before the phi2 slice is staged
  const int64_t xs = (int64_t)a.rx1 * a.m;   // stride of b2 in x
  const double* xa;
  {
    const double* xv = a.x + (int64_t)v * a.ldx;
    const int r = wave * 16 + lc;
    const int jr = r / a.BB, blr = r - jr * a.BB;
    xa = (r < R1 && blr < bb) ? xv + (b0 + blr) + (int64_t)a.rx1 * jr : xv;
  }
  d4_t acc1[TN1];
#pragma unroll
  for (int t = 0; t < TN1; ++t) acc1[t] = d4_t{0, 0, 0, 0};
  double buf1[kPF1];
  int kl = kq;   // the k of this lane's next x load
  auto load1 = [&]() {
    const double t = xa[kl < K1 ? xs * kl : 0];
    kl += 4;
    return t;
  };
#pragma unroll
  for (int p = 0; p < kPF1; ++p) buf1[p] = load1();

  // ---- phi2 columns (a2, c2) for c2 in the tile are contiguous in memory;
  // zero rows k >= K1 and columns beyond the valid c2.  kIn loads per
  // thread are issued before their stores.
  {
    const double* p2 = a.phi2 + (int64_t)K1 * a.ra2 * c0;
    const int nvc = a.ra2 * wc;   // valid columns
    const int nall = a.K1p * N1;
    const int dc = kNT / a.K1p, dk = kNT - dc * a.K1p;   // (k, c) step per kNT elements
    int c = tid / a.K1p, k = tid - c * a.K1p;
    for (int e0 = 0; e0 < nall; e0 += kIn * kNT) {
      double t[kIn];
      int kk[kIn], cc[kIn];
#pragma unroll
      for (int u = 0; u < kIn; ++u) {
        kk[u] = k;
        cc[u] = c;
        const bool ok = c < nvc && k < K1;
        t[u] = p2[ok ? k + (int64_t)K1 * c : 0];
        k += dk;
        c += dc;
        const bool w = k >= a.K1p;
        k = w ? k - a.K1p : k;
        c = w ? c + 1 : c;
      }
#pragma unroll
      for (int u = 0; u < kIn; ++u) {
        const bool ok = cc[u] < nvc && kk[u] < K1;
        if (e0 + u * kNT + tid < nall) P2s[(kk[u] ^ sw(cc[u])) + a.K1p * cc[u]] = ok ? t[u] : 0.0;
      }
    }
  }
  __syncthreads();

  // ---- stage 1: T1 (R1 x N1) = x rows (b1 in range, j) x phi2 slice
  {
    // columns N1..16 TN1 of the B fragments read stale LDS: they only feed
    // result columns that are never stored
    const double* pb = P2s + kq + a.K1p * lc;
    const int swl = sw(lc);
    if (wave * 16 < R1) {   // uniform: waves past the rows only stage data
      for (int k0 = 0; k0 < a.K1p; k0 += 4 * kPF1) {
#pragma unroll
        for (int p = 0; p < kPF1; ++p) {
          const double af = buf1[p];
          buf1[p] = load1();
          double bf[TN1];
#pragma unroll
          for (int t = 0; t < TN1; ++t) bf[t] = pb[k0 + ((4 * p) ^ swl) + a.K1p * 16 * t];
#pragma unroll
          for (int t = 0; t < TN1; ++t) acc1[t] = mfma16(af, bf[t], acc1[t]);
        }
      }
      // lane l, register r holds row l/16 + 4r, column l%16 of each block;
      // rows of b1 outside this work-group's range are stored as zeros
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = wave * 16 + kq + 4 * q;
        const int j = row / a.BB;
        const bool okrow = row - j * a.BB < bb;
#pragma unroll
        for (int t = 0; t < TN1; ++t) {
          const int col = t * 16 + lc;
          if (row < R1 && col < N1) T1s[(row ^ sw(col)) + a.ldt1 * col] = okrow ? acc1[t][q] : 0.0;
        }
      }
    }
  }
  __syncthreads();

  // the first phi1 loads of stage 3 go out before stage 2
  const int c1 = wave * 16 + lc;
  const bool okc = c1 < a.ry1;
  const double* pa = a.phi1 + (okc ? c1 : 0);
  // this lane's k = q = bl + BB a1, advanced by 4 per k-step (BB is a
  // multiple of 4, so the four k lanes wrap together); rows of T2 outside
  // the valid (bl, a1) are zero, so the clamped phi1 loads add nothing
  const int64_t s_b1 = a.ry1, wrap = (int64_t)a.ry1 * a.rx1 - (int64_t)a.BB * a.ry1;
  int bl3 = kq, a13 = 0;
  int64_t off3 = s_b1 * (b0 + kq);
  auto load3 = [&]() {
    const double t = pa[(a13 < a.ra1 && bl3 < bb) ? off3 : 0];
    bl3 += 4;
    off3 += 4 * s_b1;
    const bool w = bl3 >= a.BB;
    bl3 = w ? bl3 - a.BB : bl3;
    a13 = w ? a13 + 1 : a13;
    off3 = w ? off3 + wrap : off3;
    return t;
  };
  double buf3[kPF3];
#pragma unroll
  for (int p = 0; p < kPF3; ++p) buf3[p] = load3();

  // ---- stage 2 on the matrix cores: T2[(a1, i), (bl, cl)] =
  // A[(a1, i), (j, a2)] T1[(j, a2), (bl, cl)], where A (ra1, n, m, ra2) is
  // read as the (ra1 n) x (m ra2) matrix it is in memory; results land at
  // T2s[(bl + BB a1) + ld2 (i + n cl)].  Rows K3..K3p of T2 are zeroed.
  // Wave w takes the column tiles w + kNW s, all k-steps of its tiles
  // issued together.
  {
    const int M2 = a.ra1 * a.n, K2 = a.m * a.ra2, N2 = a.BB * a.WC;
    for (int e = tid; e < (a.K3p - K3) * 16; e += kNT) {
      const int pcol = e & 15, q = K3 + (e >> 4);
      T2s[(q ^ sw(pcol)) + a.K3p * pcol] = 0.0;
    }
    int bl[kT2], cl[kT2];
    bool okn[kT2];
#pragma unroll
    for (int s2 = 0; s2 < kT2; ++s2) {
      const int n0 = (wave + kNW * s2) * 16 + lc;
      okn[s2] = n0 < N2;
      const int nc = okn[s2] ? n0 : 0;
      cl[s2] = nc / a.BB;
      bl[s2] = nc - cl[s2] * a.BB;
    }
    d4_t acc[kT2];
#pragma unroll
    for (int s2 = 0; s2 < kT2; ++s2) acc[s2] = d4_t{0, 0, 0, 0};
    const int rowA = lc < M2 ? lc : 0;
    for (int k0 = 0; k0 < K2; k0 += 4) {
      const int k = k0 + kq;
      const bool okk = k < K2;
      const int kc = okk ? k : 0;
      const int a2 = kc / a.m, j = kc - a2 * a.m;
      const double av = a.A[rowA + M2 * kc];
      const double af = (okk && lc < M2) ? av : 0.0;
      double bf[kT2];
#pragma unroll
      for (int s2 = 0; s2 < kT2; ++s2) {
        const int col = a2 + a.ra2 * cl[s2];
        bf[s2] = T1s[((bl[s2] + a.BB * j) ^ sw(col)) + a.ldt1 * col];
      }
#pragma unroll
      for (int s2 = 0; s2 < kT2; ++s2) acc[s2] = mfma16(af, bf[s2], acc[s2]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rho = kq + 4 * r;   // = a1 + ra1 i
      if (rho < M2) {
        const int i = rho / a.ra1, a1 = rho - i * a.ra1;
#pragma unroll
        for (int s2 = 0; s2 < kT2; ++s2)
          if (okn[s2]) {
            const int pcol = i + a.n * cl[s2];
            T2s[((bl[s2] + a.BB * a1) ^ sw(pcol)) + a.K3p * pcol] = acc[s2][r];
          }
      }
    }
  }
  __syncthreads();

  // ---- stage 3: y (ry1 x N3) = phi1 (c1, (b1 in range, a1)) x T2, two
  // accumulation chains (even and odd k-steps) per wave
  {
    const int N3 = a.n * a.WC;
    if (wave * 16 < a.ry1) {   // uniform
      d4_t acc0 = d4_t{0, 0, 0, 0}, acc1b = d4_t{0, 0, 0, 0};
      const double* pb = T2s + kq + a.K3p * lc;
      const int swl = sw(lc);
      for (int k0 = 0; k0 < a.K3p; k0 += 4 * kPF3) {
#pragma unroll
        for (int p = 0; p < kPF3; p += 2) {
          const double af0 = buf3[p], af1 = buf3[p + 1];
          buf3[p] = load3();
          buf3[p + 1] = load3();
          // swapped operands: lane l, register r holds y(c1 = l%16, p = l/16 + 4r)
          acc0 = mfma16(pb[k0 + ((4 * p) ^ swl)], af0, acc0);
          acc1b = mfma16(pb[k0 + ((4 * p + 4) ^ swl)], af1, acc1b);
        }
      }
      const d4_t acc = acc0 + acc1b;
      double* yv = a.y + (int64_t)v * a.ldy + (int64_t)a.ry1 * a.n * c0;
      const int pmax = min(N3, a.n * wc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = kq + 4 * r;
        if (okc && p < pmax) {
          double* dst = yv + c1 + (int64_t)a.ry1 * p;
          if (a.nsplit > 1) unsafeAtomicAdd(dst, acc[r]);
          else *dst = acc[r];
        }
      }
    }
  }
}

// y(:, c) = 0 for the nvec columns of `rows` entries (leading dimension
// ldy): blockIdx.y is the column, each thread stores pairs of doubles.
__global__ __launch_bounds__(256) void zero_cols_kernel(int64_t rows, double* __restrict__ y, int64_t ldy) {
  double* col = y + (int64_t)blockIdx.y * ldy;
  const int64_t step = (int64_t)gridDim.x * 512;
  if ((reinterpret_cast<uintptr_t>(col) & 15) == 0) {
    for (int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 2; i < rows; i += step) {
      if (i + 1 < rows) *reinterpret_cast<double2*>(col + i) = double2{0.0, 0.0};
      else col[i] = 0.0;
    }
  } else {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows; i += (int64_t)gridDim.x * 256) col[i] = 0.0;
  }
}

constexpr size_t kMaxLds = 160 * 1024;
constexpr int kMaxTN1 = 4;

int round_up(int v, int m) { return (v + m - 1) / m * m; }

bool plan(const Shape& s, int nvec, FusedArgs* a) {
  if (nvec < 1 || nvec > 65535 || s.n > 16 || s.ry1 > 16 * kNW || s.ra1 * s.n > 16 || s.m * s.ra2 > 64) return false;
  const int WC = std::max(1, 16 / s.n);
  if (s.ra2 * WC > 16 * kMaxTN1) return false;
  const int bmax = (16 * kNW) / s.m;   // R1 = BB m <= 128
  const int ntc2 = (s.ry2 + WC - 1) / WC;
  int nsplit = 1, BB = round_up(s.rx1, 4);
  // half the b1 range per work-group when one range does not fit or the
  // launch would leave most compute units idle
  if (BB > bmax || nvec * ntc2 < 192) {
    nsplit = 2;
    BB = round_up((s.rx1 + 1) / 2, 4);
  }
  if (BB > bmax || nsplit * BB < s.rx1 || BB * WC > 16 * kNW * kT2) return false;
  const int R1 = BB * s.m, N1 = s.ra2 * WC, K3 = BB * s.ra1;
  const int TN1 = (N1 + 15) / 16;
  a->K1p = round_up(s.rx2, 4 * kPF1);
  a->K3p = round_up(K3, 4 * kPF3);
  a->ldt1 = round_up(R1, 32);
  // phi2 slice and T2 share the first region; stage 1 reads up to 16 TN1
  // columns of the slice, which stay inside the allocation
  const size_t region0 = std::max((size_t)a->K1p * N1, (size_t)a->K3p * 16);
  a->off_t1 = (int)region0;
  const size_t ndouble = std::max(region0 + (size_t)a->ldt1 * N1, (size_t)a->K1p * 16 * TN1);
  const size_t bytes = sizeof(double) * ndouble;
  if (bytes > kMaxLds) return false;
  a->rx1 = s.rx1; a->m = s.m; a->rx2 = s.rx2; a->ry1 = s.ry1; a->n = s.n; a->ry2 = s.ry2;
  a->ra1 = s.ra1; a->ra2 = s.ra2;
  a->nvec = nvec; a->WC = WC; a->BB = BB; a->nsplit = nsplit; a->ntc2 = ntc2;
  return true;
}

size_t lds_bytes(const FusedArgs& a) {
  const int N1 = a.ra2 * a.WC, TN1 = (N1 + 15) / 16;
  return sizeof(double) * std::max((size_t)a.off_t1 + (size_t)a.ldt1 * N1, (size_t)a.K1p * 16 * TN1);
}

bool fused_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("TTF_BFUN3_FUSED");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

}  // namespace

bool bfun3_fused_fits(const Shape& s, int nvec) {
  FusedArgs a{};
  return fused_enabled() && plan(s, nvec, &a);
}

bool bfun3_fused(const Shape& s, int nvec, const double* phi1, const double* A, const double* phi2,
                 const double* x, int64_t ldx, double* y, int64_t ldy, hipStream_t st) {
  FusedArgs a{};
  if (!fused_enabled() || !plan(s, nvec, &a)) return false;
  a.phi1 = phi1; a.A = A; a.phi2 = phi2; a.x = x; a.ldx = ldx; a.y = y; a.ldy = ldy;
  const int64_t ny = (int64_t)s.ry1 * s.n * s.ry2;
  if (a.nsplit > 1) {
    const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(ny, 512), 1), 1024);
    hipLaunchKernelGGL(zero_cols_kernel, dim3(gx, (unsigned)nvec), dim3(256), 0, st, ny, y, ldy);
    TTF_HIP(hipGetLastError());
  }
  static bool attr = false;   // racing first calls only repeat the call
  if (!attr) {
    TTF_HIP(hipFuncSetAttribute((const void*)bfun3_fused_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxLds));
    TTF_HIP(hipFuncSetAttribute((const void*)bfun3_fused_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxLds));
    TTF_HIP(hipFuncSetAttribute((const void*)bfun3_fused_kernel<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxLds));
    TTF_HIP(hipFuncSetAttribute((const void*)bfun3_fused_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxLds));
    attr = true;
  }
  const unsigned grid = (unsigned)(a.nvec * a.ntc2 * a.nsplit);
  const int tn1 = (a.ra2 * a.WC + 15) / 16;
  const size_t lds = lds_bytes(a);
  switch (tn1) {
    case 1: hipLaunchKernelGGL(bfun3_fused_kernel<1>, dim3(grid), dim3(kNT), lds, st, a); break;
    case 2: hipLaunchKernelGGL(bfun3_fused_kernel<2>, dim3(grid), dim3(kNT), lds, st, a); break;
    case 3: hipLaunchKernelGGL(bfun3_fused_kernel<3>, dim3(grid), dim3(kNT), lds, st, a); break;
    default: hipLaunchKernelGGL(bfun3_fused_kernel<4>, dim3(grid), dim3(kNT), lds, st, a); break;
  }
  TTF_HIP(hipGetLastError());
  return true;
}

}  // namespace ttf

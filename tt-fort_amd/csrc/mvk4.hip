// tt_mvk4 on the device: the approximate product y ~ A x of a TT-matrix and
// a TT vector by alternating projections with enrichment, the sweep of
// tt_adapt_als.f90:67-500.
//
// All cores of A, x and y and the interfaces (phi, the y'Ax projections of
// the reference's dphi_left / dphi_right) stay in HBM; the host steers the
// ranks and the stopping rule.
//   * y0 is left-orthogonalised (:197-222) with the left interfaces built;
//   * each local step sets the core to its projection bfun3(phi_k, A_k,
//     phi_k+1, x_k) and records the relative change (:243-255);
//   * left-to-right steps truncate by SVD with my_chop3 at eps / sqrt(d)
//     (a QR when kickrank = 0) and move the factor right (:269-341);
//   * right-to-left steps truncate likewise and enrich the kept basis with
//     min(kickrank, n ry) directions from the pivoted Cholesky (duchol) of
//     the half product A_k Phi_k+1 x_k, then re-orthogonalise (:349-426);
//   * a sweep ends at the last core; the run stops after the first sweep
//     whose largest change is below eps (:428-447), or after nswp sweeps.
// The reference's rmax only sizes its work arrays (ranks beyond it overrun
// them); here it sets the capacity of the rank buffers, and a rank that
// would exceed the capacity is an error instead.
#include "amen.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

namespace ttf {
int my_chop3(const std::vector<double>& s, int n, double eps);
}

using namespace ttf;

namespace {

// out = x - y, then the two norms |x - y|^2 and |y|^2 summed on the device
__global__ __launch_bounds__(256) void diff_norms_kernel(int64_t n, const double* __restrict__ x,
                                                         const double* __restrict__ y, double* __restrict__ out) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const double d = x[e] - y[e];
    a += d * d;
    b += y[e] * y[e];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    out[1] = s2[0];
  }
}

void gemm_cm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  gemm_launch(colmajor<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc), st);
}

struct Buf {
  double* p = nullptr;
  size_t cap = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  double* get(size_t n) {
    if (n > cap) {
      if (p) TTF_HIP(hipFree(p));
      p = nullptr;
      TTF_HIP(hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)));
      cap = n;
    }
    return p;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" int ttf_tt_mvk4(int d, const int* n, const int* m, const int* rx, const int* ra, const double* crA,
                           const double* crX, const double* crY0, int* ry, double eps, int rmax, int kickrank,
                           int nswp, int verb, double* result, int64_t result_cap, int64_t* result_size,
                           ttf_mvk4_stats* stats) {
  return guarded([&] {
    TTF_REQUIRE(d >= 2, "ttf_tt_mvk4: d=%d (need d >= 2)", d);
    TTF_REQUIRE(rx[0] == 1 && rx[d] == 1 && ry[0] == 1 && ry[d] == 1 && ra[0] == 1 && ra[d] == 1,
                "ttf_tt_mvk4: boundary ranks must be 1");
    TTF_REQUIRE(kickrank >= 0, "ttf_tt_mvk4: kickrank=%d", kickrank);
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const double eps2 = eps / std::sqrt((double)d);   // :115
    const int nswp0 = nswp > 0 ? nswp : 20;
    // capacities: ranks of y up to R, of x and A as given
    int R = rmax;
    for (int k = 0; k <= d; ++k) R = std::max(R, ry[k]);
    int nmax = 1, rxmax = 1, ramax = 1;
    for (int k = 0; k < d; ++k) nmax = std::max({nmax, n[k], m[k]});
    for (int k = 0; k <= d; ++k) {
      rxmax = std::max(rxmax, rx[k]);
      ramax = std::max(ramax, ra[k]);
    }
    std::vector<int64_t> pa(d + 1, 0), px(d + 1, 0);
    for (int k = 0; k < d; ++k) {
      pa[k + 1] = pa[k] + (int64_t)ra[k] * n[k] * m[k] * ra[k + 1];
      px[k + 1] = px[k] + (int64_t)rx[k] * m[k] * rx[k + 1];
    }
    Buf bA, bX;
    double* dA = bA.get(pa[d]);
    double* dX = bX.get(px[d]);
    TTF_HIP(hipMemcpyAsync(dA, crA, sizeof(double) * pa[d], hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dX, crX, sizeof(double) * px[d], hipMemcpyHostToDevice, st));
    const int kmax = kickrank;
    const size_t core_cap = (size_t)R * nmax * (R + kmax);
    std::vector<Buf> cr(d), phi(d + 1);
    for (int k = 0; k < d; ++k) cr[k].get(core_cap);
    for (int k = 0; k <= d; ++k) phi[k].get((size_t)R * rxmax * ramax);
    {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)ry[k] * n[k] * ry[k + 1];
        TTF_HIP(hipMemcpyAsync(cr[k].p, crY0 + off, sizeof(double) * sz, hipMemcpyHostToDevice, st));
        off += sz;
      }
    }
    const double one = 1.0;
    TTF_HIP(hipMemcpyAsync(phi[0].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phi[d].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    Buf bcur, bU, bR, bR2, bQ, bS, bz, bdu, btmp, bnrm;
    const size_t wide = (size_t)R * nmax * (R + kmax);
    double* cur = bcur.get(core_cap);
    double* U = bU.get(wide);
    double* Rm = bR.get(wide);
    double* R2 = bR2.get(wide);
    double* Q = bQ.get(wide);
    double* Sv = bS.get((size_t)R * nmax + 64);
    double* z = bz.get((size_t)rxmax * ramax * nmax * R);
    double* du = bdu.get((size_t)nmax * R * std::max(kmax, 1));
    double* tmp = btmp.get(core_cap);
    double* nrm = bnrm.get(4);
    auto shape = [&](int k) { return Shape{rx[k], m[k], rx[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1]}; };
    double flops = 0.0;
    // ---- left-orthogonalisation of y0 (:197-222)
    for (int k = 0; k < d - 1; ++k) {
      const int mrow = ry[k] * n[k], ncol = ry[k + 1];
      const int rnew = std::min(mrow, ncol);
      orth_factor(mrow, ncol, cr[k].p, Q, Rm, Sv, st);
      TTF_HIP(hipMemcpyAsync(cr[k].p, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
      const int ncol2 = n[k + 1] * ry[k + 2];
      gemm_cm('N', 'N', rnew, ncol2, ncol, 1.0, Rm, rnew, cr[k + 1].p, ncol, 0.0, tmp, rnew, st);
      TTF_HIP(hipMemcpyAsync(cr[k + 1].p, tmp, sizeof(double) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice, st));
      ry[k + 1] = rnew;
      phi_left_chain<double>(shape(k), phi[k].p, dA + pa[k], dX + px[k], cr[k].p, phi[k + 1].p, st);
    }
    // ---- sweeps (:228-461)
    int i = d - 1, dir = -1, swp = 1, last_swp = 0;
    double err_max = 0.0, last_err = 0.0;
    bool converged = false;
    std::vector<double> hs, h2(2);
    while (swp <= nswp0) {
      const Shape s = shape(i);
      const int64_t N = (int64_t)ry[i] * n[i] * ry[i + 1];
      bfun3_chain<double>(s, 1, phi[i].p, dA + pa[i], phi[i + 1].p, dX + px[i], (int64_t)rx[i] * m[i] * rx[i + 1],
                          cur, N, st);
      flops += bfun3_flops(s);
      hipLaunchKernelGGL(diff_norms_kernel, dim3(1), dim3(256), 0, st, N, cr[i].p, cur, nrm);
      TTF_HIP(hipGetLastError());
      read_back(h2.data(), nrm, sizeof(double) * 2, st);
      const double err = h2[1] > 0.0 ? std::sqrt(h2[0] / h2[1]) : 0.0;   // :253-255
      err_max = std::max(err_max, err);
      if (verb >= 2)
        std::fprintf(stderr, "mvk4: iteration:[%d,%d] err:%.3e ry(i):%d err_max:%.3e\n", swp, i + 1, err, ry[i],
                     err_max);
      if (dir > 0 && i < d - 1) {   // :269-341
        const int mrow = ry[i] * n[i], ncol = ry[i + 1];
        int rnew;
        if (kickrank > 0) {
          int kpos = 0;
          svd_vectors(mrow, ncol, cur, mrow, true, Sv, U, mrow, hs, &kpos, st);
          const int k = std::min(mrow, ncol);
          rnew = my_chop3(hs, k, eps2);
          complete_columns(mrow, kpos, rnew, U, mrow, st);
          gemm_cm('T', 'N', rnew, ncol, mrow, 1.0, U, mrow, cur, mrow, 0.0, Rm, rnew, st);   // S V^T
          TTF_HIP(hipMemcpyAsync(cr[i].p, U, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
        } else {
          rnew = std::min(mrow, ncol);
          orth_factor(mrow, ncol, cur, Q, Rm, Sv, st);
          TTF_HIP(hipMemcpyAsync(cr[i].p, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
        }
        TTF_REQUIRE(rnew <= R, "ttf_tt_mvk4: rank %d exceeds the capacity %d (raise rmax)", rnew, R);
        const int ncol2 = n[i + 1] * ry[i + 2];
        gemm_cm('N', 'N', rnew, ncol2, ncol, 1.0, Rm, rnew, cr[i + 1].p, ncol, 0.0, tmp, rnew, st);
        TTF_HIP(hipMemcpyAsync(cr[i + 1].p, tmp, sizeof(double) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice,
                               st));
        ry[i + 1] = rnew;
        phi_left_chain<double>(shape(i), phi[i].p, dA + pa[i], dX + px[i], cr[i].p, phi[i + 1].p, st);
      }
      if (dir < 0 && i > 0) {   // :349-426
        const int mrow = ry[i], ncol = n[i] * ry[i + 1];
        int rnew;
        const double* Lt;   // rnew x ry_i: core(i-1) <- core(i-1) Lt^T
        if (kickrank > 0) {
          int kpos = 0;
          svd_vectors(mrow, ncol, cur, mrow, false, Sv, U, ncol, hs, &kpos, st);   // V (ncol x k)
          const int k = std::min(mrow, ncol);
          const int nn = my_chop3(hs, k, eps2);
          complete_columns(ncol, kpos, nn, U, ncol, st);
          // R (nn x ry_i) = V^T cur^T = (cur V)^T
          gemm_cm('T', 'T', nn, mrow, ncol, 1.0, U, ncol, cur, mrow, 0.0, R2, nn, st);
          // enrichment from the half product A_k Phi_k+1 x_k, (rx ra) x (n ry2)
          bfun3_right_half(s, dA + pa[i], phi[i + 1].p, dX + px[i], z, st);
          const int mm = duchol('N', rx[i] * ra[i], ncol, z, std::min(kickrank, ncol), du, nullptr, st);
          // W = [V du] -> Q R3; R = R3(:, 0:nn) R2
          TTF_HIP(hipMemcpyAsync(tmp, U, sizeof(double) * (size_t)ncol * nn, hipMemcpyDeviceToDevice, st));
          TTF_HIP(hipMemcpyAsync(tmp + (size_t)ncol * nn, du, sizeof(double) * (size_t)ncol * mm,
                                 hipMemcpyDeviceToDevice, st));
          rnew = std::min(ncol, nn + mm);
          TTF_REQUIRE(rnew <= R, "ttf_tt_mvk4: rank %d exceeds the capacity %d (raise rmax)", rnew, R);
          orth_factor(ncol, nn + mm, tmp, Q, Rm, Sv, st);
          gemm_cm('N', 'N', rnew, mrow, nn, 1.0, Rm, rnew, R2, nn, 0.0, U, rnew, st);
          Lt = U;
        } else {
          rnew = std::min(mrow, ncol);
          transpose(mrow, ncol, cur, mrow, tmp, ncol, st);
          orth_factor(ncol, mrow, tmp, Q, Rm, Sv, st);
          Lt = Rm;
        }
        TTF_REQUIRE(rnew <= R, "ttf_tt_mvk4: rank %d exceeds the capacity %d (raise rmax)", rnew, R);
        transpose(ncol, rnew, Q, ncol, cr[i].p, rnew, st);   // core i = Q^T (rnew x n ry2)
        const int mp = ry[i - 1] * n[i - 1];
        gemm_cm('N', 'T', mp, rnew, mrow, 1.0, cr[i - 1].p, mp, Lt, rnew, 0.0, tmp, mp, st);
        TTF_HIP(hipMemcpyAsync(cr[i - 1].p, tmp, sizeof(double) * (size_t)mp * rnew, hipMemcpyDeviceToDevice, st));
        ry[i] = rnew;
        phi_right_chain<double>(shape(i), phi[i + 1].p, dA + pa[i], dX + px[i], cr[i].p, phi[i].p, st);
      }
      bool stop = false;
      if (dir > 0 && i == d - 1) {   // :428-447
        TTF_HIP(hipMemcpyAsync(cr[i].p, cur, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        dir = -1;
        i = d;
        last_swp = swp;
        ++swp;
        last_err = err_max;
        if (verb >= 1) std::fprintf(stderr, "mvk4: sweep %d err_max %.3e\n", last_swp, err_max);
        if (err_max < eps) {
          converged = true;
          stop = true;
        }
        err_max = 0.0;
      } else if (dir < 0 && i == 0) {   // :448-461
        TTF_HIP(hipMemcpyAsync(cr[i].p, cur, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        dir = 1;
        i = -1;
      }
      if (stop) break;
      i += dir;
    }
    int64_t need = 0;
    for (int k = 0; k < d; ++k) need += (int64_t)ry[k] * n[k] * ry[k + 1];
    if (result_size) *result_size = need;
    TTF_REQUIRE(result == nullptr || result_cap >= need, "ttf_tt_mvk4: result capacity %lld < %lld",
                (long long)result_cap, (long long)need);
    if (result) {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)ry[k] * n[k] * ry[k + 1];
        TTF_HIP(hipMemcpyAsync(result + off, cr[k].p, sizeof(double) * sz, hipMemcpyDeviceToHost, st));
        off += sz;
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
    if (stats) {
      stats->sweeps = last_swp;
      stats->last_err_max = last_err;
      stats->converged = converged ? 1 : 0;
      int mx = 0;
      for (int k = 0; k <= d; ++k) mx = std::max(mx, ry[k]);
      stats->max_rank = mx;
      stats->flops = flops;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

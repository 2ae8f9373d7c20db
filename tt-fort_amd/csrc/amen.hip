// tt_amen_solve on the device: the AMEn sweep of tt_adapt_als.f90:513-1132
// for A x = y with A, y and the initial x in TT format.
//
// All cores, operator cores and interface matrices (phiA for A, phiy for y)
// stay in HBM; the host steers ranks, tolerances and the sweep logic.
//   * initial left-to-right orthogonalisation with phi updates (:604-645);
//   * sweeps starting right-to-left: local rhs = phiy_L Y_k phiy_R, local
//     residual, and when it is above eps2 / 2 a right-preconditioned GMRES
//     correction (:654-797);
//   * SVD of the updated core and the residual-based rank choice: the
//     smallest rank whose truncated core still has a local residual within
//     max(2 res_new, eps2) (:807-833, :928-949) - the candidate ranks are
//     tested in blocks of up to 8 by one block matvec, scanning from the top
//     exactly as the reference's serial loop does;
//   * the kick: kickrank directions from the pivoted incomplete Cholesky of
//     [phiy Y | A x] projected on the open side, appended and
//     re-orthogonalised (:835-877, :954-999);
//   * stop after the sweep following the first one whose largest local
//     residual is below eps, which runs without kick (:1056-1101).
#include "amen.h"
#include "mlp.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

namespace ttf {

namespace {

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double s() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// A device buffer that grows (contents are not preserved on growth).
struct DevBuf {
  double* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  double* get(size_t n) {
    if (n > cap) {
      if (p) TTF_HIP(hipFree(p));
      p = nullptr;
      TTF_HIP(hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)));
      cap = n;
    }
    return p;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// out[c] = || Y(:, c) - rhs ||^2 for c < nc, one work-group per column
__global__ __launch_bounds__(256) void resid_norms_kernel(int64_t N, const double* __restrict__ Y, int64_t ldy,
                                                         const double* __restrict__ rhs, double* __restrict__ out) {
  __shared__ double red[256];
  const double* y = Y + ldy * blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
    const double v = y[i] - rhs[i];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void axpy_kernel(int64_t n, double a, const double* __restrict__ x, double* y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] += a * x[e];
}

// y = x - y
__global__ void rsub_kernel(int64_t n, const double* __restrict__ x, double* y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] = x[e] - y[e];
}

int grid_for(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 16384); }

double dev_norm(int64_t N, const double* v, double* scratch, hipStream_t st) {
  tsgram((int)N, 1, 1, v, N, v, N, scratch, 1, st);
  double h = 0.0;
  read_back(&h, scratch, sizeof(double), st);
  return std::sqrt(std::max(h, 0.0));
}

void gemm_nn(int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
             double beta, double* C, int64_t ldc, hipStream_t st, char ta = 'N', char tb = 'N') {
  if (M <= 0 || N <= 0) return;
  GemmOp<double> g = colmajor<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  gemm_launch(g, st);
}

}  // namespace

// Stopping test of the residual-based truncation (tt_adapt_als.f90:829-831,
// :945-947): a candidate rank ends the downward scan when its relative local
// residual exceeds both twice the solve's residual and eps2.
static inline bool residual_rank_stop(double e, double res_new, double eps2) {
  return e > res_new * 2.0 && e > eps2;
}

// The scan itself (tt_adapt_als.f90:819-833 forward, :937-950 backward):
// candidates k-1, k-2, ..., 1 from the top; the rank is one more than the
// first candidate that stops the scan, and 1 when none does (the DO loop
// ends with nn = 0).  eval(c0, nc, e) fills e[t] with the relative residual
// of candidate c0 - t for t < nc; candidates are asked for in batches of up
// to `batch`, in scan order, so the decision is the serial loop's.
template <class Eval>
static int residual_rank_scan(int k, int batch, double res_new, double eps2, Eval&& eval) {
  std::vector<double> e(std::max(batch, 1));
  for (int c0 = k - 1; c0 >= 1; c0 -= batch) {
    const int nc = std::min(batch, c0);
    eval(c0, nc, e.data());
    for (int t = 0; t < nc; ++t)
      if (residual_rank_stop(e[t], res_new, eps2)) return c0 - t + 1;
  }
  return 1;
}

constexpr int kCand = 8;   // candidate ranks per block matvec

// Residual-based truncation rank of the local solution on the device.
// dir > 0: candidate c is U(:, 0:c) Rm(0:c, :) with U (mrow x k) the left
// singular vectors and Rm = S V^T (k x ncol, leading dimension k).
// dir < 0: candidate c is Rm(:, 0:c) U(:, 0:c)^T with Rm = U S (mrow x k)
// and U the right singular vectors (ncol x k).
// cand and candy hold kCand local vectors each, scr at least kCand doubles.
// errs (host, k entries, optional) receives the residual of every candidate
// evaluated and -1 for the others.
static int residual_rank_search(const LocalOp& op, int dir, int mrow, int ncol, int k, const double* U,
                                const double* Rm, const double* rhs, double norm_rhs, double res_new, double eps2,
                                double* cand, double* candy, double* scr, long* matvecs, double* flops,
                                double* errs, hipStream_t st) {
  const int64_t N = (int64_t)mrow * ncol;
  if (errs)
    for (int c = 0; c < k; ++c) errs[c] = -1.0;
  return residual_rank_scan(k, kCand, res_new, eps2, [&](int c0, int nc, double* e) {
    for (int t = 0; t < nc; ++t) {
      if (dir > 0)
        gemm_nn(mrow, ncol, c0 - t, 1.0, U, mrow, Rm, k, 0.0, cand + (size_t)N * t, mrow, st);
      else
        gemm_nn(mrow, ncol, c0 - t, 1.0, Rm, mrow, U, ncol, 0.0, cand + (size_t)N * t, mrow, st, 'N', 'T');
    }
    op.apply(nc, cand, N, candy, N, st);
    if (matvecs) *matvecs += nc;
    if (flops) *flops += bfun3_flops(op.s) * nc;
    hipLaunchKernelGGL(resid_norms_kernel, dim3(nc), dim3(256), 0, st, N, candy, N, rhs, scr);
    std::vector<double> hr(nc);
    read_back(hr.data(), scr, sizeof(double) * nc, st);
    for (int t = 0; t < nc; ++t) {
      e[t] = std::sqrt(std::max(hr[t], 0.0)) / norm_rhs;
      if (errs) errs[c0 - t] = e[t];
    }
  });
}

}  // namespace ttf

using namespace ttf;

// Test hooks: the integer decisions of the truncations, callable alone.
extern "C" int ttf_residual_rank_rule(int k, const double* errs, double res_new, double eps2) {
  if (k < 1 || !errs) return -1;
  return residual_rank_scan(k, kCand, res_new, eps2, [&](int c0, int nc, double* e) {
    for (int t = 0; t < nc; ++t) e[t] = errs[c0 - t];
  });
}

extern "C" int ttf_amen_truncation_rank(int dir, int rx1, int n, int rx2, int ra1, int ra2, const double* phi1,
                                        const double* A, const double* phi2, const double* x, const double* rhs,
                                        double res_new, double eps2, int* rank, double* errs) {
  return guarded([&] {
    TTF_REQUIRE(dir == 1 || dir == -1, "ttf_amen_truncation_rank: dir=%d", dir);
    TTF_REQUIRE(rx1 >= 1 && n >= 1 && rx2 >= 1 && ra1 >= 1 && ra2 >= 1 && rank,
                "ttf_amen_truncation_rank: bad arguments");
    hipStream_t st = stream();
    const int mrow = dir > 0 ? rx1 * n : rx1, ncol = dir > 0 ? rx2 : n * rx2;
    const int k = std::min(mrow, ncol);
    const int64_t N = (int64_t)rx1 * n * rx2;
    DevBuf bU, bR, bS, bc, bcy, bsc, bx;
    double* U = bU.get((size_t)std::max(mrow, ncol) * k);
    double* Rm = bR.get((size_t)std::max(mrow, ncol) * k);
    double* Sv = bS.get((size_t)k + 64);
    double* cand = bc.get((size_t)N * kCand);
    double* candy = bcy.get((size_t)N * kCand);
    double* scr = bsc.get(256);
    double* xc = bx.get((size_t)N);
    TTF_HIP(hipMemcpyAsync(xc, x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    std::vector<double> hs;
    int kpos = 0;
    if (dir > 0) {
      svd_vectors(mrow, ncol, xc, mrow, true, Sv, U, mrow, hs, &kpos, st);
      gemm_nn(k, ncol, mrow, 1.0, U, mrow, xc, mrow, 0.0, Rm, k, st, 'T', 'N');
    } else {
      svd_vectors(mrow, ncol, xc, mrow, false, Sv, U, ncol, hs, &kpos, st);
      gemm_nn(mrow, k, ncol, 1.0, xc, mrow, U, ncol, 0.0, Rm, mrow, st);
    }
    const double norm_rhs = dev_norm(N, rhs, scr, st);
    TTF_REQUIRE(norm_rhs > 0.0, "ttf_amen_truncation_rank: zero right-hand side");
    LocalOp op{Shape{rx1, n, rx2, rx1, n, rx2, ra1, ra2}, phi1, A, phi2};
    *rank = residual_rank_search(op, dir, mrow, ncol, k, U, Rm, rhs, norm_rhs, res_new, eps2, cand, candy, scr,
                                 nullptr, nullptr, errs, st);
    TTF_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int ttf_tt_amen_solve(int d, const int* n, const int* m, const int* ry, const int* ra, const double* crA,
                                 const double* crY, const double* crX0, int* rx, double eps, int kickrank, int nswp,
                                 int verb, char prec, int nrestart, int niters, int rmax, double* result,
                                 int64_t result_cap, int64_t* result_size, ttf_amen_stats* stats) {
  return guarded([&] {
    TTF_REQUIRE(d >= 2, "ttf_tt_amen_solve: d=%d (need d >= 2)", d);
    for (int k = 0; k < d; ++k) TTF_REQUIRE(n[k] == m[k], "ttf_tt_amen_solve: operator must be square");
    TTF_REQUIRE(rx[0] == 1 && rx[d] == 1 && ry[0] == 1 && ry[d] == 1 && ra[0] == 1 && ra[d] == 1,
                "ttf_tt_amen_solve: boundary ranks must be 1");
    TTF_REQUIRE(kickrank >= -1, "ttf_tt_amen_solve: kickrank=%d", kickrank);
    TTF_REQUIRE(nrestart >= 1 && niters >= 1, "ttf_tt_amen_solve: nrestart=%d niters=%d", nrestart, niters);
    const char ptype = norm_ptype(prec);
    TTF_REQUIRE(ptype == 'n' || ptype == 'c' || ptype == 'l' || ptype == 'r' || ptype == 'm',
                "ttf_tt_amen_solve: prec '%c'", prec);
    TTF_REQUIRE(ptype != 'm' || mlp_ready(), "ttf_tt_amen_solve: prec 'm' needs the network parameters set first");
    Timer total;
    hipStream_t st = stream();
    const double eps2 = eps / std::sqrt((double)d);   // :576
    int kick = kickrank;
    // ---- capacities
    int nmax = 1, ramax = 1, rymax = 1;
    for (int k = 0; k < d; ++k) nmax = std::max(nmax, n[k]);
    for (int k = 0; k <= d; ++k) { ramax = std::max(ramax, ra[k]); rymax = std::max(rymax, ry[k]); }
    int R = rmax > 0 ? rmax : 256;
    for (int k = 0; k <= d; ++k) R = std::max(R, rx[k]);
    // every buffer below is sized for ranks up to R: truncation and kick
    // keep each new rank within rcap (rmax when given, else the capacity)
    const int rcap = rmax > 0 ? std::min(rmax, R) : R;
    const int kmax = std::max(kickrank, 0);
    const size_t Ncap = (size_t)R * nmax * R;
    std::vector<int64_t> pa(d + 1, 0), py(d + 1, 0);
    for (int k = 0; k < d; ++k) {
      pa[k + 1] = pa[k] + (int64_t)ra[k] * n[k] * m[k] * ra[k + 1];
      py[k + 1] = py[k] + (int64_t)ry[k] * n[k] * ry[k + 1];
    }
    DevBuf bA, bY;
    double* dA = bA.get(pa[d]);
    double* dY = bY.get(py[d]);
    TTF_HIP(hipMemcpyAsync(dA, crA, sizeof(double) * pa[d], hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dY, crY, sizeof(double) * py[d], hipMemcpyHostToDevice, st));
    std::vector<DevBuf> core(d), phiA(d + 1), phiy(d + 1);
    for (int k = 0; k < d; ++k) core[k].get((size_t)R * nmax * (R + kmax));
    for (int k = 0; k <= d; ++k) {
      phiA[k].get((size_t)R * ramax * R);
      phiy[k].get((size_t)R * rymax);
    }
    {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)rx[k] * n[k] * rx[k + 1];
        TTF_HIP(hipMemcpyAsync(core[k].p, crX0 + off, sizeof(double) * sz, hipMemcpyHostToDevice, st));
        off += sz;
      }
    }
    const double one = 1.0;
    TTF_HIP(hipMemcpyAsync(phiA[0].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phiA[d].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phiy[0].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phiy[d].p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    // work buffers
    DevBuf bwork, brhs, bres, bsol, bdx, bkick, bkick2, bjacs, bU, bR, bR2, bQ, bS, bcand, bcy, btmp, bdu, bsc;
    const size_t kick_cap = ((size_t)R * ramax + rymax) * nmax * R;
    double* work = bwork.get(std::max(Ncap, (size_t)R * nmax * rymax));
    double* rhs = brhs.get(Ncap);
    double* res = bres.get(Ncap);
    double* sol = bsol.get(Ncap);
    double* dx = bdx.get(Ncap);
    double* kblk = bkick.get(kick_cap);
    double* kblk2 = bkick2.get(2 * kick_cap);   // A phi2 x rows, then the assembled block
    double* U = bU.get((size_t)R * nmax * (R + kmax));
    double* Rm = bR.get((size_t)(R + kmax) * R * nmax);
    double* R2 = bR2.get((size_t)(R + kmax) * R * nmax);
    double* Q = bQ.get((size_t)R * nmax * (R + kmax));
    double* Sv = bS.get((size_t)R * nmax + 64);
    double* cand = bcand.get(Ncap * kCand);
    double* candy = bcy.get(Ncap * kCand);
    double* tmp = btmp.get(Ncap + (size_t)R * nmax * (R + kmax));
    double* du = bdu.get((size_t)R * nmax * std::max(kmax, 1));
    double* scr = bsc.get(256);
    std::vector<double> hs;
    double t_solve = 0, t_trunc = 0, t_phi = 0;
    long mv_total = 0;
    double flops_total = 0.0;   // algorithmic flops of those local matvecs
    long mlp_solves = 0, mlp_fallbacks = 0;
    // ---- initial QR, left to right (:604-645)
    for (int i = 0; i < d - 1; ++i) {
      const int mrow = rx[i] * n[i], ncol = rx[i + 1];
      const int rnew = std::min(mrow, ncol);
      orth_factor(mrow, ncol, core[i].p, Q, Rm, Sv, st);
      TTF_HIP(hipMemcpyAsync(core[i].p, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
      const int ncol2 = n[i + 1] * rx[i + 2];
      gemm_nn(rnew, ncol2, ncol, 1.0, Rm, rnew, core[i + 1].p, ncol, 0.0, tmp, rnew, st);
      TTF_HIP(hipMemcpyAsync(core[i + 1].p, tmp, sizeof(double) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice, st));
      rx[i + 1] = rnew;
      Shape s{rx[i], m[i], rx[i + 1], rx[i], n[i], rx[i + 1], ra[i], ra[i + 1]};
      phi_left_chain<double>(s, phiA[i].p, dA + pa[i], core[i].p, core[i].p, phiA[i + 1].p, st);
      phi2_left_chain<double>(ry[i], ry[i + 1], rx[i], n[i], rx[i + 1], phiy[i].p, dY + py[i], core[i].p,
                              phiy[i + 1].p, st);
    }
    // ---- sweeps
    double err_max = 0.0, res_max = 0.0, last_res_max = 0.0, last_err_max = 0.0;
    int i = d - 1, dir = -1, swp = 1, sweeps_done = 0;
    while (swp <= nswp) {
      const int r1 = rx[i], nn_i = n[i], r2 = rx[i + 1];
      const int64_t N = (int64_t)r1 * nn_i * r2;
      double* x = core[i].p;
      const double* Yk = dY + py[i];
      // local rhs (:657-695); the phiy Y part is kept for the kick
      if (dir < 0) {
        gemm_nn(ry[i] * nn_i, r2, ry[i + 1], 1.0, Yk, (int64_t)ry[i] * nn_i, phiy[i + 1].p, ry[i + 1], 0.0, work,
                (int64_t)ry[i] * nn_i, st);
        TTF_HIP(hipMemcpyAsync(kblk, work, sizeof(double) * (size_t)ry[i] * nn_i * r2, hipMemcpyDeviceToDevice, st));
        gemm_nn(r1, nn_i * r2, ry[i], 1.0, phiy[i].p, r1, work, ry[i], 0.0, rhs, r1, st);
      } else {
        gemm_nn(r1, nn_i * ry[i + 1], ry[i], 1.0, phiy[i].p, r1, Yk, ry[i], 0.0, work, r1, st);
        TTF_HIP(hipMemcpyAsync(kblk, work, sizeof(double) * (size_t)r1 * nn_i * ry[i + 1], hipMemcpyDeviceToDevice,
                               st));
        gemm_nn(r1 * nn_i, r2, ry[i + 1], 1.0, work, (int64_t)r1 * nn_i, phiy[i + 1].p, ry[i + 1], 0.0, rhs,
                (int64_t)r1 * nn_i, st);
      }
      LocalOp op{Shape{r1, m[i], r2, r1, nn_i, r2, ra[i], ra[i + 1]}, phiA[i].p, dA + pa[i], phiA[i + 1].p};
      Timer ts;
      // residual of the current core (:733-740)
      op.apply(1, x, N, res, N, st);
      ++mv_total;
      flops_total += bfun3_flops(op.s);
      hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, rhs, res);
      const double norm_rhs = dev_norm(N, rhs, scr, st);
      TTF_REQUIRE(norm_rhs > 0.0, "ttf_tt_amen_solve: zero local right-hand side at core %d", i + 1);
      double err = dev_norm(N, res, scr, st) / norm_rhs;
      res_max = std::max(res_max, err);
      double res_new;
      const double gtol = eps2 / (err * 2.0);
      GmresStats gs;
      if (gtol < 1.0) {
        const int nrs = (int)std::min<int64_t>(nrestart, N);
        MixedOp mixop;
        const MixedOp* mix = nullptr;
        if (local_matvec_bits() == 32) {
          mixop = make_mixed(op, st);
          mix = &mixop;
        }
        if (ptype == 'm') {
          // learned block preconditioner in flexible GMRES; plain GMRES takes
          // over from its iterate when it misses the tolerance
          const double* p1 = phiA[i].p;
          const double* Ai = dA + pa[i];
          const double* p2 = phiA[i + 1].p;
          const int ra1 = ra[i], ra2 = ra[i + 1];
          PrecFn M = [&, p1, Ai, p2, ra1, ra2, r1, nn_i, r2](const double* v, double* z) {
            mlp_precond(r1, nn_i, r2, ra1, ra2, p1, Ai, p2, v, z, st);
          };
          gs = gmres_flex(op, M, res, dx, nrs, gtol, niters, st, mix);
          mv_total += gs.matvecs;
          flops_total += bfun3_flops(op.s) * gs.matvecs;
          ++mlp_solves;
          if (!(gs.relres < gtol)) {
            ++mlp_fallbacks;
            if (!std::isfinite(gs.relres)) TTF_HIP(hipMemsetAsync(dx, 0, sizeof(double) * N, st));
            // remaining residual r2 = res - B dx (in candy), solved by plain
            // GMRES to the same absolute target (correction in cand)
            op.apply(1, dx, N, candy, N, st);
            hipLaunchKernelGGL(rsub_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, res, candy);
            const double n0 = dev_norm(N, res, scr, st), n2 = dev_norm(N, candy, scr, st);
            int extra = 1;
            if (n2 > 0.0) {
              GmresStats g2 = gmres_solve(op, JacPrec{}, candy, cand, nrs, std::min(1.0, gtol * n0 / n2), niters, st,
                                          mix);
              hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, 1.0, cand, dx);
              extra += g2.matvecs;
              gs.relres = g2.relres * n2 / n0;
            }
            mv_total += extra;
            flops_total += bfun3_flops(op.s) * extra;
            gs.matvecs += extra;
          }
        } else {
          JacPrec P;
          if (ptype != 'n') {
            double* jacs = bjacs.get(jac_size(ptype, r1, nn_i, r2));
            jac_gen(ptype, r1, nn_i, r2, ra[i], ra[i + 1], phiA[i].p, dA + pa[i], phiA[i + 1].p, jacs, st);
            P = JacPrec{ptype, r1, nn_i, r2, jacs, ra[i], ra[i + 1], phiA[i].p, dA + pa[i], phiA[i + 1].p};
          }
          gs = gmres_solve(op, P, res, sol, nrs, gtol, niters, st, mix);
          mv_total += gs.matvecs;
          flops_total += bfun3_flops(op.s) * gs.matvecs;
          if (ptype != 'n') jac_apply(ptype, r1, nn_i, r2, P.jacs, sol, dx, st);
          else TTF_HIP(hipMemcpyAsync(dx, sol, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, dx, x);
        op.apply(1, x, N, res, N, st);
        ++mv_total;
        flops_total += bfun3_flops(op.s);
        hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, rhs, res);
        res_new = dev_norm(N, res, scr, st) / norm_rhs;
        const double nx = dev_norm(N, x, scr, st);
        err = nx > 0.0 ? dev_norm(N, dx, scr, st) / nx : 0.0;
        err_max = std::max(err_max, err);
      } else {
        res_new = err;
        err = 0.0;
      }
      t_solve += ts.s();
      if (verb >= 2)
        std::fprintf(stderr, "amen: iteration:[%d,%d] err:%.3e rx(i):%d err_max:%.3e res_max:%.3e gmres:[%d,%d]\n",
                     swp, i + 1, err, r1, err_max, res_max, gs.it, gs.j);
      Timer tt;
      // ---- truncation and kick
      if (dir > 0 && i < d - 1) {
        const int mrow = r1 * nn_i, ncol = r2;
        int rnew = std::min(mrow, ncol);
        const double* R_out = nullptr;   // rnew x ncol factor for the next core
        if (kick > 0 || kick == -1) {
          int kpos = 0;
          svd_vectors(mrow, ncol, x, mrow, true, Sv, U, mrow, hs, &kpos, st);   // U (mrow x k)
          const int k = rnew;
          gemm_nn(k, ncol, mrow, 1.0, U, mrow, x, mrow, 0.0, Rm, k, st, 'T', 'N');   // S V^T
          // residual-based rank (:819-833): scan candidates from the top
          int nn = residual_rank_search(op, 1, mrow, ncol, k, U, Rm, rhs, norm_rhs, res_new, eps2, cand, candy,
                                        scr, &mv_total, &flops_total, nullptr, st);
          nn = std::min(nn, rcap);
          nn = std::min(nn, k);
          complete_columns(mrow, kpos, nn, U, mrow, st);
          if (kick > 0) {
            // kick block [phiy Y | Phi1 A x] of size mrow x (ry_{i+1} + r2 ra_{i+1})
            Shape s{r1, m[i], r2, r1, nn_i, r2, ra[i], ra[i + 1]};
            bfun3_left_half(s, phiA[i].p, dA + pa[i], x, kblk + (size_t)mrow * ry[i + 1], st);
            const int kcols = r2 * ra[i + 1] + ry[i + 1];
            int mm = duchol('T', kcols, mrow, kblk, std::min(kick, mrow), du, nullptr, st);
            mm = std::min(mm, std::max(rcap - nn, 0));
            // [U(:, 0:nn) du] -> Q R2, R = R2(:, 0:nn) (S V^T)(0:nn, :)
            TTF_HIP(hipMemcpyAsync(tmp, U, sizeof(double) * (size_t)mrow * nn, hipMemcpyDeviceToDevice, st));
            TTF_HIP(hipMemcpyAsync(tmp + (size_t)mrow * nn, du, sizeof(double) * (size_t)mrow * mm,
                                   hipMemcpyDeviceToDevice, st));
            rnew = std::min(mrow, nn + mm);
            orth_factor(mrow, nn + mm, tmp, Q, R2, Sv, st);
            gemm_nn(rnew, ncol, nn, 1.0, R2, rnew, Rm, k, 0.0, work, rnew, st);
            TTF_HIP(hipMemcpyAsync(x, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
            R_out = work;
          } else {
            rnew = nn;
            TTF_HIP(hipMemcpyAsync(x, U, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
            TTF_HIP(hipMemcpy2DAsync(work, sizeof(double) * rnew, Rm, sizeof(double) * k, sizeof(double) * rnew, ncol,
                                     hipMemcpyDeviceToDevice, st));
            R_out = work;
          }
        } else {
          orth_factor(mrow, ncol, x, Q, Rm, Sv, st);
          TTF_HIP(hipMemcpyAsync(x, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
          R_out = Rm;
        }
        TTF_REQUIRE(rnew <= R, "ttf_tt_amen_solve: rank %d exceeds the capacity %d", rnew, R);
        // next core <- R * core(i+1)
        const int ncol2 = n[i + 1] * rx[i + 2];
        gemm_nn(rnew, ncol2, ncol, 1.0, R_out, rnew, core[i + 1].p, ncol, 0.0, tmp, rnew, st);
        TTF_HIP(hipMemcpyAsync(core[i + 1].p, tmp, sizeof(double) * (size_t)rnew * ncol2, hipMemcpyDeviceToDevice,
                               st));
        rx[i + 1] = rnew;
        t_trunc += tt.s();
        Timer tp;
        Shape s{rx[i], m[i], rx[i + 1], rx[i], n[i], rx[i + 1], ra[i], ra[i + 1]};
        phi_left_chain<double>(s, phiA[i].p, dA + pa[i], core[i].p, core[i].p, phiA[i + 1].p, st);
        phi2_left_chain<double>(ry[i], ry[i + 1], rx[i], n[i], rx[i + 1], phiy[i].p, dY + py[i], core[i].p,
                                phiy[i + 1].p, st);
        TTF_HIP(hipStreamSynchronize(st));
        t_phi += tp.s();
      }
      if (dir < 0 && i > 0) {
        const int mrow = r1, ncol = nn_i * r2;
        int rnew = std::min(mrow, ncol);
        const double* Lt = nullptr;   // rnew x r1 factor: core(i-1) <- core(i-1) Lt^T
        if (kick > 0 || kick == -1) {
          int kpos = 0;
          svd_vectors(mrow, ncol, x, mrow, false, Sv, U, ncol, hs, &kpos, st);   // V (ncol x k)
          const int k = rnew;
          gemm_nn(mrow, k, ncol, 1.0, x, mrow, U, ncol, 0.0, Rm, mrow, st);      // U S (mrow x k)
          // residual-based rank (:937-950)
          int nn = residual_rank_search(op, -1, mrow, ncol, k, U, Rm, rhs, norm_rhs, res_new, eps2, cand, candy,
                                        scr, &mv_total, &flops_total, nullptr, st);
          nn = std::min(nn, rcap);
          nn = std::min(nn, k);
          complete_columns(ncol, kpos, nn, U, ncol, st);
          // (U S)(:, 0:nn)^T: nn x mrow
          transpose(mrow, nn, Rm, mrow, R2, nn, st);
          if (kick > 0) {
            // kick block rows [phiy Y ; A phi2 x] of size (ry_i + r1 ra_i) x ncol
            Shape s{r1, m[i], r2, r1, nn_i, r2, ra[i], ra[i + 1]};
            double* axr = kblk2 + (size_t)ry[i] * ncol;   // staged right after the Y rows' space
            bfun3_right_half(s, dA + pa[i], phiA[i + 1].p, x, axr, st);
            const int krows = ry[i] + r1 * ra[i];
            TTF_HIP(hipMemcpy2DAsync(kblk2 + (size_t)krows * ncol, sizeof(double) * krows, kblk, sizeof(double) * ry[i],
                                     sizeof(double) * ry[i], ncol, hipMemcpyDeviceToDevice, st));
            TTF_HIP(hipMemcpy2DAsync(kblk2 + (size_t)krows * ncol + ry[i], sizeof(double) * krows, axr,
                                     sizeof(double) * r1 * ra[i], sizeof(double) * r1 * ra[i], ncol,
                                     hipMemcpyDeviceToDevice, st));
            int mm = duchol('N', krows, ncol, kblk2 + (size_t)krows * ncol, std::min(kick, ncol), du, nullptr, st);
            mm = std::min(mm, std::max(rcap - nn, 0));
            TTF_HIP(hipMemcpyAsync(tmp, U, sizeof(double) * (size_t)ncol * nn, hipMemcpyDeviceToDevice, st));
            TTF_HIP(hipMemcpyAsync(tmp + (size_t)ncol * nn, du, sizeof(double) * (size_t)ncol * mm,
                                   hipMemcpyDeviceToDevice, st));
            rnew = std::min(ncol, nn + mm);
            orth_factor(ncol, nn + mm, tmp, Q, Rm, Sv, st);
            gemm_nn(rnew, mrow, nn, 1.0, Rm, rnew, R2, nn, 0.0, work, rnew, st);
            Lt = work;
          } else {
            rnew = nn;
            TTF_HIP(hipMemcpyAsync(Q, U, sizeof(double) * (size_t)ncol * rnew, hipMemcpyDeviceToDevice, st));
            Lt = R2;
          }
        } else {
          transpose(mrow, ncol, x, mrow, tmp, ncol, st);
          orth_factor(ncol, mrow, tmp, Q, Rm, Sv, st);
          Lt = Rm;
        }
        TTF_REQUIRE(rnew <= R, "ttf_tt_amen_solve: rank %d exceeds the capacity %d", rnew, R);
        // core i = Q^T (rnew x ncol); core(i-1) <- core(i-1) Lt^T
        transpose(ncol, rnew, Q, ncol, x, rnew, st);
        const int mp = rx[i - 1] * n[i - 1];
        gemm_nn(mp, rnew, r1, 1.0, core[i - 1].p, mp, Lt, rnew, 0.0, tmp, mp, st, 'N', 'T');
        TTF_HIP(hipMemcpyAsync(core[i - 1].p, tmp, sizeof(double) * (size_t)mp * rnew, hipMemcpyDeviceToDevice, st));
        rx[i] = rnew;
        t_trunc += tt.s();
        Timer tp;
        Shape s{rx[i], m[i], rx[i + 1], rx[i], n[i], rx[i + 1], ra[i], ra[i + 1]};
        phi_right_chain<double>(s, phiA[i + 1].p, dA + pa[i], core[i].p, core[i].p, phiA[i].p, st);
        phi2_right_chain<double>(ry[i], ry[i + 1], rx[i], n[i], rx[i + 1], phiy[i + 1].p, dY + py[i], core[i].p,
                                 phiy[i].p, st);
        TTF_HIP(hipStreamSynchronize(st));
        t_phi += tp.s();
      }
      // ---- sweep bookkeeping (:1056-1101)
      bool finish = false;
      if (dir > 0 && i == d - 1) {
        int mx = 0;
        for (int q = 0; q <= d; ++q) mx = std::max(mx, rx[q]);
        if (verb >= 1)
          std::fprintf(stderr, "amen: iteration:%d(%d) max(rx):%d err_max:%.3e res_max:%.3e\n", swp, dir, mx, err_max,
                       res_max);
        last_res_max = res_max;
        last_err_max = err_max;
        sweeps_done = swp;
        dir = -1;
        i = i + 1;
        if (kick < 0) finish = true;
        else {
          if (res_max < eps) kick = -1;
          ++swp;
          err_max = 0.0;
          res_max = 0.0;
        }
      } else if (dir < 0 && i == 0) {
        int mx = 0;
        for (int q = 0; q <= d; ++q) mx = std::max(mx, rx[q]);
        if (verb >= 1)
          std::fprintf(stderr, "amen: iteration:%d(%d) max(rx):%d err_max:%.3e res_max:%.3e\n", swp, dir, mx, err_max,
                       res_max);
        last_res_max = res_max;
        last_err_max = err_max;
        sweeps_done = swp;
        if (kick < 0) finish = true;
        else {
          if (res_max < eps) kick = -1;
          dir = 1;
          i = i - 1;
          err_max = 0.0;
          res_max = 0.0;
        }
      }
      if (finish) break;
      i += dir;
    }
    // ---- result
    int64_t total_sz = 0;
    for (int k = 0; k < d; ++k) total_sz += (int64_t)rx[k] * n[k] * rx[k + 1];
    if (result_size) *result_size = total_sz;
    if (result) {
      TTF_REQUIRE(result_cap >= total_sz, "ttf_tt_amen_solve: result buffer holds %lld doubles, need %lld",
                  (long long)result_cap, (long long)total_sz);
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)rx[k] * n[k] * rx[k + 1];
        TTF_HIP(hipMemcpyAsync(result + off, core[k].p, sizeof(double) * sz, hipMemcpyDeviceToHost, st));
        off += sz;
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
    if (stats) {
      stats->sweeps = sweeps_done;
      stats->total_matvecs = mv_total;
      stats->flops = flops_total;
      stats->learned_solves = mlp_solves;
      stats->learned_fallbacks = mlp_fallbacks;
      stats->last_res_max = last_res_max;
      stats->last_err_max = last_err_max;
      int mx = 0;
      for (int q = 0; q <= d; ++q) mx = std::max(mx, rx[q]);
      stats->max_rank = mx;
      stats->seconds = total.s();
      stats->solve_seconds = t_solve;
      stats->trunc_seconds = t_trunc;
      stats->phi_seconds = t_phi;
    }
  });
}

// ---------------------------------------------------------------- C ABI of the building blocks
extern "C" {

int ttf_djac_gen(char ptype, int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                 const double* phi2, double* jacs) {
  return guarded([&] { jac_gen(ptype, rx1, n, rx2, ra1, ra2, phi1, A, phi2, jacs, stream()); });
}

int64_t ttf_djac_size(char ptype, int rx1, int n, int rx2) { return (int64_t)jac_size(ptype, rx1, n, rx2); }

int ttf_djac_apply(char ptype, int rx1, int n, int rx2, const double* jacs, const double* x, double* y) {
  return guarded([&] { jac_apply(ptype, rx1, n, rx2, jacs, x, y, stream()); });
}

int ttf_dgmresr(const double* phi1, const double* A, const double* phi2, const double* rhs, int rx1, int n, int rx2,
                int ra1, int ra2, int nrestart, double tol, int niters, char ptype, const double* jacs, double* sol,
                int* iters, double* relres) {
  return guarded([&] {
    LocalOp op{Shape{rx1, n, rx2, rx1, n, rx2, ra1, ra2}, phi1, A, phi2};
    JacPrec P;
    const char p = norm_ptype(ptype);
    if (p != 'n') P = JacPrec{p, rx1, n, rx2, jacs};
    GmresStats gs = gmres_solve(op, P, rhs, sol, nrestart, tol, niters, stream());
    TTF_HIP(hipStreamSynchronize(stream()));
    if (iters) { iters[0] = gs.it; iters[1] = gs.j; }
    if (relres) *relres = gs.relres;
  });
}

int ttf_duchol(char trans, int m, int n, const double* dA, int numvec, double* du, double* lam, int* real_numvec) {
  return guarded([&] {
    int r = duchol(trans, m, n, dA, numvec, du, lam, stream());
    TTF_HIP(hipStreamSynchronize(stream()));
    if (real_numvec) *real_numvec = r;
  });
}

int ttf_dbfun3_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, const double* phi1,
                    const double* A, const double* x, double* y) {
  return guarded([&] {
    bfun3_left_half(Shape{rx1, m, rx2, ry1, n, ry2, ra1, ra2}, phi1, A, x, y, stream());
  });
}

int ttf_dbfun3_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, const double* A,
                     const double* phi2, const double* x, double* y) {
  return guarded([&] {
    bfun3_right_half(Shape{rx1, m, rx2, ry1, n, ry2, ra1, ra2}, A, phi2, x, y, stream());
  });
}

}  // extern "C"

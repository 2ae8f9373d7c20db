// dtt_amen_solve on the device: the AMEn linear solver of ttamen.f90:29-252
// (the typed interface of tt-fort), for A x = y with A, y and the initial x
// in TT format.  Unlike tt_amen_solve (tt_adapt_als.f90, amen.hip) it
// enriches from an auxiliary low-rank TT z that tracks the residual:
//   * each sweep first orthogonalises x (and z) right to left, updating the
//     right interfaces x'Ax, x'y, z'Ax, z'y; from the second sweep on each
//     core of z is first replaced by the projection Z'(A x - y) (:83-93);
//   * then left to right: local rhs x'y, local residual, and above
//     real_tol = tol / sqrt(d - 1) / 2 a correction from a dense solve
//     (local size below max_full_size) or right-preconditioned GMRES
//     (:122-173);
//   * the core is truncated (dtt_trunc, ttnodeop.f90:227-302): by local
//     residual (trunc_norm 1, the default) or by the Frobenius chop of its
//     singular values (trunc_norm 0);
//   * with kickrank > 0 the truncated core is enriched by the projection of
//     the residual onto z's right interface, z's core is updated to
//     Z'(A x - y), and both move right by QR (:186-219, dtt_enrich
//     ttnodeop.f90:143-182);
//   * stop when the largest local residual (trunc_norm 1) or the largest
//     relative correction (trunc_norm 0) of a sweep is below tol (:239-243).
// The reference's z starts from its Fortran random number generator
// (random(z), :75); here z's cores are N(0,1) from a caller-given seed.
#include "amen.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

namespace ttf {

// Frobenius-norm chop of singular values s (svd.f90:86-111, chop): the
// smallest rank r whose discarded tail satisfies
// sum_{k>r} s_k^2 < (tol ||s||)^2, scanning from the end as the reference
// does; rmax (when > 0) caps r first and its tail counts as discarded.
int chop_rank(const double* s, int n, double tol, int rmax) {
  if (n <= 0) return 0;
  int r = n;
  double er2 = 0.0;
  if (rmax > 0 && rmax < r) {
    for (int k = rmax; k < n; ++k) er2 += s[k] * s[k];
    r = rmax;
  }
  if (tol >= 0.0) {
    double nrm = 0.0;
    for (int k = 0; k < n; ++k) nrm += s[k] * s[k];
    const double bound = tol * tol * nrm;
    double er = er2 + s[r - 1] * s[r - 1];
    while (er < bound) {
      er2 = er;
      --r;
      if (r == 0) break;
      er += s[r - 1] * s[r - 1];
    }
  }
  return r;
}

}  // namespace ttf

using namespace ttf;

extern "C" int ttf_chop(int n, const double* s, double tol, int rmax) {
  if (n < 0 || (n > 0 && !s)) return -1;
  return chop_rank(s, n, tol, rmax);
}

namespace {

struct Buf {
  double* p = nullptr;
  size_t cap = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
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

void gemm_cm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  gemm_launch(colmajor<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc), st);
}

__global__ void axpy_kernel(int64_t n, double a, const double* __restrict__ x, double* y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] += a * x[e];
}

// C (m x (c1 + c2)) = [A (m x c1) | B (m x c2)]; D ((c1 + c2) x n) = [E (c1 x n) ; 0]
__global__ void widen_kernel(int64_t m, int c1, int c2, int64_t n, const double* __restrict__ E,
                             double* __restrict__ D) {
  const int rows = c1 + c2;
  const int64_t total = (int64_t)rows * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % rows);
    const int64_t j = e / rows;
    D[e] = i < c1 ? E[i + (int64_t)c1 * j] : 0.0;
  }
  (void)m;
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 16384); }

double dev_norm(int64_t N, const double* v, double* scratch, hipStream_t st) {
  tsgram((int)N, 1, 1, v, N, v, N, scratch, 1, st);
  double h = 0.0;
  read_back(&h, scratch, sizeof(double), st);
  return std::sqrt(std::max(h, 0.0));
}

// A TT of device cores with a rank capacity.
struct DevTT {
  int d = 0;
  std::vector<int> r;   // d + 1 ranks
  std::vector<int> n;
  std::vector<Buf> c;
};

}  // namespace

extern "C" int ttf_dtt_amen_solve(int d, const int* n, const int* ra, const double* crA, const int* ry,
                                  const double* crY, double tol, int* rx, const double* crX0, int kickrank,
                                  int nswp, char local_prec, int local_iters, int local_restart, int trunc_norm,
                                  int max_full_size, int verb, int rmax, unsigned seed, double* result,
                                  int64_t result_cap, int64_t* result_size, ttf_amen_stats* stats) {
  return guarded([&] {
    TTF_REQUIRE(d >= 2, "ttf_dtt_amen_solve: d=%d (need d >= 2)", d);
    TTF_REQUIRE(rx[0] == 1 && rx[d] == 1 && ry[0] == 1 && ry[d] == 1 && ra[0] == 1 && ra[d] == 1,
                "ttf_dtt_amen_solve: boundary ranks must be 1");
    TTF_REQUIRE(kickrank >= 0 && local_iters >= 1 && local_restart >= 1 && (trunc_norm == 0 || trunc_norm == 1),
                "ttf_dtt_amen_solve: kickrank=%d local_iters=%d local_restart=%d trunc_norm=%d", kickrank,
                local_iters, local_restart, trunc_norm);
    const char ptype = norm_ptype(local_prec);
    TTF_REQUIRE(ptype == 'n' || ptype == 'c' || ptype == 'l' || ptype == 'r', "ttf_dtt_amen_solve: prec '%c'",
                local_prec);
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = stream();
    const double resid_damp = 2.0;                                          // ttamen.f90:25
    const double real_tol = tol / std::sqrt((double)(d - 1)) / resid_damp;  // :62-63
    const int nswp0 = nswp > 0 ? nswp : 20;
    const int kick = kickrank;
    // ---- capacities
    int nmax = 1, ramax = 1, rymax = 1;
    for (int k = 0; k < d; ++k) nmax = std::max(nmax, n[k]);
    for (int k = 0; k <= d; ++k) {
      ramax = std::max(ramax, ra[k]);
      rymax = std::max(rymax, ry[k]);
    }
    int R = rmax > 0 ? rmax : 256;
    for (int k = 0; k <= d; ++k) R = std::max(R, rx[k]);
    const int Rw = R + kick;   // a rank before its enrichment is truncated back
    std::vector<int64_t> pa(d + 1, 0), py(d + 1, 0);
    for (int k = 0; k < d; ++k) {
      pa[k + 1] = pa[k] + (int64_t)ra[k] * n[k] * n[k] * ra[k + 1];
      py[k + 1] = py[k] + (int64_t)ry[k] * n[k] * ry[k + 1];
    }
    Buf bA, bY;
    double* dA = bA.get(pa[d]);
    double* dY = bY.get(py[d]);
    TTF_HIP(hipMemcpyAsync(dA, crA, sizeof(double) * pa[d], hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(dY, crY, sizeof(double) * py[d], hipMemcpyHostToDevice, st));
    DevTT x, z;
    x.d = z.d = d;
    x.r.assign(rx, rx + d + 1);
    x.n.assign(n, n + d);
    x.c.resize(d);
    for (int k = 0; k < d; ++k) x.c[k].get((size_t)Rw * nmax * Rw);
    {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)rx[k] * n[k] * rx[k + 1];
        TTF_HIP(hipMemcpyAsync(x.c[k].p, crX0 + off, sizeof(double) * sz, hipMemcpyHostToDevice, st));
        off += sz;
      }
    }
    if (kick > 0) {   // z: ranks (1, kick, ..., kick, 1), N(0,1) cores (:71-75)
      z.r.assign(d + 1, kick);
      z.r[0] = z.r[d] = 1;
      z.n = x.n;
      z.c.resize(d);
      std::mt19937_64 gen(seed);
      std::normal_distribution<double> nd(0.0, 1.0);
      for (int k = 0; k < d; ++k) {
        const size_t sz = (size_t)z.r[k] * n[k] * z.r[k + 1];
        z.c[k].get((size_t)std::max(kick, 1) * nmax * std::max(kick, 1));
        std::vector<double> h(sz);
        for (auto& v : h) v = nd(gen);
        TTF_HIP(hipMemcpyAsync(z.c[k].p, h.data(), sizeof(double) * sz, hipMemcpyHostToDevice, st));
        TTF_HIP(hipStreamSynchronize(st));
      }
    }
    const int kz = std::max(kick, 1);
    // interfaces: x'Ax (left (rx, rx, ra), right (rx, ra, rx)), x'y (left
    // (rx, ry), right (ry, rx)), z'Ax (left (rz, rx, ra), right (rx, ra, rz)),
    // z'y (left (rz, ry), right (ry, rz))
    std::vector<Buf> pxaL(d + 1), pxaR(d + 1), pxyL(d + 1), pxyR(d + 1), pzaL(d + 1), pzaR(d + 1), pzyL(d + 1),
        pzyR(d + 1);
    for (int k = 0; k <= d; ++k) {
      pxaL[k].get((size_t)Rw * Rw * ramax);
      pxaR[k].get((size_t)Rw * Rw * ramax);
      pxyL[k].get((size_t)Rw * rymax);
      pxyR[k].get((size_t)Rw * rymax);
      if (kick > 0) {
        pzaL[k].get((size_t)kz * Rw * ramax);
        pzaR[k].get((size_t)kz * Rw * ramax);
        pzyL[k].get((size_t)kz * rymax);
        pzyR[k].get((size_t)kz * rymax);
      }
    }
    const double one = 1.0;
    for (Buf* b : {&pxaL[0], &pxaR[d], &pxyL[0], &pxyR[d]})
      TTF_HIP(hipMemcpyAsync(b->p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    if (kick > 0)
      for (Buf* b : {&pzaL[0], &pzaR[d], &pzyL[0], &pzyR[d]})
        TTF_HIP(hipMemcpyAsync(b->p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    // work buffers
    const size_t Ncap = (size_t)Rw * nmax * Rw;
    Buf brhs, bres, bsol, bxn, bU, bR, bQ, bS, bt, bt2, bjac, bB, bcand, bcy, bsc;
    double* rhs = brhs.get(Ncap);
    double* res = bres.get(Ncap);
    double* sol = bsol.get(Ncap);
    double* xnew = bxn.get(Ncap);
    double* U = bU.get(Ncap);
    double* Rm = bR.get(Ncap);
    double* Q = bQ.get(Ncap);
    double* Sv = bS.get((size_t)Rw * nmax + 64);
    double* t1 = bt.get((size_t)std::max(Rw, rymax) * nmax * std::max(Rw, rymax));
    double* t2 = bt2.get(Ncap);
    double* scr = bsc.get(256);
    std::vector<double> hs;
    double t_solve = 0, t_trunc = 0, t_phi = 0;
    long mv_total = 0;
    double flops_total = 0.0;
    auto secs = [](std::chrono::steady_clock::time_point a) {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };
    // QR moving right to left (dtt_qr dir=-1): core i = Q^T, core i-1 <- core i-1 R^T
    auto qr_left = [&](DevTT& T, int i) {
      const int r1 = T.r[i], c = T.n[i] * T.r[i + 1];
      const int rnew = std::min(r1, c);
      transpose(r1, c, T.c[i].p, r1, t2, c, st);
      orth_factor(c, r1, t2, Q, Rm, Sv, st);   // Q (c x rnew), Rm (rnew x r1)
      transpose(c, rnew, Q, c, T.c[i].p, rnew, st);
      const int mp = T.r[i - 1] * T.n[i - 1];
      gemm_cm('N', 'T', mp, rnew, r1, 1.0, T.c[i - 1].p, mp, Rm, rnew, 0.0, t2, mp, st);
      TTF_HIP(hipMemcpyAsync(T.c[i - 1].p, t2, sizeof(double) * (size_t)mp * rnew, hipMemcpyDeviceToDevice, st));
      T.r[i] = rnew;
    };
    // QR moving left to right (dtt_qr dir=+1): core i = Q, core i+1 <- R core i+1
    auto qr_right = [&](DevTT& T, int i) {
      const int mrow = T.r[i] * T.n[i], c = T.r[i + 1];
      const int rnew = std::min(mrow, c);
      orth_factor(mrow, c, T.c[i].p, Q, Rm, Sv, st);
      TTF_HIP(hipMemcpyAsync(T.c[i].p, Q, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
      const int nc2 = T.n[i + 1] * T.r[i + 2];
      gemm_cm('N', 'N', rnew, nc2, c, 1.0, Rm, rnew, T.c[i + 1].p, c, 0.0, t2, rnew, st);
      TTF_HIP(hipMemcpyAsync(T.c[i + 1].p, t2, sizeof(double) * (size_t)rnew * nc2, hipMemcpyDeviceToDevice, st));
      T.r[i + 1] = rnew;
    };
    // out (rL n rR) = pL (rL x ry1) Y_i (ry1, n, ry2) pR (ry2 x rR)
    auto proj_y = [&](int i, int rL, const double* pL, int rR, const double* pR, double* out) {
      gemm_cm('N', 'N', rL, n[i] * ry[i + 1], ry[i], 1.0, pL, rL, dY + py[i], ry[i], 0.0, t1, rL, st);
      gemm_cm('N', 'N', rL * n[i], rR, ry[i + 1], 1.0, t1, (int64_t)rL * n[i], pR, ry[i + 1], 0.0, out,
              (int64_t)rL * n[i], st);
    };
    auto xshape = [&](int i) {
      return Shape{x.r[i], n[i], x.r[i + 1], x.r[i], n[i], x.r[i + 1], ra[i], ra[i + 1]};
    };
    double res_max = 0.0, err_max = 0.0, last_res = 0.0, last_err = 0.0;
    int sweeps = 0;
    bool converged = false;
    for (int swp = 1; swp <= nswp0; ++swp) {
      // ---- orthogonalisation right to left (:79-108)
      auto tp = std::chrono::steady_clock::now();
      for (int i = d - 1; i >= 1; --i) {
        if (kick > 0 && swp > 1) {
          // z_i = Z'(A x - y) at core i
          const Shape sz{x.r[i], n[i], x.r[i + 1], z.r[i], n[i], z.r[i + 1], ra[i], ra[i + 1]};
          bfun3_chain<double>(sz, 1, pzaL[i].p, dA + pa[i], pzaR[i + 1].p, x.c[i].p,
                              (int64_t)x.r[i] * n[i] * x.r[i + 1], z.c[i].p, (int64_t)z.r[i] * n[i] * z.r[i + 1], st);
          proj_y(i, z.r[i], pzyL[i].p, z.r[i + 1], pzyR[i + 1].p, rhs);
          hipLaunchKernelGGL(axpy_kernel, dim3(grid_for((int64_t)z.r[i] * n[i] * z.r[i + 1])), dim3(256), 0, st,
                             (int64_t)z.r[i] * n[i] * z.r[i + 1], -1.0, rhs, z.c[i].p);
        }
        qr_left(x, i);
        phi_right_chain<double>(xshape(i), pxaR[i + 1].p, dA + pa[i], x.c[i].p, x.c[i].p, pxaR[i].p, st);
        phi2_right_chain<double>(ry[i], ry[i + 1], x.r[i], n[i], x.r[i + 1], pxyR[i + 1].p, dY + py[i], x.c[i].p,
                                 pxyR[i].p, st);
        if (kick > 0) {
          qr_left(z, i);
          const Shape sz{x.r[i], n[i], x.r[i + 1], z.r[i], n[i], z.r[i + 1], ra[i], ra[i + 1]};
          phi_right_chain<double>(sz, pzaR[i + 1].p, dA + pa[i], x.c[i].p, z.c[i].p, pzaR[i].p, st);
          phi2_right_chain<double>(ry[i], ry[i + 1], z.r[i], n[i], z.r[i + 1], pzyR[i + 1].p, dY + py[i], z.c[i].p,
                                   pzyR[i].p, st);
        }
      }
      t_phi += secs(tp);
      res_max = 0.0;
      err_max = 0.0;
      // ---- optimisation left to right (:110-235)
      for (int i = 0; i < d; ++i) {
        const int rx1 = x.r[i], rx2 = x.r[i + 1];
        const int64_t N = (int64_t)rx1 * n[i] * rx2;
        const Shape s = xshape(i);
        LocalOp op{s, pxaL[i].p, dA + pa[i], pxaR[i + 1].p};
        auto ts = std::chrono::steady_clock::now();
        proj_y(i, rx1, pxyL[i].p, rx2, pxyR[i + 1].p, rhs);
        op.apply(1, x.c[i].p, N, res, N, st);
        ++mv_total;
        flops_total += bfun3_flops(s);
        hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, rhs, res);
        const double norm_rhs = dev_norm(N, rhs, scr, st);
        TTF_REQUIRE(norm_rhs > 0.0, "ttf_dtt_amen_solve: zero local right-hand side at core %d", i + 1);
        double err = dev_norm(N, res, scr, st) / norm_rhs;
        res_max = std::max(res_max, err);
        if (err > real_tol) {
          if (N < std::min(max_full_size, 129)) {   // dense local solve (:138-146), orders up to 128
            double* B = bB.get((size_t)N * N);
            fullmat_launch(rx1, n[i], rx2, ra[i], ra[i + 1], pxaL[i].p, dA + pa[i], pxaR[i + 1].p, B, st);
            dense_inverse((int)N, B, st);
            gemm_cm('N', 'N', (int)N, 1, (int)N, 1.0, B, N, res, N, 0.0, sol, N, st);
          } else {   // GMRES (:147-168); the rhs is the residual, sol the negative correction
            JacPrec P;
            if (ptype != 'n') {
              double* jacs = bjac.get(jac_size(ptype, rx1, n[i], rx2));
              jac_gen(ptype, rx1, n[i], rx2, ra[i], ra[i + 1], pxaL[i].p, dA + pa[i], pxaR[i + 1].p, jacs, st);
              P = JacPrec{ptype, rx1, n[i], rx2, jacs};
            }
            GmresStats gs = gmres_solve(op, P, res, sol, local_restart, real_tol / err, local_iters, st);
            mv_total += gs.matvecs;
            flops_total += bfun3_flops(s) * gs.matvecs;
            if (ptype != 'n') {
              jac_apply(ptype, rx1, n[i], rx2, P.jacs, sol, t2, st);
              TTF_HIP(hipMemcpyAsync(sol, t2, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            }
          }
          hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, sol, x.c[i].p);
          const double nx = dev_norm(N, x.c[i].p, scr, st);
          err_max = std::max(err_max, nx > 0.0 ? dev_norm(N, sol, scr, st) / nx : 0.0);
          op.apply(1, x.c[i].p, N, res, N, st);
          ++mv_total;
          flops_total += bfun3_flops(s);
          hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, rhs, res);
          err = dev_norm(N, res, scr, st) / norm_rhs;
        }
        t_solve += secs(ts);
        if (i == d - 1) break;
        auto tt = std::chrono::steady_clock::now();
        if (kick > 0) {
          // ---- dtt_trunc (ttnodeop.f90:227-302) of core i, dir +1
          const int mrow = rx1 * n[i], rr = rx2;
          const int rmin = std::min(rr, mrow);
          int kpos = 0;
          svd_vectors(mrow, rr, x.c[i].p, mrow, true, Sv, U, mrow, hs, &kpos, st);   // U (mrow x rmin)
          complete_columns(mrow, kpos, rmin, U, mrow, st);
          gemm_cm('T', 'N', rmin, rr, mrow, 1.0, U, mrow, x.c[i].p, mrow, 0.0, Rm, rmin, st);   // S V^T
          int rnew;
          if (trunc_norm == 1) {   // by local residual, tol max(err, real_tol resid_damp) (:187-189)
            const double ttol = std::max(err, real_tol * resid_damp);
            rnew = 0;
            for (int c = rmin - 1; c >= 1; --c) {
              gemm_cm('N', 'N', mrow, rr, c, 1.0, U, mrow, Rm, rmin, 0.0, t2, mrow, st);
              op.apply(1, t2, N, res, N, st);
              ++mv_total;
              flops_total += bfun3_flops(s);
              hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(N)), dim3(256), 0, st, N, -1.0, rhs, res);
              const double e = dev_norm(N, res, scr, st) / norm_rhs;
              if (e > ttol) {
                rnew = c;
                break;
              }
            }
            rnew += 1;
            rnew = std::min(rnew, rmin);
          } else {   // Frobenius chop (:191-192)
            rnew = chop_rank(hs.data(), rmin, real_tol * resid_damp, 0);
          }
          // xnew = U(:, :rnew) (S V^T)(:rnew, :) -- the truncated core, old ranks
          gemm_cm('N', 'N', mrow, rr, rnew, 1.0, U, mrow, Rm, rmin, 0.0, xnew, mrow, st);
          // core i = U(:, :rnew); core i+1 <- (S V^T)(:rnew, :) core i+1
          TTF_HIP(hipMemcpyAsync(x.c[i].p, U, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
          const int nc2 = n[i + 1] * x.r[i + 2];
          gemm_cm('N', 'N', rnew, nc2, rr, 1.0, Rm, rmin, x.c[i + 1].p, rr, 0.0, t2, rnew, st);
          TTF_HIP(hipMemcpyAsync(x.c[i + 1].p, t2, sizeof(double) * (size_t)rnew * nc2, hipMemcpyDeviceToDevice,
                                 st));
          x.r[i + 1] = rnew;
          // ---- enrichment (:196-219): sol = x'A z-projection of xnew minus x'y z-projection
          const int rz1 = z.r[i], rz2 = z.r[i + 1];
          const Shape sxz{rx1, n[i], rr, rx1, n[i], rz2, ra[i], ra[i + 1]};
          bfun3_chain<double>(sxz, 1, pxaL[i].p, dA + pa[i], pzaR[i + 1].p, xnew, N, sol,
                              (int64_t)rx1 * n[i] * rz2, st);
          const Shape szz{rx1, n[i], rr, rz1, n[i], rz2, ra[i], ra[i + 1]};
          bfun3_chain<double>(szz, 1, pzaL[i].p, dA + pa[i], pzaR[i + 1].p, xnew, N, z.c[i].p,
                              (int64_t)rz1 * n[i] * rz2, st);
          mv_total += 2;
          flops_total += bfun3_flops(sxz) + bfun3_flops(szz);
          proj_y(i, rx1, pxyL[i].p, rz2, pzyR[i + 1].p, t2);
          hipLaunchKernelGGL(axpy_kernel, dim3(grid_for((int64_t)rx1 * n[i] * rz2)), dim3(256), 0, st,
                             (int64_t)rx1 * n[i] * rz2, -1.0, t2, sol);
          proj_y(i, rz1, pzyL[i].p, rz2, pzyR[i + 1].p, t2);
          hipLaunchKernelGGL(axpy_kernel, dim3(grid_for((int64_t)rz1 * n[i] * rz2)), dim3(256), 0, st,
                             (int64_t)rz1 * n[i] * rz2, -1.0, t2, z.c[i].p);
          // dtt_enrich: core i <- [core i | sol], core i+1 <- [core i+1 ; 0]
          const int re = rnew + rz2;
          TTF_REQUIRE(re <= Rw, "ttf_dtt_amen_solve: rank %d exceeds the capacity %d (raise rmax)", re, Rw);
          TTF_HIP(hipMemcpyAsync(x.c[i].p + (size_t)mrow * rnew, sol, sizeof(double) * (size_t)mrow * rz2,
                                 hipMemcpyDeviceToDevice, st));
          hipLaunchKernelGGL(widen_kernel, dim3(grid_for((int64_t)re * nc2)), dim3(256), 0, st, (int64_t)0, rnew, rz2,
                             (int64_t)nc2, x.c[i + 1].p, t2);
          TTF_HIP(hipMemcpyAsync(x.c[i + 1].p, t2, sizeof(double) * (size_t)re * nc2, hipMemcpyDeviceToDevice, st));
          x.r[i + 1] = re;
          qr_right(x, i);
          qr_right(z, i);
          const Shape sz{x.r[i], n[i], x.r[i + 1], z.r[i], n[i], z.r[i + 1], ra[i], ra[i + 1]};
          phi_left_chain<double>(sz, pzaL[i].p, dA + pa[i], x.c[i].p, z.c[i].p, pzaL[i + 1].p, st);
          phi2_left_chain<double>(ry[i], ry[i + 1], z.r[i], n[i], z.r[i + 1], pzyL[i].p, dY + py[i], z.c[i].p,
                                  pzyL[i + 1].p, st);
        } else {
          qr_right(x, i);
        }
        TTF_REQUIRE(x.r[i + 1] <= Rw, "ttf_dtt_amen_solve: rank %d exceeds the capacity %d (raise rmax)",
                    x.r[i + 1], Rw);
        t_trunc += secs(tt);
        auto tp2 = std::chrono::steady_clock::now();
        phi_left_chain<double>(xshape(i), pxaL[i].p, dA + pa[i], x.c[i].p, x.c[i].p, pxaL[i + 1].p, st);
        phi2_left_chain<double>(ry[i], ry[i + 1], x.r[i], n[i], x.r[i + 1], pxyL[i].p, dY + py[i], x.c[i].p,
                                pxyL[i + 1].p, st);
        t_phi += secs(tp2);
      }
      sweeps = swp;
      last_res = res_max;
      last_err = err_max;
      if (verb > 0) {
        int mx = 0;
        for (int k = 0; k <= d; ++k) mx = std::max(mx, x.r[k]);
        std::fprintf(stderr, "amen_solve: swp=%d, max_dx=%.3e, max_res=%.3e, max_rank=%d\n", swp, err_max, res_max,
                     mx);
      }
      if ((trunc_norm == 1 ? res_max : err_max) < tol) {   // :239-243
        converged = true;
        break;
      }
    }
    for (int k = 0; k <= d; ++k) rx[k] = x.r[k];
    int64_t need = 0;
    for (int k = 0; k < d; ++k) need += (int64_t)rx[k] * n[k] * rx[k + 1];
    if (result_size) *result_size = need;
    TTF_REQUIRE(result == nullptr || result_cap >= need, "ttf_dtt_amen_solve: result capacity %lld < %lld",
                (long long)result_cap, (long long)need);
    if (result) {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        const int64_t sz = (int64_t)rx[k] * n[k] * rx[k + 1];
        TTF_HIP(hipMemcpyAsync(result + off, x.c[k].p, sizeof(double) * sz, hipMemcpyDeviceToHost, st));
        off += sz;
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
    if (stats) {
      *stats = ttf_amen_stats{};
      stats->sweeps = sweeps;
      stats->total_matvecs = mv_total;
      stats->last_res_max = last_res;
      stats->last_err_max = last_err;
      int mx = 0;
      for (int k = 0; k <= d; ++k) mx = std::max(mx, rx[k]);
      stats->max_rank = mx;
      stats->seconds = secs(t0);
      stats->solve_seconds = t_solve;
      stats->trunc_seconds = t_trunc;
      stats->phi_seconds = t_phi;
      stats->flops = flops_total;
    }
    (void)converged;
  });
}

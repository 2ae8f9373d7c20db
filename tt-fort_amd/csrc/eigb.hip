// tt_eigb on the device: the block eigensolver sweep of tt_eigb.f90:110-436.
//
// All cores, interface matrices (phi) and operator cores stay in HBM for the
// whole run; the host only steers (ranks, chop decisions, convergence).
//   * initial left-to-right orthogonalisation with phi_left updates
//     (tt_eigb.f90:201-217);
//   * sweeps starting right-to-left (dir = -1), one local eigenproblem per
//     core with the block index B carried by the current core
//     (tt_eigb.f90:230-412): dense path for small local size, else the
//     device block eigensolver (block_eig);
//   * truncation by SVD of the block core, rank = min(my_chop3(s, eps2),
//     rmax) (tt_eigb.f90:305-394; my_chop3 = tt_adapt_als.f90:40-61), the
//     singular values/vectors moving the B index to the neighbour core;
//   * phi_right / phi_left updates of the interfaces;
//   * stop when the largest relative change of sum(lambda) over a full
//     sweep is below eps (tt_eigb.f90:397-405), or after nswp sweeps.
#include "solvers.h"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <vector>

namespace ttf {

// Rank chosen by the reference's relative chop rule (tt_adapt_als.f90:40-61).
int my_chop3(const std::vector<double>& s, int n, double eps) {
  double nrm = 0.0;
  for (int i = 0; i < n; ++i) nrm += s[i] * s[i];
  nrm = std::sqrt(nrm);
  nrm = (nrm * eps) * (nrm * eps);
  double cursum = 0.0;
  int i = n;
  while (i > 0) {
    cursum += s[i - 1] * s[i - 1];
    if (cursum > nrm) break;
    --i;
  }
  return std::min(i + 1, n);
}

namespace {

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

}  // namespace

}  // namespace ttf

using namespace ttf;

// Test hook: the chop rule exactly as the sweep applies it to the singular
// values it reads back.
extern "C" int ttf_my_chop3(int n, const double* s, double eps) {
  if (n < 0 || (n > 0 && !s)) return -1;
  std::vector<double> v(s, s + n);
  return my_chop3(v, n, eps);
}

// Guard vectors of the local block eigensolver (block size B + guard):
// TTF_EIG_GUARD overrides the default 3 (0..16).
static int eig_guard_vectors() {
  static const int v = [] {
    const char* e = std::getenv("TTF_EIG_GUARD");
    const int x = e ? std::atoi(e) : 3;
    return std::min(std::max(x, 0), 16);
  }();
  return v;
}

extern "C" int ttf_tt_eigb(int d, const int* n, const int* m, const int* ra, const double* crA,
                           const double* crY0, int* ry, double eps, int rmax, double* lambda, int B,
                           int kickrank, int nswp, int max_full_size, int verb, double* result,
                           int64_t result_cap, int64_t* result_size, ttf_eigb_stats* stats) {
  return guarded([&] {
    (void)kickrank;   // accepted and ignored, as in the reference (tt_eigb.f90:165-168)
    TTF_REQUIRE(d >= 2, "ttf_tt_eigb: d=%d (need d >= 2)", d);
    TTF_REQUIRE(B >= 1 && ry[d] == B,
                "ttf_tt_eigb: inconsistent number of eigenvalues sought (B=%d) and last rank %d", B, ry[d]);
    for (int k = 0; k < d; ++k) TTF_REQUIRE(n[k] == m[k], "ttf_tt_eigb: operator must be square");
    TTF_REQUIRE(ry[0] == 1 && ra[0] == 1 && ra[d] == 1, "ttf_tt_eigb: boundary ranks must be 1");
    Timer total;
    hipStream_t st = stream();
    const double eps2 = eps / std::sqrt((double)d);
    const double resid_damp = 10.0;      // tt_eigb.f90:154
    const int max_matvecs = 500;         // tt_eigb.f90:152
    // LOBPCG guard vectors beyond the B wanted (TTF_EIG_GUARD overrides)
    const int guard = eig_guard_vectors();
    const int nswp0 = nswp > 0 ? nswp : 20;
    const bool mixed = local_matvec_bits() == 32;
    const int mfs = max_full_size > 0 ? max_full_size : 2000;
    // operator cores on the device
    std::vector<int64_t> pa(d + 1, 0);
    for (int k = 0; k < d; ++k) pa[k + 1] = pa[k] + (int64_t)ra[k] * n[k] * m[k] * ra[k + 1];
    double* dA = nullptr;
    TTF_HIP(hipMallocAsync((void**)&dA, sizeof(double) * std::max<int64_t>(pa[d], 1), st));
    TTF_HIP(hipMemcpyAsync(dA, crA, sizeof(double) * pa[d], hipMemcpyHostToDevice, st));
    int R = rmax;
    for (int k = 0; k <= d; ++k) R = std::max(R, ry[k]);
    int nmax = 1, ramax = 1;
    for (int k = 0; k < d; ++k) nmax = std::max(nmax, n[k]);
    for (int k = 0; k <= d; ++k) ramax = std::max(ramax, ra[k]);
    // cores: capacity for (R x nmax x R) x B; phis: R x ramax x R
    const size_t core_cap = (size_t)R * nmax * R * B;
    const size_t phi_cap = (size_t)R * ramax * R;
    std::vector<double*> core(d), phi(d + 1);
    for (int k = 0; k < d; ++k) TTF_HIP(hipMallocAsync((void**)&core[k], sizeof(double) * core_cap, st));
    for (int k = 0; k <= d; ++k) TTF_HIP(hipMallocAsync((void**)&phi[k], sizeof(double) * phi_cap, st));
    double* tmp = nullptr;
    TTF_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * core_cap, st));
    // SVD outputs (U: mu x k, S: k, VT: k x nv) for the largest unfolding
    // every unfolding is at most (max(B, n) R) x (max(B, n) R)
    const size_t wide = (size_t)std::max(B, nmax) * R;
    const size_t svd_cap = wide * wide;
    double* U = nullptr;
    TTF_HIP(hipMallocAsync((void**)&U, sizeof(double) * svd_cap, st));
    double* VT = nullptr;
    TTF_HIP(hipMallocAsync((void**)&VT, sizeof(double) * svd_cap, st));
    double* S = nullptr;
    TTF_HIP(hipMallocAsync((void**)&S, sizeof(double) * (size_t)R * std::max(B, nmax) * 4, st));
    {
      int64_t off = 0;
      for (int k = 0; k < d; ++k) {
        int64_t sz = (int64_t)ry[k] * n[k] * ry[k + 1];
        TTF_HIP(hipMemcpyAsync(core[k], crY0 + off, sizeof(double) * sz, hipMemcpyHostToDevice, st));
        off += sz;
      }
    }
    const double one = 1.0;
    TTF_HIP(hipMemcpyAsync(phi[0], &one, sizeof(double), hipMemcpyHostToDevice, st));
    TTF_HIP(hipMemcpyAsync(phi[d], &one, sizeof(double), hipMemcpyHostToDevice, st));
    double t_svd = 0, t_solve = 0, t_phi = 0;
    // ---- initial left-to-right orthogonalisation (tt_eigb.f90:201-217)
    for (int k = 0; k < d; ++k) {
      const int mrow = ry[k] * n[k], ncol = ry[k + 1];
      const int rnew = std::min(mrow, ncol);
      orth_factor(mrow, ncol, core[k], U, VT, S, st);   // U = Q, VT = R (rnew x ncol)
      TTF_HIP(hipMemcpyAsync(core[k], U, sizeof(double) * (size_t)mrow * rnew, hipMemcpyDeviceToDevice, st));
      if (k < d - 1) {
        const int ncol2 = n[k + 1] * ry[k + 2];
        GemmOp<double> g = colmajor<double>('N', 'N', rnew, ncol2, ry[k + 1], 1.0, VT, rnew, core[k + 1],
                                            ry[k + 1], 0.0, tmp, rnew);
        gemm_launch(g, st);
        TTF_HIP(hipMemcpyAsync(core[k + 1], tmp, sizeof(double) * (size_t)rnew * ncol2,
                               hipMemcpyDeviceToDevice, st));
        ry[k + 1] = rnew;
        Shape s{ry[k], m[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1]};
        phi_left_chain<double>(s, phi[k], dA + pa[k], core[k], core[k], phi[k + 1], st);
      }
    }
    ry[d] = 1;   // the block index now rides with the current core
    // ---- sweeps
    int k = d - 1, dir = -1, swp = 1;
    double ermax = 0.0, fvold = 0.0;
    long total_mv = 0, total_mv32 = 0, total_refresh = 0, block_solves = 0, block_fallbacks = 0;
    std::vector<double> lam(B, 0.0), hs;
    int last_swp = 0;
    std::vector<double> rec_sec, rec_flops, rec_er, rec_fv;
    std::vector<long> rec_mv, rec_mv32;
    // ermax of the last completed sweep and whether it met eps
    double last_er = 0.0;
    bool converged = false;
    Timer tsweep;
    double sweep_flops = 0.0;
    long sweep_mv = 0, sweep_mv32 = 0;
    while (swp <= nswp0) {
      const int N = ry[k] * n[k] * ry[k + 1];
      LocalOp op{Shape{ry[k], m[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1]}, phi[k],
                 dA + pa[k], phi[k + 1]};
      Timer ts;
      const int dense_max = std::min(mfs - 1, 64);
      // mixed precision: fp32 search-direction products, FP64 refinement
      MixedOp mixop;
      const MixedOp* mix = nullptr;
      if (mixed) {
        mixop = make_mixed(op, st, B + guard);
        mix = &mixop;
      }
      EigStats es = block_eig(op, B, guard, core[k], lam.data(), eps2 / resid_damp, max_matvecs,
                              dense_max, st, mix);
      total_mv32 += es.matvecs32;
      sweep_mv32 += es.matvecs32;
      total_refresh += es.refreshes;
      block_solves += es.block_prec > 0 ? 1 : 0;
      block_fallbacks += es.block_prec < 0 ? 1 : 0;
      t_solve += ts.ms();
      double fv = 0.0;
      for (int b = 0; b < B; ++b) fv += lam[b];
      const double erloc = (fvold - fv) / std::fabs(fv);
      ermax = std::max(ermax, erloc);
      if (verb > 0)
        std::fprintf(stderr, "swp: %d i: [%d/%d] loc_size: %d matvecs: %d res: %.3e ermax: %.3e dfv: %.3e\n", swp,
                    k + 1, d, N, es.matvecs, es.res, ermax, fvold - fv);
      total_mv += es.matvecs;
      sweep_mv += es.matvecs;
      sweep_flops += bfun3_flops(op.s) * es.matvecs;
      fvold = fv;
      if (dir < 0 && k > 0) {
        // (tt_eigb.f90:305-355) M = X^T viewed as (B*ry_k) x (n_k*ry_{k+1})
        Timer tv;
        const int mr = B * ry[k], nc = n[k] * ry[k + 1];
        const int kk = std::min(mr, nc);
        transpose(N, B, core[k], N, tmp, B, st);
        // right singular vectors V (nc x kk) into VT's storage
        int kpos = 0;
        const int sw = svd_vectors(mr, nc, tmp, mr, false, S, VT, nc, hs, &kpos, st);
        const int nn = std::min(my_chop3(hs, kk, eps2), rmax);
        if (verb > 1) std::fprintf(stderr, "svd: %d x %d sweeps %d rank %d\n", mr, nc, sw, nn);
        complete_columns(nc, kpos, nn, VT, nc, st);
        // core k = V(:, 0:nn)^T; US = M V(:, 0:nn) (mr x nn)
        transpose(nc, nn, VT, nc, core[k], nn, st);
        {
          GemmOp<double> g = colmajor<double>('N', 'N', mr, nn, nc, 1.0, tmp, mr, VT, nc, 0.0, U, mr);
          gemm_launch(g, st);
        }
        // core k-1 <- core(k-1) * US with (b, r_k, j) -> (r_k, j, b)
        const int mp = ry[k - 1] * n[k - 1];
        GemmOp<double> g{};
        g.M = mp; g.N = nn; g.K = ry[k];
        g.A = core[k - 1]; g.sAi = 1; g.sAk = mp;
        g.B = U; g.sBk = B; g.sBj = (int64_t)mr; g.sBb0 = 1;
        g.C = tmp; g.sCi = 1; g.sCj = mp; g.sCb0 = (int64_t)mp * nn;
        g.nb0 = B; g.nbatch = B; g.alpha = 1.0; g.beta = 0.0; g.ksplit = 1;
        gemm_launch(g, st);
        TTF_HIP(hipMemcpyAsync(core[k - 1], tmp, sizeof(double) * (size_t)mp * nn * B, hipMemcpyDeviceToDevice, st));
        ry[k] = nn;
        t_svd += tv.ms();
        Timer tp;
        Shape s{ry[k], m[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1]};
        phi_right_chain<double>(s, phi[k + 1], dA + pa[k], core[k], core[k], phi[k], st);
        t_phi += tp.ms();
      }
      if (dir > 0 && k < d - 1) {
        // (tt_eigb.f90:357-394) M = X viewed as (ry_k*n_k) x (ry_{k+1}*B)
        Timer tv;
        const int mr = ry[k] * n[k], nc = ry[k + 1] * B;
        const int kk = std::min(mr, nc);
        // left singular vectors U (mr x kk)
        int kpos = 0;
        const int sw = svd_vectors(mr, nc, core[k], mr, true, S, U, mr, hs, &kpos, st);
        const int nn = std::min(my_chop3(hs, kk, eps2), rmax);
        if (verb > 1) std::fprintf(stderr, "svd: %d x %d sweeps %d rank %d\n", mr, nc, sw, nn);
        complete_columns(mr, kpos, nn, U, mr, st);
        // SVT = U(:, 0:nn)^T M (nn x nc), layout (j, r_{k+1}, b)
        {
          GemmOp<double> g = colmajor<double>('T', 'N', nn, nc, mr, 1.0, U, mr, core[k], mr, 0.0, VT, nn);
          gemm_launch(g, st);
        }
        TTF_HIP(hipMemcpyAsync(core[k], U, sizeof(double) * (size_t)mr * nn, hipMemcpyDeviceToDevice, st));
        // core k+1 <- SVT * core(k+1): (j, n, r, b)
        const int ncol2 = n[k + 1] * ry[k + 2];
        GemmOp<double> g{};
        g.M = nn; g.N = ncol2; g.K = ry[k + 1];
        g.A = VT; g.sAi = 1; g.sAk = nn; g.sAb0 = (int64_t)nn * ry[k + 1];
        g.B = core[k + 1]; g.sBk = 1; g.sBj = ry[k + 1];
        g.C = tmp; g.sCi = 1; g.sCj = nn; g.sCb0 = (int64_t)nn * ncol2;
        g.nb0 = B; g.nbatch = B; g.alpha = 1.0; g.beta = 0.0; g.ksplit = 1;
        gemm_launch(g, st);
        TTF_HIP(hipMemcpyAsync(core[k + 1], tmp, sizeof(double) * (size_t)nn * ncol2 * B,
                               hipMemcpyDeviceToDevice, st));
        ry[k + 1] = nn;
        t_svd += tv.ms();
        Timer tp;
        Shape s{ry[k], m[k], ry[k + 1], ry[k], n[k], ry[k + 1], ra[k], ra[k + 1]};
        phi_left_chain<double>(s, phi[k], dA + pa[k], core[k], core[k], phi[k + 1], st);
        t_phi += tp.ms();
      }
      if (dir > 0 && k == d - 2) {
        dir = -1;
        k = d - 1;
        if (verb > 0) {
          int mx = 0;
          for (int q = 0; q < d; ++q) mx = std::max(mx, ry[q]);
          std::fprintf(stderr, "swp: %d er = %.3e rmax:%d\n", swp, ermax, mx);
        }
        last_swp = swp;
        ++swp;
        TTF_HIP(hipStreamSynchronize(st));
        rec_sec.push_back(tsweep.ms() * 1e-3);
        rec_flops.push_back(sweep_flops);
        rec_mv.push_back(sweep_mv);
        rec_mv32.push_back(sweep_mv32);
        rec_er.push_back(ermax);
        rec_fv.push_back(fvold);
        tsweep = Timer();
        sweep_flops = 0.0;
        sweep_mv = 0;
        sweep_mv32 = 0;
        last_er = ermax;
        if (ermax < eps) {
          converged = true;
          break;
        }
        ermax = 0.0;
      } else if (dir < 0 && k == 1) {
        dir = 1;
        k = 0;
      } else {
        k += dir;
      }
    }
    ry[d] = B;
    for (int b = 0; b < B; ++b) lambda[b] = lam[b];
    int64_t need = 0;
    for (int q = 0; q < d; ++q) need += (int64_t)ry[q] * n[q] * ry[q + 1];
    if (result_size) *result_size = need;
    TTF_REQUIRE(result == nullptr || result_cap >= need,
                "ttf_tt_eigb: result capacity %lld < %lld", (long long)result_cap, (long long)need);
    if (result) {
      int64_t off = 0;
      for (int q = 0; q < d; ++q) {
        int64_t sz = (int64_t)ry[q] * n[q] * ry[q + 1];
        TTF_HIP(hipMemcpyAsync(result + off, core[q], sizeof(double) * sz, hipMemcpyDeviceToHost, st));
        off += sz;
      }
    }
    TTF_HIP(hipStreamSynchronize(st));
    // stream-ordered frees: no device-wide synchronisation, so concurrent
    // solves on other streams keep running
    for (auto p : core) TTF_HIP(hipFreeAsync(p, st));
    for (auto p : phi) TTF_HIP(hipFreeAsync(p, st));
    TTF_HIP(hipFreeAsync(tmp, st));
    TTF_HIP(hipFreeAsync(U, st));
    TTF_HIP(hipFreeAsync(VT, st));
    TTF_HIP(hipFreeAsync(S, st));
    TTF_HIP(hipFreeAsync(dA, st));
    if (stats) {
      stats->sweeps = last_swp > 0 ? last_swp : nswp0;
      stats->total_matvecs = total_mv;
      stats->last_ermax = last_er;
      stats->converged = converged ? 1 : 0;
      stats->total_matvecs32 = total_mv32;
      stats->refreshes = total_refresh;
      stats->block_prec_solves = block_solves;
      stats->block_prec_fallbacks = block_fallbacks;
      int mx = 0;
      for (int q = 0; q <= d; ++q) mx = std::max(mx, ry[q]);
      stats->max_rank = mx;
      stats->seconds = total.ms() * 1e-3;
      stats->solve_seconds = t_solve * 1e-3;
      stats->trunc_seconds = t_svd * 1e-3;
      stats->phi_seconds = t_phi * 1e-3;
      stats->nrec = (int)std::min<size_t>(rec_sec.size(), 64);
      for (int q = 0; q < stats->nrec; ++q) {
        stats->sweep_seconds[q] = rec_sec[q];
        stats->sweep_flops[q] = rec_flops[q];
        stats->sweep_matvecs[q] = rec_mv[q];
        stats->sweep_ermax[q] = rec_er[q];
        stats->sweep_lambda_sum[q] = rec_fv[q];
        stats->sweep_matvecs32[q] = rec_mv32[q];
      }
    }
  });
}

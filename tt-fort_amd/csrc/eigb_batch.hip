// Many independent tt_eigb problems (chains of different length, ranks and
// operators) solved concurrently on one GPU.  Worker threads, each with its
// own stream and scratch arenas (ScopedContext), take chains from a shared
// queue ordered by decreasing estimated cost (d * rmax^3: the largest first,
// so the tail of the batch is made of short solves) and run the device
// tt_eigb driver on them.  The single-chain solves are latency bound
// (small dense steps, one work-group kernels, host convergence checks);
// running many at once fills the chip while each waits on its own stream.
#include "ttf_common.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using namespace ttf;

extern "C" int ttf_tt_eigb_batch(int nchains, const int* d, const int* n_all, const int* ra_all,
                                 const double* crA_all, const int64_t* crA_off, const double* crY0_all,
                                 const int64_t* crY0_off, int* ry_all, const double* eps, const int* rmax, int B,
                                 int nswp, int nworkers, double* lambda_all, double* result_all,
                                 const int64_t* result_off, int64_t* result_size, ttf_eigb_stats* stats,
                                 int* status) {
  return guarded([&] {
    TTF_REQUIRE(nchains >= 0 && B >= 1, "ttf_tt_eigb_batch: nchains=%d B=%d", nchains, B);
    if (nchains == 0) return;
    TTF_REQUIRE(d && n_all && ra_all && crA_all && crA_off && crY0_all && crY0_off && ry_all && eps && rmax &&
                    lambda_all && result_all && result_off && status,
                "ttf_tt_eigb_batch: null argument");
    ensure_init();
    // offsets of each chain's mode sizes (sum d) and ranks (sum d + 1)
    std::vector<int64_t> on(nchains + 1, 0), orr(nchains + 1, 0);
    for (int c = 0; c < nchains; ++c) {
      TTF_REQUIRE(d[c] >= 2, "ttf_tt_eigb_batch: chain %d has d=%d", c, d[c]);
      on[c + 1] = on[c] + d[c];
      orr[c + 1] = orr[c] + d[c] + 1;
    }
    for (int c = 0; c < nchains; ++c) status[c] = 1;   // not run yet
    std::vector<int> order(nchains);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int c) { return (double)d[c] * rmax[c] * rmax[c] * rmax[c]; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    const int nw = std::max(1, std::min(nworkers > 0 ? nworkers : 8, nchains));
    // TTF_PROGRESS set: one line on stderr per finished chain
    const bool progress = std::getenv("TTF_PROGRESS") != nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    std::atomic<int> next{0}, done{0};
    std::vector<std::string> errs(nchains);
    auto work = [&]() {
      try {
        ScopedContext ctx;
        for (;;) {
          const int k = next.fetch_add(1);
          if (k >= nchains) break;
          const int c = order[k];
          const int* nc = n_all + on[c];
          int64_t sz = 0;
          status[c] = ttf_tt_eigb(d[c], nc, nc, ra_all + orr[c], crA_all + crA_off[c], crY0_all + crY0_off[c],
                                  ry_all + orr[c], eps[c], rmax[c], lambda_all + (int64_t)B * c, B, 0, nswp, 0, 0,
                                  result_all + result_off[c], result_off[c + 1] - result_off[c], &sz,
                                  stats ? stats + c : nullptr);
          if (result_size) result_size[c] = sz;
          if (status[c] != 0) {
            errs[c] = ttf_last_error();
            ttf_clear_error();
          }
          const int nd = done.fetch_add(1) + 1;
          if (progress) {
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "ttf_tt_eigb_batch: chain %d (d=%d, rmax=%d) %s, %d of %d done, %.1f s\n", c,
                         d[c], rmax[c], status[c] == 0 ? "solved" : "failed", nd, nchains, sec);
          }
        }
      } catch (...) {
        // the worker could not start (stream creation): its chains stay
        // with the other workers
      }
    };
    std::vector<std::thread> threads;
    for (int t = 0; t < nw; ++t) threads.emplace_back(work);
    for (auto& t : threads) t.join();
    TTF_REQUIRE(next.load() >= nchains, "ttf_tt_eigb_batch: no worker could start");
    int nfail = 0, first = -1;
    for (int c = 0; c < nchains; ++c) {
      if (status[c] != 0) {
        ++nfail;
        if (first < 0) first = c;
      }
    }
    TTF_REQUIRE(nfail == 0, "ttf_tt_eigb_batch: %d of %d chains failed; chain %d: %s", nfail, nchains, first,
                first >= 0 ? errs[first].c_str() : "");
  });
}

// GEMM launchers: runtime autotuning of tile shape and split-K, and the
// small-panel kernel for products with a short K and a narrow shared B.
#include "gemm_mfma.h"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace ttf {

// ---------------------------------------------------------------- small panel
// C[b](i,j) = alpha sum_k A[b](i,k) B(k,j) + beta C[b](i,j) with K <= 64,
// N <= 32, B shared by all batches and A, C row-contiguous (sAi = sCi = 1).
// This is the shape of the middle bfun3 / phi stage for small mode sizes
// (K = m*ra2, N = ra1*n): memory-bound, so one thread per row i of one batch
// accumulates all N outputs while B sits in LDS.
// NN is N rounded up to the kernel's column bucket; B is zero-padded in LDS
// to (64 x NN) so the inner loops carry no guards.
template <typename T, int NN>
__global__ __launch_bounds__(256) void panel_kernel(GemmOp<T> op) {
  __shared__ T Bs[64][NN];
  const int K = op.K, N = op.N;
  for (int t = threadIdx.x; t < 64 * NN; t += blockDim.x) {
    int k = t / NN, j = t % NN;
    Bs[k][j] = (k < K && j < N) ? op.B[k * op.sBk + j * op.sBj] : T(0);
  }
  __syncthreads();
  const int64_t total = (int64_t)op.M * op.nbatch;
  // 32-bit unsigned index arithmetic whenever it fits (a 64-bit division is
  // a long instruction sequence on the vector ALU)
  const bool narrow = total <= 0x7fffffff;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int i, b;
    if (narrow) {
      i = (int)((unsigned)g % (unsigned)op.M);
      b = (int)((unsigned)g / (unsigned)op.M);
    } else {
      i = (int)(g % op.M);
      b = (int)(g / op.M);
    }
    int b0 = 0, b1 = b;
    if (op.nb0 != 1) {
      b0 = (int)((unsigned)b % (unsigned)op.nb0);
      b1 = (int)((unsigned)b / (unsigned)op.nb0);
    }
    const T* a = op.A + b0 * op.sAb0 + b1 * op.sAb1 + i;
    T* c = op.C + b0 * op.sCb0 + b1 * op.sCb1 + i;
    T acc[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) acc[j] = T(0);
    // the kernel is bound by memory latency: each chunk issues its 8 loads
    // together before the arithmetic that consumes them
    for (int k0 = 0; k0 < K; k0 += 8) {
      T av[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) av[u] = (k0 + u < K) ? a[(k0 + u) * op.sAk] : T(0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int j = 0; j < NN; ++j) acc[j] += av[u] * Bs[k0 + u][j];
    }
    // with beta != 0 every C element of the row is loaded before any is
    // used (one load under its own condition, used inside it, becomes a
    // branch with a full wait per column); columns past N read column 0
    T cv[NN];
    if (op.beta != T(0)) {
#pragma unroll
      for (int j = 0; j < NN; ++j) cv[j] = c[(j < N ? j : 0) * op.sCj];
    }
#pragma unroll
    for (int j = 0; j < NN; ++j)
      if (j < N) {
        T v = op.alpha * acc[j];
        if (op.beta != T(0)) v += op.beta * cv[j];
        c[j * op.sCj] = v;
      }
  }
}

template <typename T>
static bool panel_eligible(const GemmOp<T>& op) {
  return op.K <= 64 && op.N <= 32 && op.sAi == 1 && op.sCi == 1 && op.sBb0 == 0 &&
         op.sBb1 == 0 && (int64_t)op.M * op.nbatch >= 256;
}

// ---------------------------------------------------------------- autotuner
// Set while this thread records a stream capture: untuned shapes then take
// the heuristic configuration instead of timing runs (not recorded, so a
// later eager call still tunes them).
static thread_local bool t_no_tune = false;
void gemm_no_tuning(bool on) { t_no_tune = on; }

struct Cfg {
  int bm, bk, ks;   // tiles bm x bm (or bm x bn), K chunk bk, split-K factor ks
  int bn = 0;       // 0: square tiles; 80: 64 x 80 tiles, four waves stacked in rows
};

static int cfg_bn(const Cfg& c) { return c.bn > 0 ? c.bn : c.bm; }

struct Key {
  int type, M, N, K, nbatch, flags;
  bool operator==(const Key& o) const {
    return type == o.type && M == o.M && N == o.N && K == o.K && nbatch == o.nbatch &&
           flags == o.flags;
  }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    size_t h = 1469598103934665603ull;
    for (int v : {k.type, k.M, k.N, k.K, k.nbatch, k.flags}) h = (h ^ (size_t)v) * 1099511628211ull;
    return h;
  }
};

namespace {
std::mutex g_tune_mu;
std::unordered_map<Key, Cfg, KeyHash> g_tuned;
std::atomic<int> g_autotune{-1};   // -1: read TTF_AUTOTUNE on first use
long g_tune_count = 0;
bool g_table_loaded = false;
}  // namespace

// Persistent tuning table: when TTF_TUNE_FILE names a file, its entries
// ("type M N K nbatch flags bm bk ks bn" per line; a line without bn means
// square tiles) are loaded on first use and
// every newly tuned shape is appended, so that separate processes (a timed
// run and its counter-collection reruns, or successive jobs) launch the same
// configuration for the same product.  Called with g_tune_mu held.
static const char* tune_file() {
  const char* f = std::getenv("TTF_TUNE_FILE");
  return (f && f[0]) ? f : nullptr;
}

static int splitk_rule(int M, int N, int K, int nbatch);

static void load_tune_table() {
  if (g_table_loaded) return;
  g_table_loaded = true;
  const char* path = tune_file();
  if (!path) return;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return;
  char line[256];
  while (std::fgets(line, sizeof(line), fp)) {
    Key k{};
    Cfg c{};
    const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &k.type, &k.M, &k.N, &k.K, &k.nbatch,
                                &k.flags, &c.bm, &c.bk, &c.ks, &c.bn);
    if (got == 9) c.bn = 0;
    const bool valid = (got == 9 || got == 10) && (c.bm == 32 || c.bm == 64) &&
                       (c.bk == 16 || c.bk == 32 || c.bk == 64) && c.ks >= 1 && c.ks <= 16 &&
                       (c.bn == 0 || (c.bn == 80 && c.bm == 64 && c.bk <= 32) || (c.bn == 16 && c.bm == 32));
    if (valid) {
      c.ks = splitk_rule(k.M, k.N, k.K, k.nbatch);   // the split is never tuned
      g_tuned[k] = c;
    }
  }
  std::fclose(fp);
}

static void record_tuned(const Key& k, const Cfg& c) {
  const char* path = tune_file();
  if (!path) return;
  FILE* fp = std::fopen(path, "a");
  if (!fp) return;
  std::fprintf(fp, "%d %d %d %d %d %d %d %d %d %d\n", k.type, k.M, k.N, k.K, k.nbatch, k.flags, c.bm, c.bk,
               c.ks, c.bn);
  std::fclose(fp);
}

static bool autotune_enabled() {
  if (g_autotune.load() < 0) {
    const char* e = std::getenv("TTF_AUTOTUNE");
    g_autotune.store((e && e[0] == '0') ? 0 : 1);
  }
  return g_autotune.load() == 1;
}

// Split-K factor of a product, a function of its shape alone: the tuner
// varies tiles and K chunks, which leave every sum's order unchanged, but
// never the split, so one shape gives the same bits in every process and
// on every stream whatever the tuning found.  Products with fewer than 128
// 64 x 64 tiles split K into power-of-two ranges of at least 64 until about
// 512 work-groups are in flight.
static int splitk_rule(int M, int N, int K, int nbatch) {
  const int64_t t64 = ceil_div(M, 64) * ceil_div(N, 64) * (int64_t)nbatch;
  if (t64 >= 128 || K < 128) return 1;
  int ks = 1;
  while (ks < 8 && K / (2 * ks) >= 64 && t64 * ks < 512) ks *= 2;
  return ks;
}

static Cfg heuristic(int M, int N, int K, int nbatch) {
  const int64_t cus = 256;
  const int ks = splitk_rule(M, N, K, nbatch);
  int64_t t64 = ceil_div(M, 64) * ceil_div(N, 64) * (int64_t)nbatch;
  if (t64 >= cus / 2) return {64, 16, ks};
  return {32, 16, ks};
}

template <typename T, bool CT>
static void launch_tiles(const GemmOp<T>& op, const Cfg& c, dim3 grid, hipStream_t st) {
  if (cfg_bn(c) == 16 && c.bk == 16)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 16, 16, 1, CT>), grid, dim3(64), 0, st, op);
  else if (cfg_bn(c) == 16 && c.bk == 32)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 16, 32, 1, CT>), grid, dim3(64), 0, st, op);
  else if (cfg_bn(c) == 16)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 16, 64, 1, CT>), grid, dim3(64), 0, st, op);
  else if (cfg_bn(c) == 80 && c.bk == 16)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 80, 16, 4, CT, 4>), grid, dim3(256), 0, st, op);
  else if (cfg_bn(c) == 80)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 80, 32, 4, CT, 4>), grid, dim3(256), 0, st, op);
  else if (c.bm == 64 && c.bk == 16)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 64, 16, 4, CT>), grid, dim3(256), 0, st, op);
  else if (c.bm == 64 && c.bk == 32)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 64, 32, 4, CT>), grid, dim3(256), 0, st, op);
  else if (c.bm == 64)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 64, 64, 4, CT>), grid, dim3(256), 0, st, op);
  else if (c.bk == 16)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 32, 16, 1, CT>), grid, dim3(64), 0, st, op);
  else if (c.bk == 32)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 32, 32, 1, CT>), grid, dim3(64), 0, st, op);
  else
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 32, 32, 64, 1, CT>), grid, dim3(64), 0, st, op);
}

template <typename T>
static void launch_cfg(GemmOp<T> op, const Cfg& c, hipStream_t st) {
  op.ksplit = c.ks;
  int64_t tiles = ceil_div(op.M, c.bm) * ceil_div(op.N, cfg_bn(c));
  dim3 grid((unsigned)tiles, (unsigned)op.nbatch, (unsigned)op.ksplit);
  // column-contiguous outputs (and the column-major split-K slices) take the
  // swapped-fragment variant, whose lanes store consecutive rows
  if (op.ksplit > 1 || op.sCi == 1)
    launch_tiles<T, true>(op, c, grid, st);
  else
    launch_tiles<T, false>(op, c, grid, st);
  TTF_HIP(hipGetLastError());
  if (op.ksplit > 1) {
    int64_t total = (int64_t)op.M * op.N * op.nbatch;
    int blocks = (int)std::min<int64_t>(ceil_div(total, 256), 4096);
    hipLaunchKernelGGL((gemm_splitk_reduce<T>), dim3(blocks), dim3(256), 0, st, op);
    TTF_HIP(hipGetLastError());
  }
}

template <typename T>
static T* splitk_work(const GemmOp<T>& op, int ks) {
  size_t need = sizeof(T) * (size_t)ks * op.nbatch * (size_t)op.M * op.N;
  return (T*)arena(kArenaSplitK, need);
}

// Times each candidate tile shape and K chunk twice with events and keeps
// the fastest; the split-K factor is fixed by splitk_rule.  Only products that overwrite C (beta == 0) are tuned, since the
// candidates run the product several times.
template <typename T>
static Cfg tune(const GemmOp<T>& op0, hipStream_t st) {
  std::vector<Cfg> cands;
  // square tiles, 64 x 80 tiles (four waves of 16 x 80) when N takes them
  // evenly: N = 640 (a bond of 128 x MPO rank 5) splits into 8 such columns,
  // so that a batch of 8 fills the 256 CUs once; and one-wave 32 x 16 tiles
  // for N <= 16 (a block of search directions), which waste half as many
  // matrix-core products on padding columns as 32 x 32
  for (int bn : {0, 80, 16})
    for (int bm : {64, 32})
      for (int bk : {16, 32, 64}) {
        if (bn == 80 && (bm != 64 || bk > 32 || op0.N % 80 != 0)) continue;
        if (bn == 16 && (bm != 32 || op0.N > 16)) continue;
        Cfg c{bm, bk, splitk_rule(op0.M, op0.N, op0.K, op0.nbatch)};
        c.bn = bn;
        cands.push_back(c);
      }
  hipEvent_t e0, e1;
  TTF_HIP(hipEventCreate(&e0));
  TTF_HIP(hipEventCreate(&e1));
  Cfg best = heuristic(op0.M, op0.N, op0.K, op0.nbatch);
  float best_ms = 1e30f;
  for (const Cfg& c : cands) {
    GemmOp<T> op = op0;
    if (c.ks > 1) op.work = splitk_work(op, c.ks);
    launch_cfg(op, c, st);   // warm-up (code object load, caches)
    TTF_HIP(hipEventRecord(e0, st));
    launch_cfg(op, c, st);
    launch_cfg(op, c, st);
    TTF_HIP(hipEventRecord(e1, st));
    TTF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    TTF_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = c; }
  }
  TTF_HIP(hipEventDestroy(e0));
  TTF_HIP(hipEventDestroy(e1));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    ++g_tune_count;
  }
  return best;
}

template <typename T> static int type_id();
template <> int type_id<double>() { return 0; }
template <> int type_id<float>() { return 1; }

template <typename T>
void gemm_launch(GemmOp<T> op, hipStream_t st) {
  if (op.M <= 0 || op.N <= 0 || op.nbatch <= 0) return;
  if (op.nb0 <= 0) op.nb0 = 1;
  TTF_REQUIRE(op.nbatch <= 65535, "gemm: batch count %d exceeds the grid limit", op.nbatch);
  if (op.K <= 0) {
    // empty sum: C = beta*C (alpha*0)
    op.K = 0;
  }
  if (panel_eligible(op)) {
    int64_t total = (int64_t)op.M * op.nbatch;
    int blocks = (int)std::min<int64_t>(ceil_div(total, 256), 65535);
    if (op.N <= 4)
      hipLaunchKernelGGL((panel_kernel<T, 4>), dim3(blocks), dim3(256), 0, st, op);
    else if (op.N <= 8)
      hipLaunchKernelGGL((panel_kernel<T, 8>), dim3(blocks), dim3(256), 0, st, op);
    else if (op.N <= 16)
      hipLaunchKernelGGL((panel_kernel<T, 16>), dim3(blocks), dim3(256), 0, st, op);
    else
      hipLaunchKernelGGL((panel_kernel<T, 32>), dim3(blocks), dim3(256), 0, st, op);
    TTF_HIP(hipGetLastError());
    return;
  }
  const double flops = 2.0 * op.M * op.N * (double)op.K * op.nbatch;
  Key key{type_id<T>(), op.M, op.N, op.K, op.nbatch,
          (op.sAi == 1 ? 1 : 0) | (op.sBj == 1 ? 2 : 0) | (op.sCi == 1 ? 4 : 0)};
  Cfg c;
  bool have = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tune_table();
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) { c = it->second; have = true; }
  }
  if (!have) {
    const bool alias = (const void*)op.C == (const void*)op.A || (const void*)op.C == (const void*)op.B;
    // no timing runs while a stream capture records this thread's launches
    const bool tuned = autotune_enabled() && !t_no_tune && op.beta == T(0) && !alias && flops >= 4e6;
    c = tuned ? tune(op, st) : heuristic(op.M, op.N, op.K, op.nbatch);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!t_no_tune) g_tuned[key] = c;
    if (tuned) record_tuned(key, c);
  }
  if (c.ks > 1) op.work = splitk_work(op, c.ks);
  launch_cfg(op, c, st);
}

template <typename T>
size_t gemm_workspace(const GemmOp<T>& op) {
  if (op.M <= 0 || op.N <= 0 || op.nbatch <= 0 || panel_eligible(op)) return 0;
  const int ks = splitk_rule(op.M, op.N, op.K, op.nbatch);
  return ks > 1 ? (size_t)ks * op.nbatch * (size_t)op.M * op.N : 0;
}

template <typename T>
void gemm_launch_ws(GemmOp<T> op, T* ws, size_t ws_cap, hipStream_t st) {
  if (op.M <= 0 || op.N <= 0 || op.nbatch <= 0) return;
  if (op.nb0 <= 0) op.nb0 = 1;
  TTF_REQUIRE(op.nbatch <= 65535, "gemm: batch count %d exceeds the grid limit", op.nbatch);
  if (op.K <= 0) op.K = 0;
  if (panel_eligible(op)) {
    gemm_launch(op, st);   // the panel kernel needs no workspace and is never tuned
    return;
  }
  Key key{type_id<T>(), op.M, op.N, op.K, op.nbatch,
          (op.sAi == 1 ? 1 : 0) | (op.sBj == 1 ? 2 : 0) | (op.sCi == 1 ? 4 : 0)};
  Cfg c = heuristic(op.M, op.N, op.K, op.nbatch);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) c = it->second;
  }
  if (c.ks > 1) {
    const size_t need = (size_t)c.ks * op.nbatch * (size_t)op.M * op.N;
    TTF_REQUIRE(ws != nullptr && need <= ws_cap, "gemm: split-K workspace of %zu elements exceeds the %zu given",
                need, ws_cap);
    op.work = ws;
  }
  launch_cfg(op, c, st);
}

template void gemm_launch<double>(GemmOp<double>, hipStream_t);
template void gemm_launch<float>(GemmOp<float>, hipStream_t);
template size_t gemm_workspace<double>(const GemmOp<double>&);
template size_t gemm_workspace<float>(const GemmOp<float>&);
template void gemm_launch_ws<double>(GemmOp<double>, double*, size_t, hipStream_t);
template void gemm_launch_ws<float>(GemmOp<float>, float*, size_t, hipStream_t);

template <typename R>
static void zgemm_launch_t(ZGemmOpT<R> op, hipStream_t st) {
  typedef typename CplxT<R>::type C_t;
  if (op.M <= 0 || op.N <= 0 || op.nbatch <= 0) return;
  if (op.nb0 <= 0) op.nb0 = 1;
  TTF_REQUIRE(op.nbatch <= 65535, "zgemm: batch count %d exceeds the grid limit", op.nbatch);
  Cfg c = heuristic(op.M, op.N, op.K, op.nbatch);
  int64_t tiles = ceil_div(op.M, c.bm) * ceil_div(op.N, c.bm);
  // few work-groups with a long K: slices of K in parallel (at least 48 deep
  // each), summed by a second launch
  op.ksplit = 1;
  op.work = nullptr;
  if (c.bm == 32 && tiles * op.nbatch < 32 && op.K >= 128) {
    const int ks = (int)std::min<int64_t>({8, op.K / 48, ceil_div(64, tiles * op.nbatch)});
    if (ks > 1) {
      op.ksplit = ks;
      op.work = (C_t*)arena(kArenaSplitK, sizeof(C_t) * (size_t)ks * op.nbatch * (size_t)op.M * op.N);
    }
  }
  dim3 grid((unsigned)tiles, (unsigned)op.nbatch, (unsigned)op.ksplit);
  // 32x32 tiles come with few work-groups, each one's K loop latency-bound
  // on its global loads: a 32-deep K chunk halves the trips
  if (c.bm == 64)
    hipLaunchKernelGGL((zgemm_mfma_kernel<R, 64, 64>), grid, dim3(256), 0, st, op);
  else if (op.K > 16)
    hipLaunchKernelGGL((zgemm_mfma_kernel<R, 32, 32, 32>), grid, dim3(256), 0, st, op);
  else
    hipLaunchKernelGGL((zgemm_mfma_kernel<R, 32, 32>), grid, dim3(256), 0, st, op);
  TTF_HIP(hipGetLastError());
  if (op.ksplit > 1) {
    const int64_t total = (int64_t)op.M * op.N * op.nbatch;
    hipLaunchKernelGGL(zgemm_splitk_reduce<R>, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 4096)),
                       dim3(256), 0, st, op);
    TTF_HIP(hipGetLastError());
  }
}

void zgemm_launch(ZGemmOp op, hipStream_t st) { zgemm_launch_t<double>(op, st); }
void cgemm_launch(CGemmOp op, hipStream_t st) { zgemm_launch_t<float>(op, st); }

long gemm_tuned_shapes() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (long)g_tuned.size();
}

}  // namespace ttf

extern "C" {

int ttf_set_autotune(int enable) {
  std::lock_guard<std::mutex> lk(ttf::g_tune_mu);
  ttf::g_autotune = enable ? 1 : 0;
  if (!enable) ttf::g_tuned.clear();
  return 0;
}

long ttf_autotune_count(void) {
  std::lock_guard<std::mutex> lk(ttf::g_tune_mu);
  return ttf::g_tune_count;
}

}  // extern "C"

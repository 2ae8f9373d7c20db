// Library runtime: error state, device/stream set-up, scratch arenas,
// memory helpers and stream timers.
#include "ttf_common.h"

#include <cstdarg>
#include <cstring>
#include <atomic>
#include <mutex>

namespace ttf {

namespace {
thread_local char g_err[1024] = {0};
std::mutex g_mu;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
struct ArenaSlot { void* p = nullptr; size_t bytes = 0; };
ArenaSlot g_arena[kArenaCount];
// pinned host staging for read_back (allocated on first use)
constexpr size_t kStageBytes = 256 * 1024;
void* g_stage = nullptr;
std::atomic<int> g_local_bits{64};

// A worker's own stream and arenas (ScopedContext); null on threads that use
// the library-wide state.
struct Context {
  hipStream_t stream = nullptr;
  ArenaSlot arena[kArenaCount];
  void* stage = nullptr;
};
thread_local Context* t_ctx = nullptr;
}  // namespace

int local_matvec_bits() { return g_local_bits.load(); }

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

bool has_error() { return g_err[0] != 0; }

static void init_locked(int device) {
  if (g_init) return;
  int count = 0;
  TTF_HIP(hipGetDeviceCount(&count));
  TTF_REQUIRE(count > 0, "no HIP device available");
  TTF_REQUIRE(device >= 0 && device < count, "device %d out of range (%d devices)", device, count);
  TTF_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  TTF_HIP(hipGetDeviceProperties(&prop, device));
  TTF_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0,
              "this library is built for gfx950 (MI355X); device %d is %s", device,
              prop.gcnArchName);
  noalloc_setup(device);
  TTF_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  TTF_HIP(hipEventCreate(&g_ev0));
  TTF_HIP(hipEventCreate(&g_ev1));
  g_device = device;
  g_init = true;
}

void ensure_init() {
  std::lock_guard<std::mutex> lk(g_mu);
  init_locked(0);
}

hipStream_t stream() {
  if (t_ctx) return t_ctx->stream;
  ensure_init();
  return g_stream;
}

void* arena(Arena a, size_t bytes) {
  ensure_init();
  ArenaSlot& s = t_ctx ? t_ctx->arena[a] : g_arena[a];
  if (bytes == 0) bytes = 256;
  if (s.bytes < bytes) {
    if (s.p) {
      TTF_HIP(hipStreamSynchronize(stream()));
      TTF_HIP(hipFree(s.p));
      s.p = nullptr;
      s.bytes = 0;
    }
    size_t nb = bytes + bytes / 4;   // grow with headroom
    nb = (nb + 4095) & ~size_t(4095);
    TTF_HIP(hipMalloc(&s.p, nb));
    s.bytes = nb;
  }
  return s.p;
}

void read_back(void* host, const void* dev, size_t bytes, hipStream_t st) {
  if (bytes == 0) {
    TTF_HIP(hipStreamSynchronize(st));
    return;
  }
  void** stage = t_ctx ? &t_ctx->stage : &g_stage;
  if (bytes <= kStageBytes) {
    if (*stage == nullptr) TTF_HIP(hipHostMalloc(stage, kStageBytes, hipHostMallocDefault));
    TTF_HIP(hipMemcpyAsync(*stage, dev, bytes, hipMemcpyDeviceToHost, st));
    TTF_HIP(hipStreamSynchronize(st));
    std::memcpy(host, *stage, bytes);
  } else {
    TTF_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
    TTF_HIP(hipStreamSynchronize(st));
  }
}

ScopedContext::ScopedContext() {
  ensure_init();
  TTF_REQUIRE(t_ctx == nullptr, "nested worker contexts");
  TTF_HIP(hipSetDevice(g_device));
  Context* c = new Context();
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    TTF_REQUIRE(false, "could not create a worker stream");
  }
  t_ctx = c;
  impl_ = c;
}

ScopedContext::~ScopedContext() {
  Context* c = static_cast<Context*>(impl_);
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& a : c->arena)
    if (a.p) (void)hipFree(a.p);
  if (c->stage) (void)hipHostFree(c->stage);
  (void)hipStreamDestroy(c->stream);
  if (t_ctx == c) t_ctx = nullptr;
  delete c;
}

}  // namespace ttf

using namespace ttf;

extern "C" {

int ttf_init(int device) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_mu);
    init_locked(device);
  });
}

int ttf_finalize(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_init) return;
    TTF_HIP(hipStreamSynchronize(g_stream));
    for (auto& s : g_arena) {
      if (s.p) TTF_HIP(hipFree(s.p));
      s.p = nullptr;
      s.bytes = 0;
    }
    if (g_stage) TTF_HIP(hipHostFree(g_stage));
    g_stage = nullptr;
    TTF_HIP(hipEventDestroy(g_ev0));
    TTF_HIP(hipEventDestroy(g_ev1));
    TTF_HIP(hipStreamDestroy(g_stream));
    g_stream = nullptr;
    g_init = false;
  });
}

int ttf_set_local_precision(int bits) {
  return guarded([&] {
    TTF_REQUIRE(bits == 64 || bits == 32, "ttf_set_local_precision: %d bits (64 or 32)", bits);
    g_local_bits.store(bits);
  });
}

int ttf_get_local_precision(void) { return g_local_bits.load(); }

const char* ttf_last_error(void) { return g_err; }
void ttf_clear_error(void) { g_err[0] = 0; }
const char* ttf_version(void) { return "ttfort_amd 0.1 (gfx950)"; }

int ttf_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    ensure_init();
    TTF_REQUIRE(ptr != nullptr, "ttf_malloc: null output pointer");
    TTF_HIP(hipMalloc(ptr, bytes ? bytes : 16));
  });
}

int ttf_free(void* ptr) {
  return guarded([&] {
    if (!ptr) return;
    TTF_HIP(hipStreamSynchronize(stream()));
    TTF_HIP(hipFree(ptr));
  });
}

int ttf_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    TTF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    TTF_HIP(hipStreamSynchronize(stream()));
  });
}

int ttf_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    TTF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    TTF_HIP(hipStreamSynchronize(stream()));
  });
}

int ttf_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    TTF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  });
}

int ttf_memset(void* dst, int value, size_t bytes) {
  return guarded([&] { TTF_HIP(hipMemsetAsync(dst, value, bytes, stream())); });
}

int ttf_sync(void) {
  return guarded([&] { TTF_HIP(hipStreamSynchronize(stream())); });
}

int ttf_timer_start(void) {
  return guarded([&] {
    hipStream_t s = stream();
    TTF_HIP(hipEventRecord(g_ev0, s));
  });
}

int ttf_timer_stop(float* ms) {
  return guarded([&] {
    hipStream_t s = stream();
    TTF_HIP(hipEventRecord(g_ev1, s));
    TTF_HIP(hipEventSynchronize(g_ev1));
    float t = 0.f;
    TTF_HIP(hipEventElapsedTime(&t, g_ev0, g_ev1));
    if (ms) *ms = t;
  });
}

}  // extern "C"

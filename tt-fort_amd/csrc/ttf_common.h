// Common definitions for the tt-fort MI355X library (gfx950, HIP).
//
// Error model: every C-ABI entry point returns an int status (0 = success)
// and records a human-readable message retrievable with ttf_last_error().
// Entry points that mirror the reference's void Fortran subroutines cannot
// return a status; they record a sticky error that ttf_last_error() reports
// and ttf_clear_error() resets.  Nothing in the library calls exit().
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_complex.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "ttfort_amd.h"

namespace ttf {

void set_error(const char* fmt, ...);
bool has_error();

struct HipFail {};

#define TTF_HIP(call)                                                          \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::ttf::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,      \
                       hipGetErrorString(_e));                                 \
      throw ::ttf::HipFail{};                                                  \
    }                                                                          \
  } while (0)

#define TTF_REQUIRE(cond, ...)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::ttf::set_error(__VA_ARGS__);                                           \
      throw ::ttf::HipFail{};                                                  \
    }                                                                          \
  } while (0)

// Runs `body` and converts library failures into an int status.
template <class Fn>
int guarded(Fn&& body) {
  try {
    body();
    return 0;
  } catch (const HipFail&) {
    return -1;
  } catch (const std::exception& e) {
    set_error("exception: %s", e.what());
    return -2;
  }
}

// Library-wide device state: the stream every kernel is launched on and a
// set of growing scratch arenas (device memory beyond the caller's inputs and
// outputs: contraction intermediates, solver bases).  Arenas grow with
// hipMalloc and are released by ttf_finalize().
hipStream_t stream();

// While alive, gives the calling host thread its own stream and scratch
// arenas (on the library's device): every library call made by that thread
// then runs on them, so that several threads can drive independent solves
// concurrently.  Destruction waits for the stream and frees the arenas.
class ScopedContext {
 public:
  ScopedContext();
  ~ScopedContext();
  ScopedContext(const ScopedContext&) = delete;
  ScopedContext& operator=(const ScopedContext&) = delete;

 private:
  void* impl_ = nullptr;
};
// Precision of the matvecs inside the iterative local solvers: 64 (default)
// or 32 (fp32 matrix cores, FP64 refinement); set by ttf_set_local_precision.
int local_matvec_bits();
void ensure_init();
// Resolves this device's copy of the fixed scratch array of the
// no-allocation entry points (noalloc.hip); called by the library's init
// with `device` current.
void noalloc_setup(int device);

enum Arena {
  kArenaBfun = 0,    // contraction intermediates (res1 / res2)
  kArenaHostIO,      // staging for the host-pointer drop-in entry points
  kArenaHostIO2,
  kArenaSolver,      // local solver bases (LOBPCG / GMRES / Krylov)
  kArenaSmall,       // small dense matrices and scalars
  kArenaTrunc,       // SVD / QR workspaces
  kArenaTrunc2,
  kArenaTrunc3,
  kArenaGram,        // partial sums of tall-skinny Gram products
  kArenaSplitK,      // split-K partial sums of the GEMM engine
  kArenaSolver2,
  kArenaDriver,      // sweep-driver staging (truncation factors)
  kArenaMlp,         // learned preconditioner: features and activations
  kArenaMlp2,        // learned preconditioner: gradients, loss partials, indices
  kArenaMlp3,        // learned preconditioner: batch targets
  kArenaZ,           // flexible GMRES: preconditioned directions
  kArenaSvdLong,     // long-side singular vectors: Cholesky QR workspace
  kArenaSvdLong2,    // long-side singular vectors: short-side vectors
  kArenaMixed,       // fp32 operator copy and vector staging of the mixed path
  kArenaTsqr,        // tall-skinny QR: row chunks and stacked R factors
  kArenaEigPrec,     // shifted block preconditioner of the local eigensolver
  kArenaEigPrec2,    // its Schur-complement workspace and pivot flags
  kArenaCount
};
void* arena(Arena a, size_t bytes);

// Copies `bytes` of device memory to host memory and waits for the stream,
// through a pinned staging buffer of the calling thread's context: a copy
// into pageable memory would be staged and waited for a second time inside
// the runtime (about 23 us per round trip against 14 us).
void read_back(void* host, const void* dev, size_t bytes, hipStream_t st);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace ttf

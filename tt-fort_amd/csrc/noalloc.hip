// The block matvec for callers that cannot allocate device memory (code
// under stream capture, tight serving loops, harnesses that forbid it):
// ttf_dbfun3 with its stage intermediates and split-K slices in one fixed
// device array instead of the library's growing arenas.
//
// The array `scratch` is a file-scope __device__ array; every GPU that runs
// the library holds its own copy until the process ends.  Its name in host
// code is only a stand-in for the symbol, so the device address of each
// GPU's copy is looked up once per device with hipGetSymbolAddress (at the
// library's init, with that device current, or at the first call on a
// device) and that address is what the kernels are given.
//
// Every call on a GPU shares that GPU's copy and leaves its contents behind;
// nothing in it needs to start at zero (each stage overwrites what it
// reads).  Calls that overlap on one GPU (several host threads, or streams
// running concurrently) race on it.
#include "contract.h"

#include <mutex>

namespace ttf {

namespace {

// 64 MiB: a block of 8 vectors at (r, n, r, R) = (128, 2, 128, 5) needs
// 4.7 M doubles (two 1.3 M intermediates and 2.1 M of split-K slices)
constexpr size_t kScratchDoubles = size_t(8) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`; `device` must be current.
double* scratch_for(int device) {
  TTF_REQUIRE(device >= 0 && device < kMaxDevices, "no-allocation scratch: device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    TTF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TTF_REQUIRE(p != nullptr, "no-allocation scratch: no device address for device %d", device);
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

void noalloc_setup(int device) { (void)scratch_for(device); }

}  // namespace ttf

using namespace ttf;

extern "C" int ttf_dbfun3_noalloc(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
                                  const double* phi1, const double* A, const double* phi2, const double* x,
                                  int64_t ldx, double* y, int64_t ldy) {
  return guarded([&] {
    TTF_REQUIRE(rx1 > 0 && m > 0 && rx2 > 0 && ry1 > 0 && n > 0 && ry2 > 0 && ra1 > 0 && ra2 > 0,
                "ttf_dbfun3_noalloc: invalid sizes rx1=%d m=%d rx2=%d ry1=%d n=%d ry2=%d ra1=%d ra2=%d", rx1, m,
                rx2, ry1, n, ry2, ra1, ra2);
    TTF_REQUIRE(ldx >= (int64_t)rx1 * m * rx2 || nvec <= 1, "ttf_dbfun3_noalloc: ldx too small");
    TTF_REQUIRE(ldy >= (int64_t)ry1 * n * ry2 || nvec <= 1, "ttf_dbfun3_noalloc: ldy too small");
    if (nvec <= 0) return;
    const Shape s{rx1, m, rx2, ry1, n, ry2, ra1, ra2};
    const size_t need = bfun3_workspace(s, nvec);
    TTF_REQUIRE(need <= kScratchDoubles,
                "ttf_dbfun3_noalloc: this product needs %zu doubles of scratch, more than the fixed %zu "
                "(64 MiB); use ttf_dbfun3 or fewer vectors per call",
                need, kScratchDoubles);
    hipStream_t st = stream();
    int device = 0;
    TTF_HIP(hipGetDevice(&device));
    double* ws = scratch_for(device);
    bfun3_chain_ws<double>(s, nvec, phi1, A, phi2, x, ldx, y, ldy, ws, kScratchDoubles, st);
  });
}

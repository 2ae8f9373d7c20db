// SDV files: tensor trains in tt-fort's binary stream format (ttio.f90).
// Host code only; no entry point here touches the device, so these calls
// work before ttf_init and on machines without a GPU.
//
// Layout (little-endian, no record markers: the reference opens the file
// with access='stream'):
//   128-byte header (ttio.f90:10-17): 'TT      ', ver(2), inf(4), a 64-byte
//     comment, i(8), all integers int32;
//   l, m                      int32 (first and last core index);
//   n(l:m), r(l-1:m)          int32 (mode sizes, ranks);
//   the cores back to back    core k is (r(k-1), n(k), r(k)) column-major,
//                             float64, or complex128 when inf(2) = 1.
//
// The writers fill the header as dtt_write / ztt_write do (ttio.f90:29-109,
// 110-192): ver = (1, 0), inf = (1023, 0 or 1, 0, 0) with 1023 the
// reference's tt_size (tt.f90:9), i(1:2) = (l, m).  The comment and
// i(3:8) are left unset there; here they are zeros.
#include "ttf_common.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstring>
#include <vector>

namespace ttf {
namespace {

constexpr int kHeaderBytes = 128;
constexpr int32_t kVersion = 1;    // ver(1) the readers accept (ttio.f90:8)
constexpr int32_t kTtSize = 1023;  // inf(1) written by the reference

// Status values, after the info codes of dtt_read (ttio.f90:196-297)
constexpr int kMissing = -1;    // the file does not exist
constexpr int kNotTT = -2;      // the header does not start with 'TT'
constexpr int kVersionErr = -3; // ver(1) is not 1
constexpr int kShort = -4;      // the file ends before its header says
constexpr int kBadArg = -5;     // a caller buffer too small, or invalid sizes
constexpr int kKind = -6;       // real / complex payload not the one asked for

struct File {
  FILE* fp = nullptr;
  ~File() {
    if (fp) std::fclose(fp);
  }
};

struct Dims {
  int32_t l = 0, m = 0;
  bool complex_ = false;
  std::vector<int32_t> n, r;   // n: d entries, r: d + 1
  int64_t count = 0;           // elements over all cores
};

bool read_exact(FILE* fp, void* p, size_t bytes) { return std::fread(p, 1, bytes, fp) == bytes; }

// Element count of the cores, or -1 when a size is negative or the total
// exceeds 2^62.
int64_t core_elements(int d, const int32_t* n, const int32_t* r) {
  unsigned __int128 total = 0;
  for (int k = 0; k < d; ++k) {
    if (n[k] < 0 || r[k] < 0 || r[k + 1] < 0) return -1;
    total += (unsigned __int128)(uint32_t)r[k] * (uint32_t)n[k] * (uint32_t)r[k + 1];
    if (total > ((unsigned __int128)1 << 62)) return -1;
  }
  return (int64_t)total;
}

// Opens `path` and parses everything up to the first core; returns 0 or a
// status from the list above (errno values for I/O failures).
int open_dims(const char* path, File& f, Dims& dm) {
  if (path == nullptr) {
    set_error("sdv: null file name");
    return kBadArg;
  }
  f.fp = std::fopen(path, "rb");
  if (f.fp == nullptr) {
    const int e = errno;
    set_error("sdv: cannot open %s: %s", path, std::strerror(e));
    return e == ENOENT ? kMissing : (e > 0 ? e : kMissing);
  }
  struct stat sb;
  if (fstat(fileno(f.fp), &sb) != 0 || S_ISDIR(sb.st_mode)) {
    const int e = S_ISDIR(sb.st_mode) ? EISDIR : errno;
    set_error("sdv: cannot read %s: %s", path, std::strerror(e));
    return e > 0 ? e : kShort;
  }
  unsigned char head[kHeaderBytes];
  if (!read_exact(f.fp, head, sizeof(head))) {
    set_error("sdv: %s is shorter than the 128-byte header", path);
    return kShort;
  }
  if (head[0] != 'T' || head[1] != 'T') {
    set_error("sdv: %s is not a TT file", path);
    return kNotTT;
  }
  int32_t ver[2], inf[4];
  std::memcpy(ver, head + 8, sizeof(ver));
  std::memcpy(inf, head + 16, sizeof(inf));
  if (ver[0] != kVersion) {
    set_error("sdv: %s has version %d.%d; version %d is supported", path, ver[0], ver[1], kVersion);
    return kVersionErr;
  }
  dm.complex_ = inf[1] == 1;
  int32_t lm[2];
  if (!read_exact(f.fp, lm, sizeof(lm))) {
    set_error("sdv: %s ends inside l, m", path);
    return kShort;
  }
  dm.l = lm[0];
  dm.m = lm[1];
  const int64_t d = (int64_t)dm.m - dm.l + 1;
  if (d < 1 || d > (1 << 20)) {
    set_error("sdv: %s has l = %d, m = %d", path, dm.l, dm.m);
    return kBadArg;
  }
  dm.n.resize(d);
  dm.r.resize(d + 1);
  if (!read_exact(f.fp, dm.n.data(), sizeof(int32_t) * d) ||
      !read_exact(f.fp, dm.r.data(), sizeof(int32_t) * (d + 1))) {
    set_error("sdv: %s ends inside the mode sizes and ranks", path);
    return kShort;
  }
  dm.count = core_elements((int)d, dm.n.data(), dm.r.data());
  if (dm.count < 0) {
    set_error("sdv: %s holds a negative or oversized mode size or rank", path);
    return kBadArg;
  }
  // the cores must be in the file: a corrupt size is reported here, before
  // a caller allocates for it
  const int64_t have = (int64_t)sb.st_size - (kHeaderBytes + 8 + 4 * (2 * d + 1));
  if (have / (dm.complex_ ? 16 : 8) < dm.count) {
    set_error("sdv: %s ends inside its cores (%lld elements expected)", path, (long long)dm.count);
    return kShort;
  }
  return 0;
}

// Reads the cores of a file of the given kind into out (cap elements of
// elem bytes each).
int read_cores(const char* path, bool want_complex, int64_t cap, void* out) {
  File f;
  Dims dm;
  const int st = open_dims(path, f, dm);
  if (st != 0) return st;
  if (dm.complex_ != want_complex) {
    set_error("sdv: %s holds %s cores; use %s", path, dm.complex_ ? "complex" : "real",
              dm.complex_ ? "ttf_ztt_read" : "ttf_dtt_read");
    return kKind;
  }
  if (cap < dm.count || (dm.count > 0 && out == nullptr)) {
    set_error("sdv: %s holds %lld elements; the buffer has room for %lld", path, (long long)dm.count,
              (long long)cap);
    return kBadArg;
  }
  const size_t elem = want_complex ? 16 : 8;
  if (dm.count > 0 && !read_exact(f.fp, out, elem * (size_t)dm.count)) {
    set_error("sdv: %s ends inside its cores (%lld elements expected)", path, (long long)dm.count);
    return kShort;
  }
  return 0;
}

int write_file(const char* path, int l, int m, const int* n, const int* r, const void* cores, bool complex_) {
  if (path == nullptr || n == nullptr || r == nullptr) {
    set_error("sdv: null file name, mode sizes or ranks");
    return kBadArg;
  }
  const int64_t d = (int64_t)m - l + 1;
  if (d < 1 || d > (1 << 20)) {
    set_error("sdv: l = %d, m = %d is not a valid core range", l, m);
    return kBadArg;
  }
  const int64_t count = core_elements((int)d, n, r);
  if (count < 0) {
    set_error("sdv: negative or oversized mode size or rank");
    return kBadArg;
  }
  if (count > 0 && cores == nullptr) {
    set_error("sdv: null core array for %lld elements", (long long)count);
    return kBadArg;
  }
  unsigned char head[kHeaderBytes];
  std::memset(head, 0, sizeof(head));
  std::memcpy(head, "TT      ", 8);
  const int32_t ver[2] = {kVersion, 0};
  const int32_t inf[4] = {kTtSize, complex_ ? 1 : 0, 0, 0};
  const int32_t idx[2] = {l, m};
  std::memcpy(head + 8, ver, sizeof(ver));
  std::memcpy(head + 16, inf, sizeof(inf));
  std::memcpy(head + 96, idx, sizeof(idx));   // i(1:2), after the 64-byte comment
  File f;
  f.fp = std::fopen(path, "wb");
  if (f.fp == nullptr) {
    const int e = errno;
    set_error("sdv: cannot create %s: %s", path, std::strerror(e));
    return e > 0 ? e : kBadArg;
  }
  const int32_t lm[2] = {l, m};
  bool ok = std::fwrite(head, 1, sizeof(head), f.fp) == sizeof(head) &&
            std::fwrite(lm, sizeof(int32_t), 2, f.fp) == 2 &&
            std::fwrite(n, sizeof(int32_t), (size_t)d, f.fp) == (size_t)d &&
            std::fwrite(r, sizeof(int32_t), (size_t)d + 1, f.fp) == (size_t)d + 1;
  const size_t elem = complex_ ? 16 : 8;
  if (ok && count > 0) ok = std::fwrite(cores, elem, (size_t)count, f.fp) == (size_t)count;
  int e = ok ? 0 : errno;
  if (std::fclose(f.fp) != 0) {
    if (ok) e = errno;
    ok = false;
  }
  f.fp = nullptr;
  if (!ok) {
    set_error("sdv: writing %s failed: %s", path, std::strerror(e));
    return e > 0 ? e : kShort;
  }
  return 0;
}

}  // namespace
}  // namespace ttf

extern "C" {

int ttf_sdv_dims(const char* path, int dcap, int* l, int* m, int* n, int* r, int64_t* count, int* is_complex) {
  ttf::File f;
  ttf::Dims dm;
  const int st = ttf::open_dims(path, f, dm);
  if (st != 0) return st;
  const int d = (int)dm.n.size();
  if (l) *l = dm.l;
  if (m) *m = dm.m;
  if (count) *count = dm.count;
  if (is_complex) *is_complex = dm.complex_ ? 1 : 0;
  if (n != nullptr || r != nullptr) {
    if (n == nullptr || r == nullptr || dcap < d) {
      ttf::set_error("sdv: %s has %d cores; n and r need room for %d and %d entries (dcap %d)",
                     path, d, d, d + 1, dcap);
      return ttf::kBadArg;
    }
    std::memcpy(n, dm.n.data(), sizeof(int32_t) * d);
    std::memcpy(r, dm.r.data(), sizeof(int32_t) * (d + 1));
  }
  return 0;
}

int ttf_dtt_read(const char* path, int64_t cap, double* cores) {
  return ttf::read_cores(path, false, cap, cores);
}

int ttf_ztt_read(const char* path, int64_t cap, ttf_zdouble* cores) {
  return ttf::read_cores(path, true, cap, cores);
}

int ttf_dtt_write(const char* path, int l, int m, const int* n, const int* r, const double* cores) {
  return ttf::write_file(path, l, m, n, r, cores, false);
}

int ttf_ztt_write(const char* path, int l, int m, const int* n, const int* r, const ttf_zdouble* cores) {
  return ttf::write_file(path, l, m, n, r, cores, true);
}

}  // extern "C"

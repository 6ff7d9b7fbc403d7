// snapshot.hip — GALX v1 snapshots (galax-io/src/lib.rs:5-92), host code.
//
// Layout, little-endian: the 4 bytes "GALX", u32 version (1), u64 n, f64
// timestamp (seconds since the Unix epoch), then n f64 each of x, y, vx, vy,
// mass, ax, ay. Errors map SnapshotError (galax-io/src/lib.rs:8-18): Io ->
// GALAX_ERR_IO, BadMagic and a short file -> GALAX_ERR_FORMAT,
// UnsupportedVersion -> GALAX_ERR_UNSUPPORTED.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"

using namespace gx;

namespace {

constexpr char kMagic[4] = {'G', 'A', 'L', 'X'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 4 + 4 + 8 + 8;

static_assert(sizeof(double) == 8 && sizeof(uint64_t) == 8, "GALX stores 8-byte values");

// the host is little-endian (x86-64), as the format is: values are copied as is
struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

int read_header(FILE* f, const char* path, uint64_t* n, double* ts) {
  unsigned char h[kHeader];
  const size_t got = fread(h, 1, kHeader, f);
  if (got < 4 && ferror(f)) return set_error(GALAX_ERR_IO, "I/O error reading %s", path);
  if (got < 4) return set_error(GALAX_ERR_FORMAT, "unexpected end of file in %s", path);
  if (std::memcmp(h, kMagic, 4) != 0)
    return set_error(GALAX_ERR_FORMAT, "invalid magic bytes (not a galax snapshot): %s", path);
  if (got < 8) return set_error(GALAX_ERR_FORMAT, "unexpected end of file in %s", path);
  uint32_t ver;
  std::memcpy(&ver, h + 4, 4);
  if (ver != kVersion) return set_error(GALAX_ERR_UNSUPPORTED, "unsupported snapshot version: %u", ver);
  if (got < kHeader) return set_error(GALAX_ERR_FORMAT, "unexpected end of file in %s", path);
  std::memcpy(n, h + 8, 8);
  std::memcpy(ts, h + 16, 8);
  return 0;
}

}  // namespace

extern "C" {

int galax_snapshot_write(const char* path, const galax_bodies_t* b, double timestamp) {
  if (!path || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  File fh;
  fh.f = fopen(path, "wb");
  if (!fh.f) return set_error(GALAX_ERR_IO, "cannot create %s", path);
  const uint64_t n = b->n;
  unsigned char h[kHeader];
  std::memcpy(h, kMagic, 4);
  std::memcpy(h + 4, &kVersion, 4);
  std::memcpy(h + 8, &n, 8);
  std::memcpy(h + 16, &timestamp, 8);
  if (fwrite(h, 1, kHeader, fh.f) != kHeader) return set_error(GALAX_ERR_IO, "write to %s failed", path);
  // write_snapshot's array order (galax-io/src/lib.rs:34-40); a NULL array is written as zeros
  const double* arrs[7] = {b->x, b->y, b->vx, b->vy, b->m, b->ax, b->ay};
  std::vector<double> zeros;
  for (const double* a : arrs) {
    if (!a && n) {
      if (zeros.empty()) zeros.assign(n, 0.0);
      a = zeros.data();
    }
    if (n && fwrite(a, 8, n, fh.f) != n) return set_error(GALAX_ERR_IO, "write to %s failed", path);
  }
  FILE* f = fh.f;
  fh.f = nullptr;
  if (fclose(f) != 0) return set_error(GALAX_ERR_IO, "closing %s failed", path);
  return 0;
}

int galax_snapshot_info(const char* path, size_t* n, double* timestamp) {
  if (!path) return set_error(GALAX_ERR_INVALID, "NULL path");
  File fh;
  fh.f = fopen(path, "rb");
  if (!fh.f) return set_error(GALAX_ERR_IO, "cannot open %s", path);
  uint64_t nn = 0;
  double ts = 0.0;
  GX_CHECK(read_header(fh.f, path, &nn, &ts));
  if (n) *n = (size_t)nn;
  if (timestamp) *timestamp = ts;
  return 0;
}

int galax_snapshot_read(const char* path, galax_bodies_t* b) {
  if (!path || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  File fh;
  fh.f = fopen(path, "rb");
  if (!fh.f) return set_error(GALAX_ERR_IO, "cannot open %s", path);
  uint64_t n = 0;
  double ts = 0.0;
  GX_CHECK(read_header(fh.f, path, &n, &ts));
  if (n != (uint64_t)b->n)
    return set_error(GALAX_ERR_INVALID, "snapshot holds %llu bodies, arrays hold %zu", (unsigned long long)n, b->n);
  // the whole body block must be present (read_exact of each array)
  if (fseek(fh.f, 0, SEEK_END) != 0) return set_error(GALAX_ERR_IO, "seek in %s failed", path);
  const long size = ftell(fh.f);
  if (size < 0) return set_error(GALAX_ERR_IO, "seek in %s failed", path);
  if ((uint64_t)size < kHeader + 56 * n) return set_error(GALAX_ERR_FORMAT, "unexpected end of file in %s", path);
  if (fseek(fh.f, (long)kHeader, SEEK_SET) != 0) return set_error(GALAX_ERR_IO, "seek in %s failed", path);
  // read_snapshot's array order (galax-io/src/lib.rs:75-81); NULL arrays are skipped
  double* arrs[7] = {b->x, b->y, b->vx, b->vy, b->m, b->ax, b->ay};
  for (double* a : arrs) {
    if (!n) break;
    if (a) {
      if (fread(a, 8, n, fh.f) != n)
        return set_error(ferror(fh.f) ? GALAX_ERR_IO : GALAX_ERR_FORMAT, "unexpected end of file in %s", path);
    } else if (fseek(fh.f, (long)(8 * n), SEEK_CUR) != 0) {
      return set_error(GALAX_ERR_IO, "seek in %s failed", path);
    }
  }
  return 0;
}

}  // extern "C"

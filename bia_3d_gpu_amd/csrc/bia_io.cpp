// bia_io.cpp -- the reference's record files for C and C++ callers (host code only).
//
// Fortran unformatted sequential files with 4-byte little-endian record markers, the layout the
// reference's programs read and write (gfortran/amdflang default):
//   .bbb  one record per realisation, 6N doubles particle-major (x, y, z, u, v, w)
//         (IC/bbb_maker_REV.f90:71; the N_calc loop of bia/bia.f90:62-64 reads one per pass);
//   .DIM  [t, 6N] (dim/dim_3D.f90:32);  .BIA  [dt, 6N deltas] (bia/bia.f90:206).
// nhead is the number of leading doubles of a record (0 for .bbb, 1 for .DIM/.BIA).  Arrays on
// the caller's side are the step entry points' structure-of-arrays layout: six rows of stride ld.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/bia_hip.h"

namespace {

// a record's payload must hold nhead doubles and whole particles of 6 doubles; the marker is a
// signed 32-bit byte count
bool payload_particles(int64_t bytes, int nhead, int64_t* n) {
  const int64_t head = 8 * (int64_t)nhead;
  if (bytes < head || (bytes - head) % 48 != 0) return false;
  *n = (bytes - head) / 48;
  return true;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

bool read_marker(FILE* f, int32_t* v) {
  unsigned char b[4];
  if (fread(b, 1, 4, f) != 4) return false;
  *v = (int32_t)((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24));
  return true;
}

bool write_marker(FILE* f, int32_t v) {
  const uint32_t u = (uint32_t)v;
  const unsigned char b[4] = {(unsigned char)u, (unsigned char)(u >> 8), (unsigned char)(u >> 16),
                              (unsigned char)(u >> 24)};
  return fwrite(b, 1, 4, f) == 4;
}

// walks the records of an open file of `size` bytes; for each, calls visit(index, offset of the
// payload, payload bytes), which returns false to stop.  Returns BIA_ERR_IO on a marker that is
// negative, does not fit the file or does not match the closing one.
template <class Visit>
int walk(FILE* f, int64_t size, Visit visit) {
  int64_t p = 0, k = 0;
  while (p < size) {
    int32_t n1, n2;
    if (p + 4 > size || fseeko(f, (off_t)p, SEEK_SET) != 0 || !read_marker(f, &n1)) return BIA_ERR_IO;
    if (n1 < 0 || p + 8 + (int64_t)n1 > size) return BIA_ERR_IO;
    if (fseeko(f, (off_t)(p + 4 + n1), SEEK_SET) != 0 || !read_marker(f, &n2) || n2 != n1) return BIA_ERR_IO;
    if (!visit(k, p + 4, (int64_t)n1)) return BIA_OK;
    p += 8 + (int64_t)n1;
    ++k;
  }
  return BIA_OK;
}

int open_sized(const char* path, File& F, int64_t* size) {
  F.f = fopen(path, "rb");
  if (!F.f) return BIA_ERR_IO;
  if (fseeko(F.f, 0, SEEK_END) != 0) return BIA_ERR_IO;
  const off_t s = ftello(F.f);
  if (s < 0) return BIA_ERR_IO;
  *size = (int64_t)s;
  return BIA_OK;
}

}  // namespace

extern "C" {

int bia_records_scan(const char* path, int nhead, long long* nrec, long long* nmax) {
  if (!path || nhead < 0 || !nrec) return BIA_ERR_ARG;
  File F;
  int64_t size = 0;
  if (int s = open_sized(path, F, &size)) return s;
  int64_t count = 0, most = 0;
  bool shape_ok = true;
  const int s = walk(F.f, size, [&](int64_t, int64_t, int64_t bytes) {
    int64_t n;
    if (!payload_particles(bytes, nhead, &n)) {
      shape_ok = false;
      return false;
    }
    ++count;
    if (n > most) most = n;
    return true;
  });
  if (s != BIA_OK) return s;
  if (!shape_ok) return BIA_ERR_IO;
  *nrec = (long long)count;
  if (nmax) *nmax = (long long)most;
  return BIA_OK;
}

int bia_record_read(const char* path, int nhead, long long rec, double* head, double* soa, long long ld,
                    long long* n) {
  if (!path || nhead < 0 || rec < 0 || !n || (nhead > 0 && soa && !head)) return BIA_ERR_ARG;
  File F;
  int64_t size = 0;
  if (int s = open_sized(path, F, &size)) return s;
  int64_t off = -1, bytes = 0;
  int s = walk(F.f, size, [&](int64_t k, int64_t o, int64_t b) {
    if (k != rec) return true;
    off = o;
    bytes = b;
    return false;
  });
  if (s != BIA_OK) return s;
  if (off < 0) return BIA_ERR_ARG;  // fewer records than rec + 1
  int64_t np;
  if (!payload_particles(bytes, nhead, &np)) return BIA_ERR_IO;
  *n = (long long)np;
  if (!soa) return BIA_OK;  // size query
  if (ld < np) return BIA_ERR_ARG;
  std::vector<double> buf((size_t)(bytes / 8));
  if (fseeko(F.f, (off_t)off, SEEK_SET) != 0 || fread(buf.data(), 8, buf.size(), F.f) != buf.size()) return BIA_ERR_IO;
  for (int h = 0; h < nhead; h++) head[h] = buf[(size_t)h];
  const double* aos = buf.data() + nhead;
  for (int64_t i = 0; i < np; i++)
    for (int c = 0; c < 6; c++) soa[(size_t)c * (size_t)ld + (size_t)i] = aos[(size_t)(6 * i + c)];
  return BIA_OK;
}

int bia_record_write(const char* path, int append, int nhead, const double* head, const double* soa, long long ld,
                     long long n) {
  if (!path || nhead < 0 || n < 0 || ld < n || (n > 0 && !soa) || (nhead > 0 && !head) || (append != 0 && append != 1))
    return BIA_ERR_ARG;
  const int64_t bytes = 8 * ((int64_t)nhead + 6 * (int64_t)n);
  if (bytes > INT32_MAX) return BIA_ERR_ARG;  // a 4-byte marker holds at most 2^31 - 1 bytes
  std::vector<double> buf((size_t)(bytes / 8));
  for (int h = 0; h < nhead; h++) buf[(size_t)h] = head[h];
  double* aos = buf.data() + nhead;
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 6; c++) aos[(size_t)(6 * i + c)] = soa[(size_t)c * (size_t)ld + (size_t)i];
  File F;
  F.f = fopen(path, append ? "ab" : "wb");
  if (!F.f) return BIA_ERR_IO;
  if (!write_marker(F.f, (int32_t)bytes) || fwrite(buf.data(), 8, buf.size(), F.f) != buf.size() ||
      !write_marker(F.f, (int32_t)bytes))
    return BIA_ERR_IO;
  if (fclose(F.f) != 0) {
    F.f = nullptr;
    return BIA_ERR_IO;
  }
  F.f = nullptr;
  return BIA_OK;
}

}  // extern "C"

// The reference driver's files, read and written without Fortran I/O:
// state files init.bin / temp.bin (Fortran unformatted sequential: each
// record framed by its byte count as a 4-byte little-endian integer before
// and after; src/shwwnl.f:180-187 reads 7 records, save_temp
// src/shwwnl_adap.f:703-724 writes them) and the mesh file cm.dat
// (list-directed text, src/cmal_s.f:7-78).  Host only.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/spoc_abi.h"
#include "error.h"

namespace {

struct File {
  FILE *f;
  explicit File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// One record of n bytes into dst; SPOC_EFILE if the framing differs.
int read_record(FILE *f, void *dst, uint64_t n) {
  int32_t head = 0, tail = 0;
  if (n > INT32_MAX) return SPOC_EFILE;  // (records of 2 GiB or more are split by
                                         // the Fortran runtimes; not supported)
  if (std::fread(&head, 4, 1, f) != 1 || (uint64_t)head != n) return SPOC_EFILE;
  if (n && std::fread(dst, 1, n, f) != n) return SPOC_EFILE;
  if (std::fread(&tail, 4, 1, f) != 1 || tail != head) return SPOC_EFILE;
  return SPOC_OK;
}

int write_record(FILE *f, const void *src, uint64_t n) {
  if (n > INT32_MAX) return SPOC_EFILE;
  const int32_t m = (int32_t)n;
  if (std::fwrite(&m, 4, 1, f) != 1 || (n && std::fwrite(src, 1, n, f) != n) ||
      std::fwrite(&m, 4, 1, f) != 1)
    return SPOC_EFILE;
  return SPOC_OK;
}

// Next list-directed value (blanks, commas and line ends separate values;
// Fortran's D exponent accepted).
bool next_token(FILE *f, char *buf, size_t cap) {
  int c;
  do c = std::fgetc(f);
  while (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ',');
  if (c == EOF) return false;
  size_t n = 0;
  while (c != EOF && c != ' ' && c != '\t' && c != '\n' && c != '\r' && c != ',') {
    if (n + 1 < cap) buf[n++] = (c == 'D' || c == 'd') ? 'e' : (char)c;
    c = std::fgetc(f);
  }
  buf[n] = 0;
  return n > 0;
}

bool read_int(FILE *f, int *v) {
  char b[64], *end;
  if (!next_token(f, b, sizeof b)) return false;
  errno = 0;
  long x = std::strtol(b, &end, 10);
  if (*end || errno || x < INT32_MIN || x > INT32_MAX) return false;
  *v = (int)x;
  return true;
}

bool read_real(FILE *f, double *v) {
  char b[96], *end;
  if (!next_token(f, b, sizeof b)) return false;
  *v = std::strtod(b, &end);
  return *end == 0;
}

}  // namespace

extern "C" {

int spoc_state_file_read(const char *path, int nm, int ne, int ld, double *time, double *u0,
                         double *v0, double *h0, double *w0, double *z0, double *hb) {
  return spoc::guarded([&]() -> int {
    if (!path || nm <= 0 || ne <= 0 || ld < nm || !time)
      return spoc::report(SPOC_EINVAL, "spoc_state_file_read: bad arguments");
    File F(path, "rb");
    if (!F.f) return spoc::report(SPOC_EFILE, "%s: %s", path, std::strerror(errno));
    if (read_record(F.f, time, 8))
      return spoc::report(SPOC_EFILE, "%s: record 1 is not one 8-byte time", path);
    std::vector<double> buf((size_t)nm * ne);
    double *out[6] = {u0, v0, h0, w0, z0, hb};
    int rec = 2;
    for (double *o : out) {
      if (read_record(F.f, buf.data(), 8 * (uint64_t)buf.size()))
        return spoc::report(SPOC_EFILE, "%s: record %d is not %d x %d doubles", path, rec, nm, ne);
      rec++;
      if (!o) continue;
      for (int e = 0; e < ne; e++)
        for (int m = 0; m < nm; m++) o[(size_t)ld * e + m] = buf[(size_t)nm * e + m];
    }
    return SPOC_OK;
  });
}

int spoc_state_file_write(const char *path, int nm, int ne, int ld, double time,
                          const double *u0, const double *v0, const double *h0, const double *w0,
                          const double *z0, const double *hb) {
  return spoc::guarded([&]() -> int {
    const double *in[6] = {u0, v0, h0, w0, z0, hb};
    if (!path || nm <= 0 || ne <= 0 || ld < nm)
      return spoc::report(SPOC_EINVAL, "spoc_state_file_write: bad arguments");
    for (const double *a : in)
      if (!a) return spoc::report(SPOC_EINVAL, "spoc_state_file_write: null field");
    File F(path, "wb");
    if (!F.f) return spoc::report(SPOC_EFILE, "%s: %s", path, std::strerror(errno));
    if (write_record(F.f, &time, 8)) return spoc::report(SPOC_EFILE, "%s: write failed", path);
    std::vector<double> buf((size_t)nm * ne);
    for (const double *a : in) {
      for (int e = 0; e < ne; e++)
        for (int m = 0; m < nm; m++) buf[(size_t)nm * e + m] = a[(size_t)ld * e + m];
      if (write_record(F.f, buf.data(), 8 * (uint64_t)buf.size()))
        return spoc::report(SPOC_EFILE, "%s: write failed (records of 2 GiB or more are not "
                            "supported)", path);
    }
    return std::fflush(F.f) == 0 ? SPOC_OK : spoc::report(SPOC_EFILE, "%s: write failed", path);
  });
}

int spoc_mesh_file_size(const char *path, int *nn, int *ne) {
  return spoc::guarded([&]() -> int {
    if (!path || !nn || !ne) return spoc::report(SPOC_EINVAL, "spoc_mesh_file_size: bad arguments");
    File F(path, "r");
    if (!F.f) return spoc::report(SPOC_EFILE, "%s: %s", path, std::strerror(errno));
    if (!read_int(F.f, nn) || !read_int(F.f, ne) || *nn <= 0 || *ne <= 0)
      return spoc::report(SPOC_EFILE, "%s: no \"nn ne\" header", path);
    return SPOC_OK;
  });
}

int spoc_mesh_file_read(const char *path, int nn, int ne, double scx, double scy, double *x,
                        double *y, int *in) {
  return spoc::guarded([&]() -> int {
    if (!path || nn <= 0 || ne <= 0 || !x || !y || !in)
      return spoc::report(SPOC_EINVAL, "spoc_mesh_file_read: bad arguments");
    File F(path, "r");
    if (!F.f) return spoc::report(SPOC_EFILE, "%s: %s", path, std::strerror(errno));
    int a, b, idx;
    if (!read_int(F.f, &a) || !read_int(F.f, &b) || a != nn || b != ne)
      return spoc::report(SPOC_EFILE, "%s: header is not \"%d %d\"", path, nn, ne);
    for (int i = 0; i < nn; i++)  // the index column is read and ignored, as the reference does
      if (!read_int(F.f, &idx) || !read_real(F.f, &x[i]) || !read_real(F.f, &y[i]))
        return spoc::report(SPOC_EFILE, "%s: node line %d unreadable", path, i + 1);
    for (int e = 0; e < ne; e++)
      if (!read_int(F.f, &idx) || !read_int(F.f, &in[3 * e]) || !read_int(F.f, &in[3 * e + 1]) ||
          !read_int(F.f, &in[3 * e + 2]))
        return spoc::report(SPOC_EFILE, "%s: triangle line %d unreadable", path, e + 1);
    for (int i = 0; i < nn; i++) {  // src/cmal_s.f:31-34
      x[i] = x[i] * scx;
      y[i] = y[i] * scy;
    }
    return SPOC_OK;
  });
}

int spoc_mesh_file_write(const char *path, int nn, int ne, const double *x, const double *y,
                         const int *in, double scx, double scy) {
  return spoc::guarded([&]() -> int {
    if (!path || nn <= 0 || ne <= 0 || !x || !y || !in)
      return spoc::report(SPOC_EINVAL, "spoc_mesh_file_write: bad arguments");
    File F(path, "w");
    if (!F.f) return spoc::report(SPOC_EFILE, "%s: %s", path, std::strerror(errno));
    std::fprintf(F.f, "%d %d\n", nn, ne);
    for (int i = 0; i < nn; i++)  // 17 significant digits: read back to the same double
      std::fprintf(F.f, "%d %.17g %.17g\n", i + 1, x[i] / scx, y[i] / scy);
    for (int e = 0; e < ne; e++)
      std::fprintf(F.f, "%d %d %d %d\n", e + 1, in[3 * e], in[3 * e + 1], in[3 * e + 2]);
    return std::fflush(F.f) == 0 ? SPOC_OK : spoc::report(SPOC_EFILE, "%s: write failed", path);
  });
}

}  // extern "C"

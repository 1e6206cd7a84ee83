// Wavefunction snapshots in the reference's two formats (host code only).
//
// printoutc.f90:69-114 writes, after `#` comment lines (skipped on input by
// titols.f90:12-37), the list-directed records
//   xmax ymax zmax hx hy hz nx ny nz / rimp(3) / vimp(3) / ninvar /
//   invar(ninvar) (complex) / psi(nx,ny,nz) (complex, column-major)
// (formatted, iprint > 0) or the same six records unformatted (iprint < 0);
// readenc.f90:151-168 reads the formatted form back with list-directed READ.
// The reader here scans the text once with strtod: complex values as
// "(re,im)", repeat counts "k*value" / "k*(re,im)", Fortran `D` exponents, any
// whitespace or comma separation.  The unformatted form is sequential records
// with 4-byte length markers and 4-byte default integers.
#pragma once
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/effpot.h"

namespace snap {
namespace {

thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

inline bool slurp(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) { std::fclose(f); return false; }
  buf.resize((size_t)n);
  size_t got = n ? std::fread(&buf[0], 1, (size_t)n, f) : 0;
  std::fclose(f);
  return got == (size_t)n;
}

// List-directed scanner over the text after the comment lines.
struct Scanner {
  const char* p;
  const char* end;
  long repeat = 0;   // pending copies of the last value(s)
  double last[2] = {0.0, 0.0};
  char tmp[64];

  static bool sep(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ','; }
  void skip() { while (p < end && sep(*p)) p++; }
  // one real number (a D exponent is read as E)
  bool number(double& v) {
    skip();
    const char* s = p;
    size_t n = 0;
    while (p < end && !sep(*p) && *p != ')' && *p != '(' && *p != '*' && n < sizeof(tmp) - 1) {
      char c = *p++;
      tmp[n++] = (c == 'D' || c == 'd') ? 'E' : c;
    }
    if (n == 0) { p = s; return false; }
    tmp[n] = 0;
    char* e = nullptr;
    errno = 0;
    v = std::strtod(tmp, &e);
    return e == tmp + n && errno != ERANGE;
  }
  // optional "k*" in front of the next value
  long count() {
    skip();
    const char* s = p;
    long k = 0;
    while (p < end && *p >= '0' && *p <= '9') k = 10 * k + (*p++ - '0');
    if (p < end && *p == '*' && p > s) { p++; return k; }
    p = s;
    return 1;
  }
  bool real(double& v) {
    if (repeat > 0) { repeat--; v = last[0]; return true; }
    long k = count();
    if (!number(v)) return false;
    last[0] = v;
    repeat = k - 1;
    return true;
  }
  bool cplx(double& re, double& im) {
    if (repeat > 0) { repeat--; re = last[0]; im = last[1]; return true; }
    long k = count();
    skip();
    if (p >= end || *p != '(') return false;
    p++;
    if (!number(re)) return false;
    skip();
    if (!number(im)) return false;
    skip();
    if (p >= end || *p != ')') return false;
    p++;
    last[0] = re; last[1] = im;
    repeat = k - 1;
    return true;
  }
  bool integer(int32_t& v) {
    double d;
    if (!real(d) || d != (double)(int32_t)d) return false;
    v = (int32_t)d;
    return true;
  }
};

inline int read_text(const std::string& buf, effpot_snapshot* h, double* psi) {
  size_t pos = 0;
  while (pos < buf.size() && buf[pos] == '#') {  // titols: leading comment lines
    size_t nl = buf.find('\n', pos);
    pos = nl == std::string::npos ? buf.size() : nl + 1;
  }
  Scanner sc{buf.data() + pos, buf.data() + buf.size()};
  double hd[6];
  for (double& v : hd) if (!sc.real(v)) return fail(EFFPOT_ERR_IO, "snapshot: bad grid record");
  if (!sc.integer(h->nx) || !sc.integer(h->ny) || !sc.integer(h->nz)) return fail(EFFPOT_ERR_IO, "snapshot: bad grid sizes");
  h->xmax = hd[0]; h->ymax = hd[1]; h->zmax = hd[2]; h->hx = hd[3]; h->hy = hd[4]; h->hz = hd[5];
  for (double& v : h->rimp) if (!sc.real(v)) return fail(EFFPOT_ERR_IO, "snapshot: bad rimp record");
  for (double& v : h->vimp) if (!sc.real(v)) return fail(EFFPOT_ERR_IO, "snapshot: bad vimp record");
  if (!sc.integer(h->ninvar) || h->ninvar < 0 || h->ninvar > EFFPOT_SNAPSHOT_MAX_INVAR)
    return fail(EFFPOT_ERR_IO, "snapshot: bad ninvar");
  for (int k = 0; k < h->ninvar; k++)
    if (!sc.cplx(h->invar[2 * k], h->invar[2 * k + 1])) return fail(EFFPOT_ERR_IO, "snapshot: bad invar record");
  if (h->nx < 1 || h->ny < 1 || h->nz < 1) return fail(EFFPOT_ERR_IO, "snapshot: bad grid sizes");
  if (!psi) return EFFPOT_OK;
  const long n = (long)h->nx * h->ny * h->nz;
  for (long i = 0; i < n; i++)
    if (!sc.cplx(psi[2 * i], psi[2 * i + 1]))
      return fail(EFFPOT_ERR_IO, "snapshot: psi ends after " + std::to_string(i) + " of " + std::to_string(n) + " values");
  return EFFPOT_OK;
}

// Sequential unformatted records: int32 length, payload, int32 length.
struct Records {
  const std::string& b;
  size_t pos = 0;
  bool next(const char** data, size_t* len) {
    if (pos + 4 > b.size()) return false;
    int32_t n;
    std::memcpy(&n, b.data() + pos, 4);
    if (n < 0 || pos + 8 + (size_t)n > b.size()) return false;
    int32_t m;
    std::memcpy(&m, b.data() + pos + 4 + n, 4);
    if (m != n) return false;
    *data = b.data() + pos + 4;
    *len = (size_t)n;
    pos += 8 + (size_t)n;
    return true;
  }
};

inline bool looks_binary(const std::string& b) {
  if (b.size() < 8 || b[0] == '#') return false;
  int32_t n;
  std::memcpy(&n, b.data(), 4);
  return n == 6 * 8 + 3 * 4 && b.size() >= 8 + (size_t)n;
}

inline int read_binary(const std::string& buf, effpot_snapshot* h, double* psi) {
  Records r{buf};
  const char* d;
  size_t n;
  if (!r.next(&d, &n) || n != 60) return fail(EFFPOT_ERR_IO, "snapshot: bad grid record");
  double hd[6];
  std::memcpy(hd, d, 48);
  std::memcpy(&h->nx, d + 48, 4); std::memcpy(&h->ny, d + 52, 4); std::memcpy(&h->nz, d + 56, 4);
  h->xmax = hd[0]; h->ymax = hd[1]; h->zmax = hd[2]; h->hx = hd[3]; h->hy = hd[4]; h->hz = hd[5];
  if (!r.next(&d, &n) || n != 24) return fail(EFFPOT_ERR_IO, "snapshot: bad rimp record");
  std::memcpy(h->rimp, d, 24);
  if (!r.next(&d, &n) || n != 24) return fail(EFFPOT_ERR_IO, "snapshot: bad vimp record");
  std::memcpy(h->vimp, d, 24);
  if (!r.next(&d, &n) || n != 4) return fail(EFFPOT_ERR_IO, "snapshot: bad ninvar record");
  std::memcpy(&h->ninvar, d, 4);
  if (h->ninvar < 0 || h->ninvar > EFFPOT_SNAPSHOT_MAX_INVAR) return fail(EFFPOT_ERR_IO, "snapshot: bad ninvar");
  if (!r.next(&d, &n) || n != (size_t)h->ninvar * 16) return fail(EFFPOT_ERR_IO, "snapshot: bad invar record");
  std::memcpy(h->invar, d, n);
  if (h->nx < 1 || h->ny < 1 || h->nz < 1) return fail(EFFPOT_ERR_IO, "snapshot: bad grid sizes");
  if (!psi) return EFFPOT_OK;
  const size_t want = (size_t)h->nx * h->ny * h->nz * 16;
  if (!r.next(&d, &n) || n != want) return fail(EFFPOT_ERR_IO, "snapshot: bad psi record");
  std::memcpy(psi, d, n);
  return EFFPOT_OK;
}

inline int read(const char* path, effpot_snapshot* h, double* psi) {
  if (!path || !h) return fail(EFFPOT_ERR_ARG, "snapshot: null argument");
  std::memset(h, 0, sizeof(*h));
  std::string buf;
  if (!slurp(path, buf)) return fail(EFFPOT_ERR_IO, std::string("snapshot: cannot read ") + path);
  if (looks_binary(buf)) {
    h->binary = 1;
    return read_binary(buf, h, psi);
  }
  return read_text(buf, h, psi);
}

inline void put_rec(FILE* f, const void* data, int32_t n) {
  std::fwrite(&n, 4, 1, f);
  if (n) std::fwrite(data, 1, (size_t)n, f);
  std::fwrite(&n, 4, 1, f);
}

inline int write(const char* path, const effpot_snapshot* h, const double* psi, int binary) {
  if (!path || !h || !psi) return fail(EFFPOT_ERR_ARG, "snapshot: null argument");
  if (h->nx < 1 || h->ny < 1 || h->nz < 1 || h->ninvar < 0 || h->ninvar > EFFPOT_SNAPSHOT_MAX_INVAR)
    return fail(EFFPOT_ERR_ARG, "snapshot: bad header");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(EFFPOT_ERR_IO, std::string("snapshot: cannot write ") + path);
  const long n = (long)h->nx * h->ny * h->nz;
  if (binary) {
    char g[60];
    const double hd[6] = {h->xmax, h->ymax, h->zmax, h->hx, h->hy, h->hz};
    std::memcpy(g, hd, 48);
    std::memcpy(g + 48, &h->nx, 4); std::memcpy(g + 52, &h->ny, 4); std::memcpy(g + 56, &h->nz, 4);
    put_rec(f, g, 60);
    put_rec(f, h->rimp, 24);
    put_rec(f, h->vimp, 24);
    put_rec(f, &h->ninvar, 4);
    put_rec(f, h->invar, h->ninvar * 16);
    if (n * 16 > 0x7fffffffL) { std::fclose(f); return fail(EFFPOT_ERR_ARG, "snapshot: psi record over 2 GiB"); }
    put_rec(f, psi, (int32_t)(n * 16));
  } else {
    std::fprintf(f, "# wavefunction snapshot (psi complex(8), nx ny nz column-major)\n");
    std::fprintf(f, " %.17g %.17g %.17g %.17g %.17g %.17g %d %d %d\n", h->xmax, h->ymax, h->zmax, h->hx, h->hy,
                 h->hz, h->nx, h->ny, h->nz);
    std::fprintf(f, " %.17g %.17g %.17g\n", h->rimp[0], h->rimp[1], h->rimp[2]);
    std::fprintf(f, " %.17g %.17g %.17g\n", h->vimp[0], h->vimp[1], h->vimp[2]);
    std::fprintf(f, " %d\n", h->ninvar);
    for (int k = 0; k < h->ninvar; k++) std::fprintf(f, " (%.17g,%.17g)", h->invar[2 * k], h->invar[2 * k + 1]);
    std::fprintf(f, "\n");
    for (long i = 0; i < n; i++)
      std::fprintf(f, (i % 4 == 3 || i == n - 1) ? " (%.17g,%.17g)\n" : " (%.17g,%.17g)", psi[2 * i], psi[2 * i + 1]);
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return fail(EFFPOT_ERR_IO, std::string("snapshot: write failed: ") + path);
  return EFFPOT_OK;
}

}  // namespace
}  // namespace snap

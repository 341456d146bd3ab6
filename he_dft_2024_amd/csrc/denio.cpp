// denio.cpp — density and wave-function files of the BCN-4He-DFT drivers.
//
// Formats (text, Fortran list-directed):
//   * leading comment lines whose first column is the comment character
//     (cchar, '#' in the examples) are skipped, as titols.f90:12-42 does;
//   * statics density (He_statics/readen.f90:74-84, printout.f90:52-86):
//       xmax ymax zmax hx hy hz nx ny nz limp ximp yimp zimp
//       den(nx,ny,nz)                          (x fastest)
//   * TDDFT wave function (Dynamics_TDDFT/readenc.f90:94-110,
//     printoutc.f90:64-74):
//       xmax ymax zmax hx hy hz nx ny nz
//       N_imp
//       rimp(N_imp,3)
//       vimp(N_imp,3)
//       psi(nx,ny,nz)                          complex: (re,im)
// The reader accepts what a Fortran list-directed READ accepts for these
// records: values separated by blanks, commas or line ends, repeat counts
// r*value, D or E exponents, logicals T/F/.TRUE./.FALSE., complex (re,im).
// Parsing runs over the whole file in memory and splits the value stream
// over host threads.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hedft.h"

namespace {

struct Text {
  std::vector<char> buf;
  size_t pos = 0;
};

int load(const char* path, Text& t) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return HEDFT_EINVAL;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  t.buf.resize((size_t)std::max(sz, 0L) + 1);
  const size_t got = sz > 0 ? std::fread(t.buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if ((long)got != sz) return HEDFT_EINVAL;
  t.buf[got] = '\0';
  t.pos = 0;
  return HEDFT_OK;
}

// skip the leading lines that start with cchar (titols)
void skip_comments(Text& t, char cchar) {
  while (t.pos < t.buf.size() - 1 && t.buf[t.pos] == cchar) {
    while (t.pos < t.buf.size() - 1 && t.buf[t.pos] != '\n') ++t.pos;
    if (t.pos < t.buf.size() - 1) ++t.pos;
  }
}

inline bool is_sep(char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n' || ch == ','; }

// one list-directed item at p: returns false at the end of the text.  A value
// may carry a repeat count (r*value); complex items are handled by the caller.
struct Item {
  const char* s;
  size_t len;
  long rep;
};

bool next_item(const char*& p, const char* end, Item& it) {
  while (p < end && is_sep(*p)) ++p;
  if (p >= end || *p == '\0') return false;
  const char* s = p;
  // up to the next separator; a parenthesised complex constant (possibly after
  // a repeat count r*) runs to its closing parenthesis, commas included
  while (p < end && !is_sep(*p) && *p != '\0') {
    if (*p == '(') {
      while (p < end && *p != ')') ++p;
      if (p < end) ++p;
      break;
    }
    ++p;
  }
  it.s = s;
  it.len = (size_t)(p - s);
  it.rep = 1;
  const char* star = (const char*)std::memchr(s, '*', it.len);
  if (star && *s != '(' && (star < s + it.len)) {
    char* e = nullptr;
    const long r = std::strtol(s, &e, 10);
    if (e == star && r > 0) {
      it.rep = r;
      it.len -= (size_t)(star + 1 - s);
      it.s = star + 1;
    }
  }
  return true;
}

bool parse_real(const char* s, size_t len, double& v) {
  char tmp[64];
  if (len == 0 || len >= sizeof tmp) return false;
  for (size_t i = 0; i < len; ++i) tmp[i] = (s[i] == 'D' || s[i] == 'd') ? 'E' : s[i];
  tmp[len] = '\0';
  char* e = nullptr;
  errno = 0;
  v = std::strtod(tmp, &e);
  return e == tmp + len;
}

bool parse_logical(const char* s, size_t len, int& v) {
  size_t i = 0;
  if (i < len && s[i] == '.') ++i;
  if (i >= len) return false;
  const char ch = s[i];
  if (ch == 'T' || ch == 't') { v = 1; return true; }
  if (ch == 'F' || ch == 'f') { v = 0; return true; }
  return false;
}

bool parse_complex(const char* s, size_t len, double& re, double& im) {
  if (len < 5 || s[0] != '(' || s[len - 1] != ')') return false;
  const char* comma = (const char*)std::memchr(s, ',', len);
  if (!comma) return false;
  auto trim = [](const char*& a, const char*& b) {
    while (a < b && (*a == ' ' || *a == '\t' || *a == '\n' || *a == '\r')) ++a;
    while (b > a && (b[-1] == ' ' || b[-1] == '\t' || b[-1] == '\n' || b[-1] == '\r')) --b;
  };
  const char *a0 = s + 1, *b0 = comma, *a1 = comma + 1, *b1 = s + len - 1;
  trim(a0, b0);
  trim(a1, b1);
  return parse_real(a0, (size_t)(b0 - a0), re) && parse_real(a1, (size_t)(b1 - a1), im);
}

// reads `count` reals (or logicals as 0/1 when logical[k]) from the text
int read_scalars(Text& t, int count, double* out, const bool* logical) {
  const char* p = t.buf.data() + t.pos;
  const char* end = t.buf.data() + t.buf.size() - 1;
  int k = 0;
  Item it;
  while (k < count) {
    if (!next_item(p, end, it)) return HEDFT_EINVAL;
    for (long r = 0; r < it.rep && k < count; ++r, ++k) {
      if (logical && logical[k]) {
        int v;
        if (!parse_logical(it.s, it.len, v)) return HEDFT_EINVAL;
        out[k] = v;
      } else if (!parse_real(it.s, it.len, out[k])) {
        return HEDFT_EINVAL;
      }
    }
  }
  // a list-directed READ consumes the rest of its last record
  while (p < end && *p != '\n') ++p;
  t.pos = (size_t)(p - t.buf.data());
  return HEDFT_OK;
}

// reads n values (complex when cplx) from the rest of the text, in parallel
// over chunks that start at item boundaries
int read_array(Text& t, size_t n, double* out, bool cplx) {
  const char* beg = t.buf.data() + t.pos;
  const char* end = t.buf.data() + t.buf.size() - 1;
  const size_t bytes = (size_t)(end - beg);
  const int nt = (int)std::max<size_t>(1, std::min<size_t>({16, std::thread::hardware_concurrency(),
                                                             bytes / (1u << 20) + 1}));
  // chunk starts moved forward to the next separator (for complex: next '(')
  std::vector<const char*> cut(nt + 1);
  cut[0] = beg;
  cut[nt] = end;
  for (int i = 1; i < nt; ++i) {
    const char* q = beg + bytes * i / nt;
    if (cplx) {  // start of the item holding the next '(' (it may carry r*)
      while (q < end && *q != '(') ++q;
      while (q > beg && !is_sep(q[-1])) --q;
    } else {
      while (q < end && !is_sep(*q)) ++q;
    }
    cut[i] = std::max(q, cut[i - 1]);
  }
  // pass 1: count values per chunk; pass 2: parse into place
  std::vector<size_t> cnt(nt, 0);
  std::vector<int> bad(nt, 0);
  auto work = [&](int i, bool store, size_t off) {
    const char* p = cut[i];
    Item it;
    size_t k = off;
    while (next_item(p, cut[i + 1], it)) {
      if (!store) {
        k += (size_t)it.rep;
        continue;
      }
      double a = 0.0, b = 0.0;
      const bool ok = cplx ? parse_complex(it.s, it.len, a, b) : parse_real(it.s, it.len, a);
      if (!ok) {
        bad[i] = 1;
        return;
      }
      for (long r = 0; r < it.rep && k < n; ++r, ++k) {
        if (cplx) {
          out[2 * k] = a;
          out[2 * k + 1] = b;
        } else {
          out[k] = a;
        }
      }
    }
    if (!store) cnt[i] = k - off;
  };
  {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(work, i, false, 0);
    for (auto& x : th) x.join();
  }
  size_t total = 0;
  std::vector<size_t> off(nt);
  for (int i = 0; i < nt; ++i) {
    off[i] = total;
    total += cnt[i];
  }
  if (total < n) return HEDFT_EINVAL;  // too short (values beyond n are ignored, as READ does)
  {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i)
      if (off[i] < n) th.emplace_back(work, i, true, off[i]);
    for (auto& x : th) x.join();
  }
  for (int b : bad)
    if (b) return HEDFT_EINVAL;
  return HEDFT_OK;
}

// list-directed style output: full precision so that a read gives back the
// same doubles
int write_values(FILE* f, size_t n, const double* v, bool cplx) {
  char line[512];
  for (size_t i = 0; i < n;) {
    int len = 0;
    const size_t per = cplx ? 2 : 4;
    for (size_t j = 0; j < per && i < n; ++j, ++i) {
      if (cplx)
        len += std::snprintf(line + len, sizeof line - len, " (%.17E,%.17E)", v[2 * i], v[2 * i + 1]);
      else
        len += std::snprintf(line + len, sizeof line - len, " %.17E", v[i]);
    }
    line[len++] = '\n';
    if (std::fwrite(line, 1, (size_t)len, f) != (size_t)len) return HEDFT_EINVAL;
  }
  return HEDFT_OK;
}

int write_comment(FILE* f, const char* comment) {
  if (!comment) return HEDFT_OK;
  const char* p = comment;
  while (*p) {  // every line of the comment gets the '#' column
    const char* nl = std::strchr(p, '\n');
    const size_t len = nl ? (size_t)(nl - p) : std::strlen(p);
    std::fprintf(f, "#  %.*s\n", (int)len, p);
    p += len + (nl ? 1 : 0);
  }
  return HEDFT_OK;
}

}  // namespace

extern "C" {

int hedft_read_density(const char* path, char cchar, hedft_denfile* hdr, double* den, size_t n) {
  if (!path || !hdr) return HEDFT_EINVAL;
  Text t;
  int rc = load(path, t);
  if (rc != HEDFT_OK) return rc;
  skip_comments(t, cchar);
  double h[13];
  const bool lg[13] = {false, false, false, false, false, false, false, false, false, true, false, false, false};
  rc = read_scalars(t, 13, h, lg);
  if (rc != HEDFT_OK) return rc;
  hdr->xmax = h[0]; hdr->ymax = h[1]; hdr->zmax = h[2];
  hdr->hx = h[3]; hdr->hy = h[4]; hdr->hz = h[5];
  hdr->nx = (int32_t)h[6]; hdr->ny = (int32_t)h[7]; hdr->nz = (int32_t)h[8];
  hdr->limp = (int32_t)h[9];
  hdr->ximp = h[10]; hdr->yimp = h[11]; hdr->zimp = h[12];
  if (!den) return HEDFT_OK;  // header only
  if (n != (size_t)hdr->nx * hdr->ny * hdr->nz) return HEDFT_EGRID;
  return read_array(t, n, den, false);
}

int hedft_write_density(const char* path, const hedft_denfile* hdr, const double* den, size_t n,
                        const char* comment) {
  if (!path || !hdr || !den || n != (size_t)hdr->nx * hdr->ny * hdr->nz) return HEDFT_EINVAL;
  FILE* f = std::fopen(path, "wb");
  if (!f) return HEDFT_EINVAL;
  write_comment(f, comment);
  std::fprintf(f, " %.17E %.17E %.17E %.17E %.17E %.17E %d %d %d %c %.17E %.17E %.17E\n", hdr->xmax,
               hdr->ymax, hdr->zmax, hdr->hx, hdr->hy, hdr->hz, hdr->nx, hdr->ny, hdr->nz,
               hdr->limp ? 'T' : 'F', hdr->ximp, hdr->yimp, hdr->zimp);
  const int rc = write_values(f, n, den, false);
  return (std::fclose(f) == 0) ? rc : HEDFT_EINVAL;
}

int hedft_read_wavefunction(const char* path, char cchar, hedft_denfile* hdr, int32_t* nimp, double* rimp,
                            double* vimp, int32_t max_imp, double* psi, size_t n) {
  if (!path || !hdr || !nimp) return HEDFT_EINVAL;
  Text t;
  int rc = load(path, t);
  if (rc != HEDFT_OK) return rc;
  skip_comments(t, cchar);
  double h[9];
  rc = read_scalars(t, 9, h, nullptr);
  if (rc != HEDFT_OK) return rc;
  std::memset(hdr, 0, sizeof *hdr);
  hdr->xmax = h[0]; hdr->ymax = h[1]; hdr->zmax = h[2];
  hdr->hx = h[3]; hdr->hy = h[4]; hdr->hz = h[5];
  hdr->nx = (int32_t)h[6]; hdr->ny = (int32_t)h[7]; hdr->nz = (int32_t)h[8];
  double ni;
  rc = read_scalars(t, 1, &ni, nullptr);
  if (rc != HEDFT_OK) return rc;
  *nimp = (int32_t)ni;
  if (*nimp < 0 || *nimp > max_imp) return HEDFT_EINVAL;
  if (*nimp > 0) {
    std::vector<double> tmp(3 * (size_t)*nimp);
    rc = read_scalars(t, 3 * *nimp, tmp.data(), nullptr);
    if (rc != HEDFT_OK) return rc;
    if (rimp) std::memcpy(rimp, tmp.data(), tmp.size() * sizeof(double));
    rc = read_scalars(t, 3 * *nimp, tmp.data(), nullptr);
    if (rc != HEDFT_OK) return rc;
    if (vimp) std::memcpy(vimp, tmp.data(), tmp.size() * sizeof(double));
  }
  if (!psi) return HEDFT_OK;
  if (n != (size_t)hdr->nx * hdr->ny * hdr->nz) return HEDFT_EGRID;
  return read_array(t, n, psi, true);
}

int hedft_write_wavefunction(const char* path, const hedft_denfile* hdr, int32_t nimp, const double* rimp,
                             const double* vimp, const double* psi, size_t n, const char* comment) {
  if (!path || !hdr || !psi || nimp < 0 || n != (size_t)hdr->nx * hdr->ny * hdr->nz) return HEDFT_EINVAL;
  if (nimp > 0 && (!rimp || !vimp)) return HEDFT_EINVAL;
  FILE* f = std::fopen(path, "wb");
  if (!f) return HEDFT_EINVAL;
  write_comment(f, comment);
  std::fprintf(f, " %.17E %.17E %.17E %.17E %.17E %.17E %d %d %d\n", hdr->xmax, hdr->ymax, hdr->zmax, hdr->hx,
               hdr->hy, hdr->hz, hdr->nx, hdr->ny, hdr->nz);
  std::fprintf(f, " %d\n", nimp);
  int rc = HEDFT_OK;
  if (nimp > 0) {
    rc = write_values(f, 3 * (size_t)nimp, rimp, false);
    if (rc == HEDFT_OK) rc = write_values(f, 3 * (size_t)nimp, vimp, false);
  }
  if (rc == HEDFT_OK) rc = write_values(f, n, psi, true);
  return (std::fclose(f) == 0) ? rc : HEDFT_EINVAL;
}

}  // extern "C"

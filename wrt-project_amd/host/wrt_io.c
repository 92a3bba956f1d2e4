/*
 * wrt_io.c -- the reference's text file formats, read and written in parallel, and a raw
 * binary format for chaining the programs without text.
 *
 * Text.  The reference reads values with fscanf "%lf\n" (one column) or
 * "%lf, %lf, %lf, %lf\n" keeping the last (radon_grid/src/init.c:107-139,
 * radon_reduction/init.c:116-164, radon_inversion/backprojection/C/src/init.c:117-153) and
 * writes them with fprintf "%lf" (radon_grid/src/radon_grid.c:248-258 and the other drivers).
 *   - Reading: the file is loaded whole; when it holds one record per line, the lines are
 *     parsed in parallel with strtod (the conversion fscanf's %lf performs).  Otherwise (or on
 *     any malformed record) a sequential scanner with fscanf's whitespace rules takes over and
 *     reports the error as the programs did.
 *   - Writing: %lf is %.6f, i.e. the exact binary value rounded to 6 decimals, ties to even.
 *     For |v| < 2^43 that rounding is done here exactly in 128-bit integers from the double's
 *     significand and exponent, so the bytes equal the C library's; other values (huge,
 *     infinite, NaN) go through snprintf.  Rows are formatted in parallel into blocks that are
 *     written in order.
 * Binary: an 8-byte magic "WRTF64\0\1", the value count as uint64, then the values as
 * little-endian binary64 in the text file's element order.
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/wrt.h"

static const char kMagic[8] = {'W', 'R', 'T', 'F', '6', '4', 0, 1};

#define LF_MAX 330 /* "%lf" of -DBL_MAX: 309 integer digits, sign, point, 6 decimals */

/* "%lf" of v into buf (at least LF_MAX bytes); returns the length */
static int fmt_lf(double v, char* buf) {
  if (!isfinite(v) || fabs(v) >= 8796093022208.0 /* 2^43 */) return snprintf(buf, LF_MAX, "%lf", v);
  int e;
  const double fr = frexp(fabs(v), &e); /* |v| = fr * 2^e, fr in [0.5, 1) or 0 */
  const uint64_t m = (uint64_t)ldexp(fr, 53); /* exact: |v| = m * 2^(e - 53) */
  const int sh = 53 - e;                      /* |v| = m / 2^sh */
  unsigned __int128 num = (unsigned __int128)m * 1000000u;
  uint64_t q; /* round-half-even(|v| * 10^6) */
  if (sh <= 0) {
    q = (uint64_t)(num << (-sh));
  } else if (sh >= 120) {
    q = 0; /* |v| < 2^-66: far below half a unit of the 6th decimal */
  } else {
    const unsigned __int128 qq = num >> sh;
    const unsigned __int128 r = num - (qq << sh);
    const unsigned __int128 half = (unsigned __int128)1 << (sh - 1);
    q = (uint64_t)qq + ((r > half || (r == half && (qq & 1))) ? 1 : 0);
  }
  char* p = buf;
  if (signbit(v)) *p++ = '-';
  const uint64_t ip = q / 1000000u;
  uint32_t fp = (uint32_t)(q % 1000000u);
  char tmp[24];
  int k = 0;
  uint64_t t = ip;
  do {
    tmp[k++] = (char)('0' + t % 10);
    t /= 10;
  } while (t);
  while (k) *p++ = tmp[--k];
  *p++ = '.';
  for (int d = 5; d >= 0; --d) {
    p[d] = (char)('0' + fp % 10);
    fp /= 10;
  }
  p += 6;
  return (int)(p - buf);
}

/* rows [r0, r1) of a write into one buffer: 4 columns "a, b, c, v" or 1 column "v" */
typedef struct {
  const double *a, *b, *c, *v;
  size_t nb, nc;
  int cols;
} rows_src;

/* growable output of one block */
typedef struct {
  char* p;
  size_t len, cap;
} text_buf;

static int fmt_rows(const rows_src* s, size_t r0, size_t r1, text_buf* tb) {
  char row[4 * (LF_MAX + 2) + 2];
  tb->len = 0;
  for (size_t r = r0; r < r1; ++r) {
    char* p = row;
    if (s->cols == 4) {
      const size_t k = r % s->nc, j = (r / s->nc) % s->nb, i = r / (s->nc * s->nb);
      p += fmt_lf(s->a[i], p);
      *p++ = ',';
      *p++ = ' ';
      p += fmt_lf(s->b[j], p);
      *p++ = ',';
      *p++ = ' ';
      p += fmt_lf(s->c[k], p);
      *p++ = ',';
      *p++ = ' ';
    }
    p += fmt_lf(s->v[r], p);
    *p++ = '\n';
    const size_t n = (size_t)(p - row);
    if (tb->len + n > tb->cap) {
      const size_t cap = 2 * (tb->len + n);
      char* q = (char*)realloc(tb->p, cap);
      if (!q) return WRT_E_NOMEM;
      tb->p = q;
      tb->cap = cap;
    }
    memcpy(tb->p + tb->len, row, n);
    tb->len += n;
  }
  return WRT_OK;
}

#define IO_BLOCK_ROWS 16384
#define IO_GROUP 64 /* blocks formatted in parallel, then written in order */

static int write_rows(const char* path, const rows_src* s, size_t count) {
  FILE* f = fopen(path, "w");
  if (!f) return WRT_E_IO;
  const size_t nblk = (count + IO_BLOCK_ROWS - 1) / IO_BLOCK_ROWS;
  /* typical rows are ~ 10 characters per value; a block grows if its values are longer */
  const size_t cap0 = (size_t)IO_BLOCK_ROWS * (s->cols == 4 ? 48 : 12);
  text_buf* tb = (text_buf*)calloc(IO_GROUP, sizeof(text_buf));
  int rc = tb ? WRT_OK : WRT_E_NOMEM;
  for (size_t g = 0; g < IO_GROUP && rc == WRT_OK; ++g) {
    tb[g].cap = cap0;
    if (!(tb[g].p = (char*)malloc(cap0))) rc = WRT_E_NOMEM;
  }
  for (size_t b0 = 0; b0 < nblk && rc == WRT_OK; b0 += IO_GROUP) {
    const size_t nb = nblk - b0 < IO_GROUP ? nblk - b0 : IO_GROUP;
    int err = WRT_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : err)
    for (long long g = 0; g < (long long)nb; ++g) {
      const size_t r0 = (b0 + (size_t)g) * IO_BLOCK_ROWS;
      const size_t r1 = r0 + IO_BLOCK_ROWS < count ? r0 + IO_BLOCK_ROWS : count;
      const int e = fmt_rows(s, r0, r1, &tb[g]);
      if (e < err) err = e;
    }
    rc = err;
    for (size_t g = 0; g < nb && rc == WRT_OK; ++g)
      if (fwrite(tb[g].p, 1, tb[g].len, f) != tb[g].len) rc = WRT_E_IO;
  }
  if (tb)
    for (size_t g = 0; g < IO_GROUP; ++g) free(tb[g].p);
  free(tb);
  if (fclose(f) != 0 && rc == WRT_OK) rc = WRT_E_IO;
  return rc;
}

int wrt_write_text_grid(const char* path, const double* a, int na, const double* b, int nb,
                        const double* c, int nc, const double* values) {
  if (!path || !a || !b || !c || !values || na < 0 || nb < 1 || nc < 1) return WRT_E_ARG;
  rows_src s = {a, b, c, values, (size_t)nb, (size_t)nc, 4};
  return write_rows(path, &s, (size_t)na * nb * nc);
}

int wrt_write_text_values(const char* path, const double* values, size_t count) {
  if (!path || (!values && count)) return WRT_E_ARG;
  rows_src s = {NULL, NULL, NULL, values, 1, 1, 1};
  return write_rows(path, &s, count);
}

/* ------------------------------------------------------------------------ reading --- */
static int is_space(char ch) {
  return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\v' || ch == '\f' || ch == '\r';
}

/* One record at *pp with fscanf's rules: "%lf" skips white space and converts; in the
 * multi-column formats each "," must follow its number directly and ", " skips any white space;
 * the trailing "\n" skips any white space.  Returns the number of values matched (cols on
 * success), or -1 at the end of the data before the first value. */
static int scan_record(const char** pp, const char* end, int cols, double* last) {
  const char* p = *pp;
  int got = 0;
  for (int k = 0; k < cols; ++k) {
    while (p < end && is_space(*p)) ++p;
    if (p >= end) {
      *pp = p;
      return got == 0 ? -1 : got;
    }
    char* q;
    const double v = strtod(p, &q);
    if (q == p) {
      *pp = p;
      return got;
    }
    p = q;
    ++got;
    *last = v;
    if (k + 1 < cols) {
      if (p >= end || *p != ',') {
        *pp = p;
        return got;
      }
      ++p;
    }
  }
  while (p < end && is_space(*p)) ++p;
  *pp = p;
  return got;
}

static int read_sequential(const char* data, size_t len, size_t count, int cols, double* out) {
  const char* p = data;
  const char* end = data + len;
  for (size_t i = 0; i < count; ++i) {
    double v = 0;
    const int got = scan_record(&p, end, cols, &v);
    if (got < 0) {
      fprintf(stderr, "Error : fscanf : reached the end of file before it was expected.\n");
      return WRT_E_IO;
    }
    if (got != cols) {
      fprintf(stderr, "Error: fscanf : matching failure, expected %d floating-point value(s).\n", cols);
      return WRT_E_IO;
    }
    out[i] = v;
  }
  return WRT_OK;
}

int wrt_read_text(const char* path, size_t count, int cols, double* out) {
  if (!path || cols < 1 || cols > 4 || (!out && count)) return WRT_E_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), path);
    return WRT_E_IO;
  }
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* data = (char*)malloc((size_t)(sz > 0 ? sz : 0) + 1);
  if (!data) {
    fclose(f);
    return WRT_E_NOMEM;
  }
  const size_t len = sz > 0 ? fread(data, 1, (size_t)sz, f) : 0;
  fclose(f);
  data[len] = 0; /* strtod stops at the terminator */
  /* line starts of the first count non-blank-free lines */
  size_t* start = (size_t*)malloc(sizeof(size_t) * (count + 1));
  int fast = start != NULL;
  size_t nl = 0, pos = 0;
  while (fast && nl < count && pos < len) {
    start[nl++] = pos;
    const char* e = (const char*)memchr(data + pos, '\n', len - pos);
    pos = e ? (size_t)(e - data) + 1 : len;
  }
  if (fast && nl < count) fast = 0;
  int bad = 0;
  if (fast) {
    start[count] = pos;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (long long i = 0; i < (long long)count; ++i) {
      const char* p = data + start[i];
      const char* e = data + start[i + 1];
      double v = 0;
      /* the record must fill its line exactly: a line with no value, or with a second
       * record, is left to the sequential scanner */
      if (scan_record(&p, e, cols, &v) != cols || p != e) {
        bad = 1;
      } else {
        out[i] = v;
      }
    }
  }
  int rc = WRT_OK;
  if (!fast || bad) rc = read_sequential(data, len, count, cols, out);
  free(start);
  free(data);
  return rc;
}

/* ------------------------------------------------------------------------- binary --- */
int wrt_write_f64(const char* path, const double* values, size_t count) {
  if (!path || (!values && count)) return WRT_E_ARG;
  FILE* f = fopen(path, "wb");
  if (!f) return WRT_E_IO;
  const uint64_t n = count;
  int rc = WRT_OK;
  if (fwrite(kMagic, 1, 8, f) != 8 || fwrite(&n, 8, 1, f) != 1 ||
      (count && fwrite(values, sizeof(double), count, f) != count))
    rc = WRT_E_IO;
  if (fclose(f) != 0) rc = WRT_E_IO;
  return rc;
}

int wrt_read_f64(const char* path, size_t count, double* out) {
  if (!path || (!out && count)) return WRT_E_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), path);
    return WRT_E_IO;
  }
  char magic[8];
  uint64_t n = 0;
  int rc = WRT_OK;
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kMagic, 8) != 0 || fread(&n, 8, 1, f) != 1) {
    fprintf(stderr, "Error : %s is not a binary data file of this library.\n", path);
    rc = WRT_E_IO;
  } else if (n != count) {
    fprintf(stderr, "Error : %s holds %llu values, %zu expected.\n", path, (unsigned long long)n, count);
    rc = WRT_E_IO;
  } else if (count && fread(out, sizeof(double), count, f) != count) {
    fprintf(stderr, "Error : %s ends before its %zu values.\n", path, count);
    rc = WRT_E_IO;
  }
  fclose(f);
  return rc;
}

/*
 * Goldilocks field and NTT oracle.  TEST INFRASTRUCTURE ONLY: the shipped
 * library never links this file; tests, the smoke check and the CPU baseline
 * of the benchmark use it as the checker.
 *
 * Pins (see tests/test_oracle_ntt.py):
 *   - field ops vs Python big ints on the edge vectors of
 *     libs/hardcaml_ntt/test/test_gf.ml:9-25, and vs the reference's own
 *     zprize/ntt/host/gf.h compiled into oracle/_ref where that exists;
 *   - roots vs the OMEGA table of zprize/ntt/host/gf.h:65-99 (fixture);
 *   - transforms vs the reference KATs (ntt_test.cpp:22-101 ==
 *     test_ntt_against_reference.ml:33-85, test_ntt_4step.ml:19-40) and
 *     vs an O(n^2) brute-force DFT and closed forms (delta, ones, 1+2w^k).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define GL_P 0xFFFFFFFF00000001ull
#define GL_EPS 0xFFFFFFFFull

/* zprize/ntt/host/gf.h:19-25: add with overflow or >= p folds back by -p. */
uint64_t orc_gl_add(uint64_t a, uint64_t b) {
  uint64_t s;
  int ovf = __builtin_add_overflow(a, b, &s);
  if (ovf || s >= GL_P) s -= GL_P;
  return s;
}

/* zprize/ntt/host/gf.h:27-33: borrow adds p back. */
uint64_t orc_gl_sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  if (a < b) d += GL_P;
  return d;
}

/* zprize/ntt/host/gf.h:35-58: 128-bit product reduced with 2^64 = eps,
 * 2^96 = -1 (mod p). */
uint64_t orc_gl_mul(uint64_t a, uint64_t b) {
  unsigned __int128 prod = (unsigned __int128)a * b;
  uint64_t lo = (uint64_t)prod;
  uint64_t hi = (uint64_t)(prod >> 64);
  uint64_t hi_hi = hi >> 32;
  uint64_t hi_lo = hi & GL_EPS;
  uint64_t t0;
  if (__builtin_sub_overflow(lo, hi_hi, &t0)) t0 -= GL_EPS;
  uint64_t t1 = hi_lo * GL_EPS;
  uint64_t r;
  if (__builtin_add_overflow(t0, t1, &r)) r += GL_EPS;
  if (r >= GL_P) r -= GL_P;
  return r;
}

uint64_t orc_gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = orc_gl_mul(r, a);
    a = orc_gl_mul(a, a);
    e >>= 1;
  }
  return r;
}

uint64_t orc_gl_inv(uint64_t a) { return orc_gl_pow(a, GL_P - 2); }

/* The 2^i-th root used by zprize/ntt/host/gf.h:65-99 (OMEGA[i]); computed
 * from its definition 7^((p-1)/2^i) rather than copied. */
uint64_t orc_gl_root(unsigned log_n) {
  if (log_n > 32) return 0;
  return orc_gl_pow(7, (GL_P - 1) >> log_n);
}

static uint64_t bitrev(uint64_t x, unsigned bits) {
  uint64_t r = 0;
  for (unsigned i = 0; i < bits; i++) {
    r = (r << 1) | ((x >> i) & 1);
  }
  return r;
}

/* zprize/ntt/host/reference.cpp:87-137: bit-reverse permutation, then log n
 * radix-2 DIT stages, w <- w * w_m within each block. */
static void dit_inplace(uint64_t *a, unsigned log_n, int inverse_roots) {
  uint64_t n = 1ull << log_n;
  for (uint64_t k = 0; k < n; k++) {
    uint64_t rk = bitrev(k, log_n);
    if (k < rk) {
      uint64_t t = a[k];
      a[k] = a[rk];
      a[rk] = t;
    }
  }
  uint64_t m = 1;
  for (unsigned s = 1; s <= log_n; s++) {
    uint64_t wm = orc_gl_root(s);
    if (inverse_roots) wm = orc_gl_inv(wm);
    for (uint64_t k = 0; k < n; k += 2 * m) {
      uint64_t w = 1;
      for (uint64_t j = 0; j < m; j++) {
        uint64_t t = orc_gl_mul(a[k + j + m], w);
        uint64_t u = a[k + j];
        a[k + j + m] = orc_gl_sub(u, t);
        a[k + j] = orc_gl_add(u, t);
        w = orc_gl_mul(w, wm);
      }
    }
    m <<= 1;
  }
}

void orc_ntt_inplace_unscaled(uint64_t *a, unsigned log_n, int dir) {
  dit_inplace(a, log_n, dir != 0);
}

void orc_ntt_inplace(uint64_t *a, unsigned log_n, int dir) {
  dit_inplace(a, log_n, dir != 0);
  if (dir != 0 && log_n > 0) {
    uint64_t n_inv = orc_gl_inv(1ull << log_n);
    uint64_t n = 1ull << log_n;
    for (uint64_t i = 0; i < n; i++) a[i] = orc_gl_mul(a[i], n_inv);
  }
}

/* libs/hardcaml_ntt/src/reference_model.ml:92-105 four_step:
 *   matrix (log_cols rows x log_rows cols) |> transpose; inverse_dit on each
 *   row; apply_twiddles with w_n (row r, col c scaled by w_n^(r*c));
 *   transpose; inverse_dit on each row; transpose; concat.
 * Here "inverse_dit" is the reference model's name for the OMEGA-rooted
 * transform (forward in zprize/ntt/host terms). */
void orc_ntt_four_step(uint64_t *a, unsigned log_n, unsigned log_rows) {
  unsigned log_cols = log_n - log_rows;
  uint64_t rows = 1ull << log_rows, cols = 1ull << log_cols, n = 1ull << log_n;
  /* matrix a log_cols log_rows: 2^log_cols rows of 2^log_rows; transposed it
   * is 2^log_rows rows (index i) of 2^log_cols entries m[i][j] = a[j*rows+i]. */
  uint64_t *m = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t i = 0; i < rows; i++)
    for (uint64_t j = 0; j < cols; j++) m[i * cols + j] = a[j * rows + i];
  for (uint64_t i = 0; i < rows; i++) dit_inplace(m + i * cols, log_cols, 0);
  uint64_t wn = orc_gl_root(log_n);
  uint64_t wr = 1; /* wn^i */
  for (uint64_t i = 0; i < rows; i++) {
    uint64_t w = 1; /* wn^(i*j) */
    for (uint64_t j = 0; j < cols; j++) {
      m[i * cols + j] = orc_gl_mul(w, m[i * cols + j]);
      w = orc_gl_mul(wr, w);
    }
    wr = orc_gl_mul(wr, wn);
  }
  /* transpose: 2^log_cols rows (index j) of 2^log_rows; row NTTs. */
  uint64_t *t = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t j = 0; j < cols; j++)
    for (uint64_t i = 0; i < rows; i++) t[j * rows + i] = m[i * cols + j];
  for (uint64_t j = 0; j < cols; j++) dit_inplace(t + j * rows, log_rows, 0);
  /* final transpose and concat: out[i*cols + j] = t[j][i]. */
  for (uint64_t i = 0; i < rows; i++)
    for (uint64_t j = 0; j < cols; j++) a[i * cols + j] = t[j * rows + i];
  free(m);
  free(t);
}

typedef struct {
  uint64_t *a;
  unsigned log_n;
  size_t begin, end;
  int dir;
} many_arg;

static void *many_worker(void *p) {
  many_arg *arg = (many_arg *)p;
  for (size_t i = arg->begin; i < arg->end; i++)
    orc_ntt_inplace(arg->a + (i << arg->log_n), arg->log_n, arg->dir);
  return NULL;
}

void orc_ntt_many(uint64_t *a, unsigned log_n, size_t count, int dir, int threads) {
  if (threads < 1) threads = 1;
  if ((size_t)threads > count) threads = (int)(count ? count : 1);
  pthread_t tid[64];
  many_arg args[64];
  if (threads > 64) threads = 64;
  size_t per = (count + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    args[t].a = a;
    args[t].log_n = log_n;
    args[t].begin = t * per < count ? t * per : count;
    args[t].end = (t + 1) * per < count ? (t + 1) * per : count;
    args[t].dir = dir;
    pthread_create(&tid[t], NULL, many_worker, &args[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* ---- REVERSE core model (zprize/ntt/host/reference.cpp:7-84): the FPGA
 * build's memory-path debug core.  Same two phases as the 4-step transform
 * with each row DFT replaced by reversing the row (do_stuff, :18-21):
 * transpose, reverse rows, transpose (phase 1, :69-74); multiply element
 * (i, j) by w^(i j), w = OMEGA[2 log_row_size] (slow_apply_twiddle, :46-57);
 * reverse rows, transpose (phase 2, :79-83). */
static void rev_transpose(uint64_t *p, uint64_t r) { /* reference.cpp:7-15 */
  for (uint64_t i = 0; i < r; i++)
    for (uint64_t j = i + 1; j < r; j++) {
      uint64_t t = p[i * r + j];
      p[i * r + j] = p[j * r + i];
      p[j * r + i] = t;
    }
}

static void rev_rows(uint64_t *p, uint64_t r) {
  for (uint64_t i = 0; i < r; i++)
    for (uint64_t a = 0, b = r - 1; a < b; a++, b--) {
      uint64_t t = p[i * r + a];
      p[i * r + a] = p[i * r + b];
      p[i * r + b] = t;
    }
}

void orc_ntt_reverse(uint64_t *dst, const uint64_t *src, unsigned log_row_size) {
  const uint64_t r = 1ull << log_row_size;
  memcpy(dst, src, sizeof(uint64_t) * r * r);
  rev_transpose(dst, r);
  rev_rows(dst, r);
  rev_transpose(dst, r);
  const uint64_t w = orc_gl_root(2 * log_row_size);
  for (uint64_t i = 0; i < r; i++)
    for (uint64_t j = 0; j < r; j++) dst[i * r + j] = orc_gl_mul(dst[i * r + j], orc_gl_pow(w, i * j));
  rev_rows(dst, r);
  rev_transpose(dst, r);
}

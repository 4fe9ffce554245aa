/*
 * BLS12-377 G1 MSM oracle.  TEST INFRASTRUCTURE ONLY: the shipped library
 * never links this file; tests, the smoke check and the CPU baseline of the
 * benchmark use it as the checker.
 *
 * Restates the reference host algorithm:
 *   field      zprize/msm_pippenger/host/msm_compute/include/bls12_377_g1/bls12_377_g1.h:90-219
 *              (GMP there; here 6x64-bit Montgomery words, R = 2^384, the
 *              arkworks wire form named by test_fpga_harness/README.md:13-17)
 *   import     bls12_377_g1.h:646-659, 406-415, 361-388, 343-346
 *   precompute bls12_377_g1.h:606-618, libs/twisted_edwards/model/twisted_edwards_curve.ml:109-123
 *   digits     zprize/msm_pippenger/hardcaml/test/top/utils.ml:243-288, config_utils.ml:7-41
 *   buckets    twisted_edwards_curve.ml:126-170 (add_unified_precomputed,
 *              buckets start at (2,2,4,0), zprize/msm_pippenger/hardcaml/src/top.ml:115)
 *   reduction  zprize/msm_pippenger/host/driver/driver.cpp:306-347, pippenger.h:25-45,
 *              bls12_377_g1.h:449-491 (hwcd-3 unified add), 582-599 (postCompute)
 *   output     bls12_377_g1.h:427-439, 347-402
 *   fallback   driver.cpp:349-372, bls12_377_g1.h:493-566, 701-742
 *
 * Pins (tests/test_oracle_msm.py): Python big-int textbook affine arithmetic
 * (tests/golden/), closed form MSM = (sum s_i k_i mod r) G for P_i = k_i G,
 * the special points of test_fpga_harness/tests/msm_unit_tests.rs:30-48,
 * brute force on tiny inputs.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef struct { uint64_t v[6]; } fq;
typedef struct { fq x, y, z, t; } xyzt;

/* Constants in plain (non-Montgomery) form, bls12_377_g1.h:24-54. */
static const uint64_t Q_W[6] = {0x8508c00000000001ull, 0x170b5d4430000000ull, 0x1ef3622fba094800ull,
                                0x1a22d9f300f5138full, 0xc63b05c06ca1493bull, 0x01ae3a4617c510eaull};
static const char *TS_HEX = "363b01df81a0405e86206f2cc504d147752b91b664d91d4c82fed217902ee874354485836fa3ccba3af539c173857b";
static const char *K_HEX = "177f95e65b6bf7fc2f8f4fbe3b4cc6e2aa705a3f51e66d9fdd06ffd95626c6a3adc8cbbf1c2c1ba4dcada97b292007";
static const char *S_HEX = "19d47d415b5ff60a87a8b7bbab25eb6427dd58ca38e47030efd1e6310ac7bf3079221bf2b4bd72c5106e9e70fcc6156";
static const char *D_HEX = "e2dcedff103e7161354a88156e4b00fe66a526a0237cfe5f683497c9afb7635d5c9307f78e160e14f2b6d4bd949004";
static const char *CA_HEX = "32d756062d349e59416ece15ccbf8e86ef0d33183465a42fe2cb65fc1664272e6bb28f0e1c7a7c9c05824ad09adc01";
static const char *CB_HEX = "19d47d415b5ff60a87a8b7bbab25eb6427dd58ca38e47030efd1e6310ac7bf3079221bf2b4bd72c5106e9e70fcc6156";
/* G1 generator, libs/ark_bls12_377_g1/test/test_ark_bls12_377_g1.ml:45-57. */
static const char *GX_HEX = "8848defe740a67c8fc6225bf87ff5485951e2caa9d41bb188282c8bd37cb5cd5481512ffcd394eeab9b16eb21be9ef";
static const char *GY_HEX = "1914a69c5102eff1f674f5d30afeec4bd7fb348ca3e52d96d182ad44fb82305c2fe3d3634a9591afd82de55559c8ea6";

static uint64_t QINV; /* -q^-1 mod 2^64 */
static fq F_ZERO, F_ONE, F_R2, F_TWO, F_THREE, F_FOUR, F_TS, F_K, F_S, F_D, F_CA, F_CB, F_ALPHA, F_INV2;
static fq F_GX, F_GY;
static pthread_once_t init_once = PTHREAD_ONCE_INIT;

/* ---- raw multiword helpers ---- */
static int geq_q(const uint64_t a[6]) {
  for (int i = 5; i >= 0; i--) {
    if (a[i] > Q_W[i]) return 1;
    if (a[i] < Q_W[i]) return 0;
  }
  return 1;
}
static void sub_q(uint64_t a[6]) {
  uint64_t br = 0;
  for (int i = 0; i < 6; i++) {
    u128 d = (u128)a[i] - Q_W[i] - br;
    a[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}

static void mont_mul(fq *r, const fq *a, const fq *b) {
  uint64_t t[8] = {0};
  for (int i = 0; i < 6; i++) {
    u128 c = 0;
    for (int j = 0; j < 6; j++) {
      c = (u128)a->v[j] * b->v[i] + t[j] + (uint64_t)(c >> 64);
      t[j] = (uint64_t)c;
    }
    c = (u128)t[6] + (uint64_t)(c >> 64);
    t[6] = (uint64_t)c;
    t[7] = (uint64_t)(c >> 64);
    uint64_t m = t[0] * QINV;
    c = (u128)m * Q_W[0] + t[0];
    for (int j = 1; j < 6; j++) {
      c = (u128)m * Q_W[j] + t[j] + (uint64_t)(c >> 64);
      t[j - 1] = (uint64_t)c;
    }
    c = (u128)t[6] + (uint64_t)(c >> 64);
    t[5] = (uint64_t)c;
    t[6] = t[7] + (uint64_t)(c >> 64);
  }
  uint64_t out[6];
  memcpy(out, t, sizeof(out));
  if (t[6] || geq_q(out)) sub_q(out);
  memcpy(r->v, out, sizeof(out));
}

static void f_add(fq *r, const fq *a, const fq *b) {
  uint64_t o[6], c = 0;
  for (int i = 0; i < 6; i++) {
    u128 s = (u128)a->v[i] + b->v[i] + c;
    o[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (c || geq_q(o)) sub_q(o);
  memcpy(r->v, o, sizeof(o));
}

static void f_sub(fq *r, const fq *a, const fq *b) {
  uint64_t o[6], br = 0;
  for (int i = 0; i < 6; i++) {
    u128 d = (u128)a->v[i] - b->v[i] - br;
    o[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    uint64_t c = 0;
    for (int i = 0; i < 6; i++) {
      u128 s = (u128)o[i] + Q_W[i] + c;
      o[i] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  memcpy(r->v, o, sizeof(o));
}

static void f_mul(fq *r, const fq *a, const fq *b) { mont_mul(r, a, b); }
static int f_eq(const fq *a, const fq *b) { return memcmp(a->v, b->v, 48) == 0; }
static int f_is_zero(const fq *a) { return f_eq(a, &F_ZERO); }

static void f_pow(fq *r, const fq *a, const uint64_t e[6]) {
  fq acc = F_ONE, base = *a;
  for (int i = 0; i < 6; i++)
    for (int b = 0; b < 64; b++) {
      if ((e[i] >> b) & 1) f_mul(&acc, &acc, &base);
      f_mul(&base, &base, &base);
    }
  *r = acc;
}

/* mpz_invert semantics (bls12_377_g1.h:152-158) via Fermat; 0 maps to 0. */
static void f_inv(fq *r, const fq *a) {
  uint64_t e[6];
  memcpy(e, Q_W, sizeof(e));
  e[0] -= 2;
  f_pow(r, a, e);
}
static void f_div(fq *r, const fq *a, const fq *b) {
  fq bi;
  f_inv(&bi, b);
  f_mul(r, a, &bi);
}

static void plain_to_fq(fq *r, const uint64_t plain[6]) {
  fq p;
  memcpy(p.v, plain, 48);
  mont_mul(r, &p, &F_R2);
}
static void fq_to_plain(uint64_t out[6], const fq *a) {
  fq one_plain = {{1, 0, 0, 0, 0, 0}};
  fq r;
  mont_mul(&r, a, &one_plain);
  memcpy(out, r.v, 48);
}

static void hex_to_words(uint64_t w[6], const char *hex) {
  memset(w, 0, 48);
  size_t len = strlen(hex);
  for (size_t i = 0; i < len; i++) {
    char ch = hex[len - 1 - i];
    uint64_t v = (ch >= '0' && ch <= '9') ? (uint64_t)(ch - '0')
                 : (ch >= 'a' && ch <= 'f') ? (uint64_t)(ch - 'a' + 10)
                                            : (uint64_t)(ch - 'A' + 10);
    w[i / 16] |= v << (4 * (i % 16));
  }
}
static void hex_to_fq(fq *r, const char *hex) {
  uint64_t w[6];
  hex_to_words(w, hex);
  plain_to_fq(r, w);
}
static void small_fq(fq *r, uint64_t v) {
  uint64_t w[6] = {v, 0, 0, 0, 0, 0};
  plain_to_fq(r, w);
}

static void do_init(void) {
  /* -q^-1 mod 2^64 by Newton iteration. */
  uint64_t inv = 1;
  for (int i = 0; i < 7; i++) inv *= 2 - Q_W[0] * inv;
  QINV = (uint64_t)0 - inv;
  /* R2 = 2^768 mod q: start at 1 and double 768 times. */
  uint64_t x[6] = {1, 0, 0, 0, 0, 0};
  for (int k = 0; k < 768; k++) {
    uint64_t c = 0;
    for (int i = 0; i < 6; i++) {
      uint64_t nc = x[i] >> 63;
      x[i] = (x[i] << 1) | c;
      c = nc;
    }
    if (c || geq_q(x)) sub_q(x);
  }
  memcpy(F_R2.v, x, 48);
  memset(F_ZERO.v, 0, 48);
  small_fq(&F_ONE, 1);
  small_fq(&F_TWO, 2);
  small_fq(&F_THREE, 3);
  small_fq(&F_FOUR, 4);
  hex_to_fq(&F_TS, TS_HEX);
  hex_to_fq(&F_K, K_HEX);
  hex_to_fq(&F_S, S_HEX);
  hex_to_fq(&F_D, D_HEX);
  hex_to_fq(&F_CA, CA_HEX);
  hex_to_fq(&F_CB, CB_HEX);
  f_sub(&F_ALPHA, &F_ZERO, &F_ONE); /* A_WORDS = q - 1 (bls12_377_g1.h:244) */
  f_inv(&F_INV2, &F_TWO);
  hex_to_fq(&F_GX, GX_HEX);
  hex_to_fq(&F_GY, GY_HEX);
}
static void ensure_init(void) { pthread_once(&init_once, do_init); }

/* ---- exported field ops (Montgomery words) ---- */
void orc_fq_mul(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]) {
  ensure_init();
  fq x, y, r;
  memcpy(x.v, a, 48);
  memcpy(y.v, b, 48);
  f_mul(&r, &x, &y);
  memcpy(out, r.v, 48);
}
void orc_fq_add(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]) {
  ensure_init();
  fq x, y, r;
  memcpy(x.v, a, 48);
  memcpy(y.v, b, 48);
  f_add(&r, &x, &y);
  memcpy(out, r.v, 48);
}
void orc_fq_sub(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]) {
  ensure_init();
  fq x, y, r;
  memcpy(x.v, a, 48);
  memcpy(y.v, b, 48);
  f_sub(&r, &x, &y);
  memcpy(out, r.v, 48);
}
void orc_fq_inv(uint64_t out[6], const uint64_t a[6]) {
  ensure_init();
  fq x, r;
  memcpy(x.v, a, 48);
  f_inv(&r, &x);
  memcpy(out, r.v, 48);
}
void orc_fq_to_mont(uint64_t out[6], const uint64_t plain[6]) {
  ensure_init();
  fq r;
  plain_to_fq(&r, plain);
  memcpy(out, r.v, 48);
}
void orc_fq_from_mont(uint64_t out[6], const uint64_t mont[6]) {
  ensure_init();
  fq a;
  memcpy(a.v, mont, 48);
  fq_to_plain(out, &a);
}

/* ---- windows and signed digits ---- */
/* config_utils.ml:7-41 with scalar_bits = 253, num_windows = 20. */
int orc_msm_window_bits(int w) { return w < 13 ? 13 : 12; }
int orc_msm_window_offset(int w) { return w <= 13 ? 13 * w : 169 + 12 * (w - 13); }
int orc_msm_num_buckets(int w) {
  if (w == 19) return (1 << orc_msm_window_bits(w)) - 1;
  return 1 << (orc_msm_window_bits(w) - 1);
}

static uint32_t slice_bits(const orc_bigint256 *s, int off, int len) {
  uint32_t v = 0;
  for (int b = 0; b < len; b++) {
    int bit = off + b;
    if (bit >= 256) break;
    v |= (uint32_t)((s->data[bit / 64] >> (bit % 64)) & 1) << b;
  }
  return v;
}

/* utils.ml:243-288 perform_scalar_reduction. */
void orc_msm_digits(const orc_bigint256 *scalar, uint32_t mag[20], uint8_t neg[20]) {
  int carry = 0;
  for (int i = 0; i < 20; i++) {
    int size = orc_msm_window_bits(i);
    int32_t slice = (int32_t)slice_bits(scalar, orc_msm_window_offset(i), size) + carry;
    if (i == 19) {
      mag[i] = (uint32_t)slice;
      neg[i] = 0;
      break;
    }
    int32_t sv;
    if (slice >= (1 << size)) {
      sv = 0;
      carry = 1;
    } else if (slice >= (1 << (size - 1))) {
      sv = slice - (1 << size);
      carry = 1;
    } else {
      sv = slice;
      carry = 0;
    }
    mag[i] = (uint32_t)(sv < 0 ? -sv : sv);
    neg[i] = sv < 0;
  }
}

/* ---- point conversions, bls12_377_g1.h:343-447 ---- */
/* affineWeierstrassToMontgomery :361-365 */
static void w_to_mont(fq *x, fq *y) {
  f_sub(x, x, &F_ALPHA);
  f_mul(x, x, &F_S);
  f_mul(y, &F_S, y);
}
/* affineMontgomeryToWeierstrass :366-374 */
static void mont_to_w(fq *x, fq *y) {
  fq tmp;
  f_mul(&tmp, &F_THREE, &F_CB);
  f_div(&tmp, &F_CA, &tmp);
  f_div(x, x, &F_CB);
  f_add(x, x, &tmp);
  f_div(y, y, &F_CB);
}
/* affineMontgomeryToTwistedEdwards :377-388 */
static int mont_to_te(fq *x, fq *y) {
  fq t1, t2, t3;
  f_mul(&t3, &F_TS, x);
  f_sub(&t1, x, &F_ONE);
  f_add(&t2, x, &F_ONE);
  if (f_is_zero(y) || f_is_zero(&t2)) return 0;
  f_div(x, &t3, y);
  f_div(y, &t1, &t2);
  return 1;
}
/* affineTwistedEdwardsToMontgomery :389-402 */
static int te_to_mont(fq *x, fq *y) {
  fq t1, t2;
  f_add(&t1, &F_ONE, y);
  f_sub(&t2, &F_ONE, y);
  if (f_is_zero(&t2) || f_is_zero(x)) return 0;
  f_div(&t1, &t1, &t2);
  f_div(&t2, &F_TS, x);
  *x = t1;
  f_mul(y, &t1, &t2);
  return 1;
}

/* hwcd-3 unified addition with k = 2d, bls12_377_g1.h:449-491. */
static void te_add(xyzt *p, const xyzt *o) {
  fq A, B, C, D, E, F, G, H, t1, t2;
  f_sub(&t1, &p->y, &p->x);
  f_sub(&t2, &o->y, &o->x);
  f_mul(&A, &t1, &t2);
  f_add(&t1, &p->y, &p->x);
  f_add(&t2, &o->y, &o->x);
  f_mul(&B, &t1, &t2);
  f_mul(&t1, &p->t, &F_K);
  f_mul(&C, &t1, &o->t);
  f_mul(&t1, &p->z, &F_TWO);
  f_mul(&D, &t1, &o->z);
  f_sub(&E, &B, &A);
  f_sub(&F, &D, &C);
  f_add(&G, &D, &C);
  f_add(&H, &B, &A);
  f_mul(&p->x, &E, &F);
  f_mul(&p->y, &G, &H);
  f_mul(&p->t, &E, &H);
  f_mul(&p->z, &F, &G);
}
static void te_identity(xyzt *p) {
  p->x = F_ZERO;
  p->y = F_ONE;
  p->z = F_ONE;
  p->t = F_ZERO;
}

/* twisted_edwards_curve.ml:145-170 add_unified_precomputed: bucket in the
 * internal form (2(y-x), 2(y+x), 4z, t), point in host form
 * ((y-x)/2, (y+x)/2, 4dt); subtract swaps x/y and negates t. */
static void te_add_precomputed(xyzt *b, const fq pt[3], int subtract) {
  fq xh, yh, th;
  if (subtract) {
    xh = pt[1];
    yh = pt[0];
    f_sub(&th, &F_ZERO, &pt[2]);
  } else {
    xh = pt[0];
    yh = pt[1];
    th = pt[2];
  }
  fq cA, cB, cC, cD, cE, cF, cG, cH, px, py;
  f_mul(&cA, &b->x, &xh);
  f_mul(&cB, &b->y, &yh);
  f_mul(&cC, &b->t, &th);
  cD = b->z;
  f_sub(&cE, &cB, &cA);
  f_sub(&cF, &cD, &cC);
  f_add(&cG, &cD, &cC);
  f_add(&cH, &cB, &cA);
  f_mul(&px, &cE, &cF);
  f_mul(&py, &cG, &cH);
  f_mul(&b->t, &cE, &cH);
  f_mul(&b->z, &cF, &cG);
  f_sub(&b->x, &py, &px);
  f_add(&b->y, &py, &px);
}

/* divBy4 via (q+1)/2 multiplications, bls12_377_g1.h:128-139. */
static void f_div4(fq *r, const fq *a) {
  f_mul(r, a, &F_INV2);
  f_mul(r, r, &F_INV2);
}

/* postComputeFPGA :582-599: ((y-x)/4, (y+x)/4, z/4, t). */
static void post_compute(xyzt *p) {
  fq a, b;
  f_sub(&a, &p->y, &p->x);
  f_div4(&a, &a);
  f_add(&b, &p->y, &p->x);
  f_div4(&b, &b);
  p->x = a;
  p->y = b;
  f_div4(&p->z, &p->z);
}

/* Weierstrass affine with infinity flag; textbook formulas as in
 * bls12_377_g1.h:493-566 (a = 0). */
typedef struct { fq x, y; int inf; } waff;

static void w_dbl(waff *p) {
  if (p->inf) return;
  fq t1, t2, t3, l;
  f_add(&t2, &p->y, &p->y);
  if (f_is_zero(&t2)) {
    p->inf = 1;
    return;
  }
  f_mul(&t1, &p->x, &p->x);
  f_mul(&t1, &t1, &F_THREE);
  f_div(&l, &t1, &t2);
  f_mul(&t3, &l, &l);
  f_sub(&t3, &t3, &p->x);
  f_sub(&t3, &t3, &p->x); /* x3 */
  fq y3;
  f_sub(&y3, &p->x, &t3);
  f_mul(&y3, &y3, &l);
  f_sub(&y3, &y3, &p->y);
  p->x = t3;
  p->y = y3;
}
static void w_add(waff *p, const waff *o) {
  if (o->inf) return;
  if (p->inf) {
    *p = *o;
    return;
  }
  if (f_eq(&p->x, &o->x)) {
    if (f_eq(&p->y, &o->y)) {
      w_dbl(p);
    } else {
      p->inf = 1;
    }
    return;
  }
  fq dy, dx, l, x3, y3;
  f_sub(&dy, &o->y, &p->y);
  f_sub(&dx, &o->x, &p->x);
  f_div(&l, &dy, &dx);
  f_mul(&x3, &l, &l);
  f_sub(&x3, &x3, &p->x);
  f_sub(&x3, &x3, &o->x);
  f_sub(&y3, &p->x, &x3);
  f_mul(&y3, &y3, &l);
  f_sub(&y3, &y3, &p->y);
  p->x = x3;
  p->y = y3;
}
/* weierstrassMultiplication :701-724: LSB-first double-and-add over 253 bits. */
static void w_mul(waff *r, const waff *p, const orc_bigint256 *k) {
  waff base = *p, acc;
  acc.inf = 1;
  acc.x = F_ZERO;
  acc.y = F_ZERO;
  for (int i = 0; i < 256; i++) {
    if ((k->data[i / 64] >> (i % 64)) & 1) w_add(&acc, &base);
    w_dbl(&base);
  }
  *r = acc;
}

/* Jacobian double-and-add used only for bulk test-point generation. */
typedef struct { fq X, Y, Z; } jac;
static void j_dbl(jac *p) {
  if (f_is_zero(&p->Z)) return;
  fq A, B, C, D, E, F, t;
  f_mul(&A, &p->X, &p->X);
  f_mul(&B, &p->Y, &p->Y);
  f_mul(&C, &B, &B);
  f_add(&t, &p->X, &B);
  f_mul(&t, &t, &t);
  f_sub(&t, &t, &A);
  f_sub(&t, &t, &C);
  f_add(&D, &t, &t);
  f_add(&E, &A, &A);
  f_add(&E, &E, &A);
  f_mul(&F, &E, &E);
  fq X3, Y3, Z3;
  f_sub(&X3, &F, &D);
  f_sub(&X3, &X3, &D);
  f_sub(&t, &D, &X3);
  f_mul(&Y3, &E, &t);
  fq c8;
  f_add(&c8, &C, &C);
  f_add(&c8, &c8, &c8);
  f_add(&c8, &c8, &c8);
  f_sub(&Y3, &Y3, &c8);
  f_mul(&Z3, &p->Y, &p->Z);
  f_add(&Z3, &Z3, &Z3);
  p->X = X3;
  p->Y = Y3;
  p->Z = Z3;
}
/* p += (x2, y2) affine, handling doubling and cancellation. */
static void j_add_aff(jac *p, const fq *x2, const fq *y2) {
  if (f_is_zero(&p->Z)) {
    p->X = *x2;
    p->Y = *y2;
    p->Z = F_ONE;
    return;
  }
  fq Z1Z1, U2, S2, H, r, t;
  f_mul(&Z1Z1, &p->Z, &p->Z);
  f_mul(&U2, x2, &Z1Z1);
  f_mul(&S2, y2, &p->Z);
  f_mul(&S2, &S2, &Z1Z1);
  f_sub(&H, &U2, &p->X);
  f_sub(&r, &S2, &p->Y);
  if (f_is_zero(&H)) {
    if (f_is_zero(&r)) {
      j_dbl(p);
    } else {
      p->Z = F_ZERO;
    }
    return;
  }
  fq HH, HHH, V, X3, Y3, Z3;
  f_mul(&HH, &H, &H);
  f_mul(&HHH, &HH, &H);
  f_mul(&V, &p->X, &HH);
  f_mul(&X3, &r, &r);
  f_sub(&X3, &X3, &HHH);
  f_sub(&X3, &X3, &V);
  f_sub(&X3, &X3, &V);
  f_sub(&t, &V, &X3);
  f_mul(&Y3, &r, &t);
  f_mul(&t, &p->Y, &HHH);
  f_sub(&Y3, &Y3, &t);
  f_mul(&Z3, &p->Z, &H);
  p->X = X3;
  p->Y = Y3;
  p->Z = Z3;
}

static void waff_from_orc(waff *w, const orc_g1_affine *a) {
  memcpy(w->x.v, a->x.data, 48);
  memcpy(w->y.v, a->y.data, 48);
  w->inf = a->infinity ? 1 : 0;
}
static void waff_to_orc(orc_g1_affine *a, const waff *w) {
  if (w->inf) {
    memset(a, 0, sizeof(*a));
    a->infinity = 1;
    return;
  }
  memcpy(a->x.data, w->x.v, 48);
  memcpy(a->y.data, w->y.v, 48);
  a->infinity = 0;
}

void orc_g1_add(orc_g1_affine *out, const orc_g1_affine *a, const orc_g1_affine *b) {
  ensure_init();
  waff p, q;
  waff_from_orc(&p, a);
  waff_from_orc(&q, b);
  w_add(&p, &q);
  waff_to_orc(out, &p);
}
void orc_g1_scalar_mul(orc_g1_affine *out, const orc_g1_affine *p, const orc_bigint256 *k) {
  ensure_init();
  waff a, r;
  waff_from_orc(&a, p);
  w_mul(&r, &a, k);
  waff_to_orc(out, &r);
}
void orc_g1_generator(orc_g1_affine *out) {
  ensure_init();
  memcpy(out->x.data, F_GX.v, 48);
  memcpy(out->y.data, F_GY.v, 48);
  out->infinity = 0;
}

typedef struct {
  orc_g1_affine *out;
  const orc_bigint256 *k;
  size_t begin, end;
} gen_arg;
static void *gen_worker(void *vp) {
  gen_arg *a = (gen_arg *)vp;
  for (size_t i = a->begin; i < a->end; i++) {
    jac acc;
    acc.X = F_ZERO;
    acc.Y = F_ONE;
    acc.Z = F_ZERO;
    for (int bit = 255; bit >= 0; bit--) {
      j_dbl(&acc);
      if ((a->k[i].data[bit / 64] >> (bit % 64)) & 1) j_add_aff(&acc, &F_GX, &F_GY);
    }
    if (f_is_zero(&acc.Z)) {
      memset(&a->out[i], 0, sizeof(orc_g1_affine));
      a->out[i].infinity = 1;
      continue;
    }
    fq zi, zi2, x, y;
    f_inv(&zi, &acc.Z);
    f_mul(&zi2, &zi, &zi);
    f_mul(&x, &acc.X, &zi2);
    f_mul(&zi2, &zi2, &zi);
    f_mul(&y, &acc.Y, &zi2);
    memcpy(a->out[i].x.data, x.v, 48);
    memcpy(a->out[i].y.data, y.v, 48);
    a->out[i].infinity = 0;
  }
  return NULL;
}
void orc_g1_mul_generator_many(orc_g1_affine *out, const orc_bigint256 *k, size_t n, int threads) {
  ensure_init();
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t tid[64];
  gen_arg args[64];
  size_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    args[t].out = out;
    args[t].k = k;
    args[t].begin = t * per < n ? t * per : n;
    args[t].end = (t + 1) * per < n ? (t + 1) * per : n;
    pthread_create(&tid[t], NULL, gen_worker, &args[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

void orc_g1_projective_to_affine(orc_g1_affine *out, const orc_g1_projective *p) {
  ensure_init();
  fq z;
  memcpy(z.v, p->z.data, 48);
  if (f_is_zero(&z)) {
    memset(out, 0, sizeof(*out));
    out->infinity = 1;
    return;
  }
  /* ark-ec 0.3 short Weierstrass projective is Jacobian: x/z^2, y/z^3. */
  fq zi, zi2, x, y;
  f_inv(&zi, &z);
  f_mul(&zi2, &zi, &zi);
  memcpy(x.v, p->x.data, 48);
  memcpy(y.v, p->y.data, 48);
  f_mul(&x, &x, &zi2);
  f_mul(&zi2, &zi2, &zi);
  f_mul(&y, &y, &zi2);
  memcpy(out->x.data, x.v, 48);
  memcpy(out->y.data, y.v, 48);
  out->infinity = 0;
}

/* Import + precompute one point.  Returns convertible flag; for a
 * non-convertible point, *w holds its Weierstrass affine form. */
static int import_point(const orc_g1_affine *p, fq host[3], waff *w) {
  if (p->infinity) {
    /* setToTwistedEdwardsIdentity, then preComputeFPGA of (0,1,1,0). */
    f_mul(&host[0], &F_ONE, &F_INV2);
    f_mul(&host[1], &F_ONE, &F_INV2);
    host[2] = F_ZERO;
    return 1;
  }
  fq x, y;
  memcpy(x.v, p->x.data, 48);
  memcpy(y.v, p->y.data, 48);
  w->x = x;
  w->y = y;
  w->inf = 0;
  w_to_mont(&x, &y);
  if (!mont_to_te(&x, &y)) return 0;
  fq t;
  f_mul(&t, &x, &y); /* twistedEdwardsAffineToExtended */
  /* preComputeFPGA :606-618 */
  fq a, b, d4;
  f_sub(&a, &y, &x);
  f_mul(&a, &a, &F_INV2);
  f_add(&b, &y, &x);
  f_mul(&b, &b, &F_INV2);
  f_mul(&d4, &F_D, &F_FOUR);
  host[0] = a;
  host[1] = b;
  f_mul(&host[2], &t, &d4);
  return 1;
}

int orc_msm_precompute_point(const orc_g1_affine *p, uint64_t out[18]) {
  ensure_init();
  fq host[3];
  waff w;
  int ok = import_point(p, host, &w);
  if (ok) {
    for (int i = 0; i < 3; i++) memcpy(out + 6 * i, host[i].v, 48);
  } else {
    memset(out, 0, 18 * 8);
  }
  return ok;
}

/* ---- Pippenger ---- */
typedef struct {
  const fq *host;      /* npoints x 3 */
  const uint8_t *conv; /* npoints */
  const int16_t *dig;  /* npoints x 20 signed digits */
  size_t npoints;
  int window;
  xyzt result; /* triangle sum of this window, standard extended coords */
} win_arg;

static void *window_worker(void *vp) {
  win_arg *a = (win_arg *)vp;
  int w = a->window;
  int nb = orc_msm_num_buckets(w);
  xyzt *buckets = (xyzt *)malloc(sizeof(xyzt) * nb);
  for (int b = 0; b < nb; b++) {
    /* to_fpga_internal_representation of the identity = (2, 2, 4, 0). */
    buckets[b].x = F_TWO;
    buckets[b].y = F_TWO;
    buckets[b].z = F_FOUR;
    buckets[b].t = F_ZERO;
  }
  for (size_t i = 0; i < a->npoints; i++) {
    if (!a->conv[i]) continue; /* scalar zeroed, driver.cpp:398-406 */
    int d = a->dig[i * 20 + w];
    if (d == 0) continue; /* libs/pippenger/src/controller.ml:128-132 */
    int mag = d < 0 ? -d : d;
    te_add_precomputed(&buckets[mag - 1], &a->host[3 * i], d < 0);
  }
  /* driver.cpp:318-339: buckets descending, running/accum triangle sums. */
  xyzt accum, running;
  te_identity(&accum);
  te_identity(&running);
  for (int b = nb - 1; b >= 0; b--) {
    post_compute(&buckets[b]);
    te_add(&running, &buckets[b]);
    te_add(&accum, &running);
  }
  a->result = accum;
  free(buckets);
  return NULL;
}

typedef struct {
  const orc_g1_affine *points;
  const orc_bigint256 *scalars;
  fq *host;
  uint8_t *conv;
  int16_t *dig;
  size_t begin, end;
} prep_arg;

static void *prep_worker(void *vp) {
  prep_arg *a = (prep_arg *)vp;
  for (size_t i = a->begin; i < a->end; i++) {
    if (a->points) {
      waff w;
      a->conv[i] = (uint8_t)import_point(&a->points[i], &a->host[3 * i], &w);
    }
    if (a->scalars) {
      uint32_t mag[20];
      uint8_t neg[20];
      orc_msm_digits(&a->scalars[i], mag, neg);
      for (int k = 0; k < 20; k++) a->dig[i * 20 + k] = (int16_t)(neg[k] ? -(int)mag[k] : (int)mag[k]);
    }
  }
  return NULL;
}

static void run_prep(const orc_g1_affine *points, const orc_bigint256 *scalars, fq *host, uint8_t *conv,
                     int16_t *dig, size_t n, int threads) {
  pthread_t tid[64];
  prep_arg args[64];
  size_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    args[t].points = points;
    args[t].scalars = scalars;
    args[t].host = host;
    args[t].conv = conv;
    args[t].dig = dig;
    args[t].begin = t * per < n ? t * per : n;
    args[t].end = (t + 1) * per < n ? (t + 1) * per : n;
    pthread_create(&tid[t], NULL, prep_worker, &args[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

void orc_msm(const orc_g1_affine *points, size_t npoints, const orc_bigint256 *scalars, size_t num_batches,
             orc_g1_projective *out, int threads) {
  ensure_init();
  if (threads < 1) threads = 1;
  if (threads > 20) threads = 20;
  size_t n = npoints ? npoints : 1;
  fq *host = (fq *)calloc(3 * n, sizeof(fq));
  uint8_t *conv = (uint8_t *)calloc(n, 1);
  int16_t *dig = (int16_t *)calloc(20 * n, sizeof(int16_t));
  run_prep(points, NULL, host, conv, dig, npoints, threads);
  for (size_t b = 0; b < num_batches; b++) {
    const orc_bigint256 *sc = scalars + b * npoints;
    run_prep(NULL, sc, host, conv, dig, npoints, threads);
    win_arg wa[20];
    pthread_t tid[20];
    for (int w = 0; w < 20; w++) {
      wa[w].host = host;
      wa[w].conv = conv;
      wa[w].dig = dig;
      wa[w].npoints = npoints;
      wa[w].window = w;
    }
    /* Up to `threads` windows at a time. */
    for (int w0 = 0; w0 < 20; w0 += threads) {
      int w1 = w0 + threads < 20 ? w0 + threads : 20;
      for (int w = w0; w < w1; w++) pthread_create(&tid[w], NULL, window_worker, &wa[w]);
      for (int w = w0; w < w1; w++) pthread_join(tid[w], NULL);
    }
    /* driver.cpp:340-347 + pippenger.h:36-45: accum doubled bit_offset
     * times, added into the final result. */
    xyzt final;
    te_identity(&final);
    for (int w = 0; w < 20; w++) {
      xyzt acc = wa[w].result;
      int off = orc_msm_window_offset(w);
      for (int i = 0; i < off; i++) {
        xyzt self = acc;
        te_add(&acc, &self);
      }
      te_add(&final, &acc);
    }
    /* extendedTwistedEdwardsToWeierstrass (affine), :427-439 / :359-372. */
    waff res;
    {
      fq x = final.x, y = final.y;
      if (f_is_zero(&final.z)) {
        x = F_ZERO;
        y = F_ONE;
      } else {
        f_div(&x, &final.x, &final.z);
        f_div(&y, &final.y, &final.z);
      }
      if (!te_to_mont(&x, &y)) {
        res.inf = 1;
        res.x = F_ZERO;
        res.y = F_ZERO;
      } else {
        mont_to_w(&x, &y);
        res.x = x;
        res.y = y;
        res.inf = 0;
      }
    }
    /* Non-convertible points: weierstrassMultiplyAndAdd, driver.cpp:349-372. */
    for (size_t i = 0; i < npoints; i++) {
      if (conv[i]) continue;
      waff pt, kp;
      waff_from_orc(&pt, &points[i]);
      w_mul(&kp, &pt, &sc[i]);
      w_add(&res, &kp);
    }
    /* copy_to_rust_type :667-690: (x, y, 1) in Montgomery form, z = 0 for infinity. */
    orc_g1_projective *o = &out[b];
    memset(o, 0, sizeof(*o));
    if (!res.inf) {
      memcpy(o->x.data, res.x.v, 48);
      memcpy(o->y.data, res.y.v, 48);
      memcpy(o->z.data, F_ONE.v, 48);
    }
  }
  free(host);
  free(conv);
  free(dig);
}

/* sum_i s_i * k_i mod r for 256-bit little-endian integers (closed-form MSM
 * checks: with P_i = k_i G the MSM equals (sum s_i k_i mod r) G). */
void orc_dot_mod_r(const orc_bigint256 *s, const orc_bigint256 *k, size_t n, uint64_t out[4]) {
  static const uint64_t R_W[4] = {0x0a11800000000001ull, 0x59aa76fed0000001ull, 0x60b44d1e5c37b001ull,
                                  0x12ab655e9a2ca556ull};
  uint64_t acc[10] = {0};
  for (size_t i = 0; i < n; i++) {
    uint64_t prod[8] = {0};
    for (int a = 0; a < 4; a++) {
      u128 c = 0;
      for (int b = 0; b < 4; b++) {
        c = (u128)s[i].data[a] * k[i].data[b] + prod[a + b] + (uint64_t)(c >> 64);
        prod[a + b] = (uint64_t)c;
      }
      prod[a + 4] = (uint64_t)(c >> 64);
    }
    u128 c = 0;
    for (int j = 0; j < 10; j++) {
      c = (u128)acc[j] + (j < 8 ? prod[j] : 0) + (uint64_t)(c >> 64);
      acc[j] = (uint64_t)c;
    }
  }
  /* acc mod r by shift-and-subtract from the top bit down. */
  uint64_t rem[5] = {0};
  for (int bit = 639; bit >= 0; bit--) {
    /* rem = rem * 2 + bit */
    uint64_t carry = (acc[bit / 64] >> (bit % 64)) & 1;
    for (int j = 0; j < 5; j++) {
      uint64_t nc = rem[j] >> 63;
      rem[j] = (rem[j] << 1) | carry;
      carry = nc;
    }
    /* if rem >= r: rem -= r */
    int ge = rem[4] != 0;
    if (!ge) {
      ge = 1;
      for (int j = 3; j >= 0; j--) {
        if (rem[j] != R_W[j]) {
          ge = rem[j] > R_W[j];
          break;
        }
      }
    }
    if (ge) {
      u128 br = 0;
      for (int j = 0; j < 5; j++) {
        u128 d = (u128)rem[j] - (j < 4 ? R_W[j] : 0) - br;
        rem[j] = (uint64_t)d;
        br = (d >> 64) & 1;
      }
    }
  }
  for (int j = 0; j < 4; j++) out[j] = rem[j];
}

/*
 * CPU oracle for the Goldilocks NTT and the BLS12-377 G1 MSM.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark use it, and only as the checker.
 *
 * Each function restates the reference's algorithm and cites the reference
 * file:line it follows (paths relative to the fyquah/hardcaml_zprize
 * checkout).  How each function is pinned (known-answer vectors, closed
 * forms, the reference's own gf.h compiled into oracle/_ref) is listed in the
 * test files that exercise it.
 */
#ifndef ZP_ORACLE_H
#define ZP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- NTT --- */

/* Goldilocks field, p = 2^64 - 2^32 + 1 (zprize/ntt/host/gf.h:6-8). */
uint64_t orc_gl_add(uint64_t a, uint64_t b);
uint64_t orc_gl_sub(uint64_t a, uint64_t b);
uint64_t orc_gl_mul(uint64_t a, uint64_t b);
uint64_t orc_gl_pow(uint64_t a, uint64_t e);
uint64_t orc_gl_inv(uint64_t a);
/* 7^((p-1)/2^i): the value zprize/ntt/host/gf.h:65-99 tabulates as OMEGA[i]. */
uint64_t orc_gl_root(unsigned log_n);

/* In-place radix-2 DIT NTT with roots OMEGA (zprize/ntt/host/reference.cpp:87-137).
 * dir = 0: forward (X_k = sum a_j w^jk).  dir = 1: inverse with roots OMEGA^-1
 * (libs/hardcaml_ntt/src/reference_model.ml:54 forward_dit) followed by the
 * n^-1 scale, so that inverse(forward(x)) == x. */
void orc_ntt_inplace(uint64_t *a, unsigned log_n, int dir);
/* Same transform, unscaled inverse (dir = 1 without the n^-1 factor). */
void orc_ntt_inplace_unscaled(uint64_t *a, unsigned log_n, int dir);
/* Bailey 4-step model (libs/hardcaml_ntt/src/reference_model.ml:92-105). */
void orc_ntt_four_step(uint64_t *a, unsigned log_n, unsigned log_rows);
/* Many independent transforms (for throughput baselines); threads <= 0 uses 1. */
void orc_ntt_many(uint64_t *a, unsigned log_n, size_t count, int dir, int threads);
/* The REVERSE debug core of the FPGA NTT build (zprize/ntt/host/reference.cpp:60-84):
 * n = 4^log_row_size elements, src and dst distinct. */
void orc_ntt_reverse(uint64_t *dst, const uint64_t *src, unsigned log_row_size);

/* ---------------------------------------------------------------- MSM --- */

/* Layouts of zprize/msm_pippenger/host/msm_compute/include/bls12_377_g1/rust_types.h:8-83.
 * Coordinates are 6 little-endian 64-bit words in Montgomery form, R = 2^384. */
typedef struct { uint64_t data[4]; } orc_bigint256;
typedef struct { uint64_t data[6]; } orc_bigint384;
typedef struct { orc_bigint384 x, y; _Bool infinity; } orc_g1_affine;
typedef struct { orc_bigint384 x, y, z; } orc_g1_projective;

/* Fq arithmetic on Montgomery-form words. */
void orc_fq_mul(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]);
void orc_fq_add(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]);
void orc_fq_sub(uint64_t out[6], const uint64_t a[6], const uint64_t b[6]);
void orc_fq_inv(uint64_t out[6], const uint64_t a[6]);
void orc_fq_to_mont(uint64_t out[6], const uint64_t plain[6]);
void orc_fq_from_mont(uint64_t out[6], const uint64_t mont[6]);

/* Signed-digit window decomposition (zprize/msm_pippenger/hardcaml/test/top/utils.ml:243-288,
 * windows from zprize/msm_pippenger/hardcaml/src/config_utils.ml:7-41).
 * mag[w] in [0, num_buckets(w)], neg[w] in {0,1}. */
void orc_msm_digits(const orc_bigint256 *scalar, uint32_t mag[20], uint8_t neg[20]);
int orc_msm_window_bits(int w);
int orc_msm_window_offset(int w);
int orc_msm_num_buckets(int w);

/* Point import and host precompute (bls12_377_g1.h:646-659, 406-415, 606-618).
 * Returns 1 if the point has a twisted-Edwards image, 0 otherwise.
 * out = ((y-x)/2, (y+x)/2, 4d*t) in Montgomery form. */
int orc_msm_precompute_point(const orc_g1_affine *p, uint64_t out[18]);

/* Full MSM, restating driver.cpp:306-375 + pippenger.h + the non-convertible
 * fallback (driver.cpp:349-372).  scalars: num_batches * npoints, out: num_batches. */
void orc_msm(const orc_g1_affine *points, size_t npoints, const orc_bigint256 *scalars,
             size_t num_batches, orc_g1_projective *out, int threads);

/* Weierstrass affine helpers (bls12_377_g1.h:493-742), Montgomery form in and out. */
void orc_g1_add(orc_g1_affine *out, const orc_g1_affine *a, const orc_g1_affine *b);
void orc_g1_scalar_mul(orc_g1_affine *out, const orc_g1_affine *p, const orc_bigint256 *k);
void orc_g1_generator(orc_g1_affine *out);
/* P_i = k_i * G for a batch (fixed-base generation used by tests). */
void orc_g1_mul_generator_many(orc_g1_affine *out, const orc_bigint256 *k, size_t n, int threads);
/* sum_i s_i k_i mod r (closed-form MSM check). */
void orc_dot_mod_r(const orc_bigint256 *s, const orc_bigint256 *k, size_t n, uint64_t out[4]);
/* Projective (x*R, y*R, R)/(z=0) -> affine with infinity flag. */
void orc_g1_projective_to_affine(orc_g1_affine *out, const orc_g1_projective *p);

#ifdef __cplusplus
}
#endif
#endif

/*
 * zpgpu: MI355X (gfx950) Goldilocks NTT and BLS12-377 G1 MSM behind a C ABI.
 *
 * Drop-in boundary for the two hot paths of fyquah/hardcaml_zprize.  Every
 * entry point names the reference interface it replaces (paths relative to
 * the reference checkout).  Plain pointers and sizes only.
 *
 * Conventions
 *   - Functions returning int return a zp_status; zp_last_error() gives the
 *     text of the calling thread's last failure.
 *   - "d_" pointers are device (HBM) pointers; "stream" is a hipStream_t
 *     passed as void* (NULL = the null stream).
 *   - One handle is not reentrant; distinct handles may be used from
 *     distinct threads.
 */
#ifndef ZPGPU_H
#define ZPGPU_H

#include <stddef.h>
#include <stdint.h>
#include <sys/types.h>

#ifdef __cplusplus
#define ZP_BOOL bool
extern "C" {
#else
#include <stdbool.h>
#define ZP_BOOL bool
#endif

typedef enum {
  ZP_OK = 0,
  ZP_ERR_INVALID_ARG = 1, /* bad size / pointer / direction */
  ZP_ERR_HIP = 2,         /* a HIP runtime call failed */
  ZP_ERR_OUT_OF_MEMORY = 3,
  ZP_ERR_NO_BUFFER = 4,   /* all in-flight buffers are in use */
  ZP_ERR_STATE = 5,       /* call out of order (e.g. free of a free buffer) */
  ZP_ERR_TOO_LARGE = 6,   /* data_length > transform size, budget exceeded */
  ZP_ERR_NO_DEVICE = 7
} zp_status;

typedef enum { ZP_NTT_FORWARD = 0, ZP_NTT_INVERSE = 1 } zp_ntt_dir;

const char *zp_last_error(void);
/* Library / kernel build identification, e.g. "zpgpu gfx950". */
const char *zp_version(void);

/* ------------------------------------------------------------------ NTT --
 * Goldilocks p = 2^64 - 2^32 + 1, natural-order input and output, canonical
 * values in [0, p).  FORWARD: X_k = sum_j a_j w^(jk), w = 7^((p-1)/n), the
 * transform of zprize/ntt/host/reference.cpp:87-152 (the hardware core's).
 * INVERSE: n^-1 * sum_j X_j w^(-jk), so INVERSE(FORWARD(a)) == a.
 */
typedef struct zp_ntt_plan zp_ntt_plan;

/* NttFpgaDriver(NttFpgaDriverArg::create_ntt_2_24 / _2_18 / _2_12) +
 * load_xclbin (zprize/ntt/host/driver.h:110-134, driver.cpp:107-138,364-454).
 * log_n in [1, 26].  Allocates twiddle tables, one device scratch vector and
 * `num_buffers` (<= 64; 0 means 8, driver.cpp:133) in-flight user buffers. */
int zp_ntt_create(unsigned log_n, unsigned num_buffers, zp_ntt_plan **out);
/* NttFpgaDriverArg::create_reverse (driver.cpp:100-104): a plan of n = 4^log_row_size
 * elements running the REVERSE debug core (driver.h:18) instead of the transform:
 * the two 4-step passes with each row DFT replaced by reversing the row
 * (zprize/ntt/host/reference.cpp:60-84).  Every plan entry point works on it;
 * dir must be 0. */
int zp_ntt_create_reverse(unsigned log_row_size, unsigned num_buffers, zp_ntt_plan **out);
int zp_ntt_destroy(zp_ntt_plan *plan);
/* NttFpgaDriver::input_vector_size (driver.h:127). */
uint64_t zp_ntt_size(const zp_ntt_plan *plan);
/* NttFpgaDriver::max_num_buffers_in_flight (driver.h:142-144). */
unsigned zp_ntt_max_buffers_in_flight(const zp_ntt_plan *plan);

/* The hot path: transform n u64 resident in HBM, in place, on `stream`
 * (the device work of enqueue_phase1_work + enqueue_phase2_work,
 * zprize/ntt/host/driver.cpp:201-231).  Asynchronous. */
int zp_ntt_inplace_device(zp_ntt_plan *plan, uint64_t *d_data, int dir, void *stream);

/* Measurement (no counterpart in the reference): runs the same in-place
 * transform `reps` times on `stream` with HIP events around each of its two
 * kernel launches (pass 1: column DFTs + 4-step twiddle; pass 2: row DFTs)
 * and returns the mean milliseconds of each.  Blocks until done.  For
 * log_n <= 12 (a single launch) ms_pass2 is 0. */
int zp_ntt_time_passes(zp_ntt_plan *plan, uint64_t *d_data, int dir, int reps, void *stream, float *ms_pass1,
                       float *ms_pass2);

/* NttFpgaDriver::simple_evaluate (driver.h:188-197, driver.cpp:332-344):
 * host in, host out, blocking; input shorter than n is zero padded;
 * data_length > n is ZP_ERR_TOO_LARGE (driver.cpp:289-297). */
int zp_ntt_simple_evaluate(zp_ntt_plan *plan, uint64_t *out, const uint64_t *in, uint64_t data_length,
                           int dir);

/* Asynchronous buffer API, NttFpgaDriver::request_buffer / transfer_* /
 * enqueue_evaluation_async / wait_for_result / free_buffer
 * (driver.h:146-186, driver.cpp:148-286).  Buffers are ids in
 * [0, max_buffers_in_flight).  Input and output are pinned host vectors.
 * An evaluation copies the input in, transforms and copies the output back;
 * copies run on their own streams, so one buffer's transfers overlap the
 * (serialized) transforms of the others.  A buffer still in flight may be
 * enqueued again (bench_throughput.cpp:76-86 does so): the evaluations run
 * in order, each from the buffer's input.  zp_ntt_wait returns once the
 * buffer's most recent evaluation is copied out; waiting again is a no-op.
 * zp_ntt_wait on a buffer never enqueued since its request is ZP_ERR_STATE. */
int zp_ntt_request_buffer(zp_ntt_plan *plan, int *buffer_id); /* ZP_ERR_NO_BUFFER when all busy */
uint64_t *zp_ntt_buffer_input(zp_ntt_plan *plan, int buffer_id);
uint64_t *zp_ntt_buffer_output(zp_ntt_plan *plan, int buffer_id);
int zp_ntt_enqueue(zp_ntt_plan *plan, int buffer_id, int dir);
int zp_ntt_wait(zp_ntt_plan *plan, int buffer_id);
int zp_ntt_free_buffer(zp_ntt_plan *plan, int buffer_id); /* ZP_ERR_STATE if not in use */

/* Batched form: `count` transforms of mixed sizes and directions, all
 * resident in HBM, in place.  Jobs are grouped by (log_n, dir) and each
 * group runs as one launch per pass with the job index on grid.y; with
 * use_graph != 0 the whole sequence is captured into a HIP graph at the
 * first launch and replayed afterwards.  log_ns[i] in [1, 24]; d_ptrs[i]
 * device pointers; dirs[i] in {0,1}.  No reference counterpart: the
 * reference runs one transform per enqueue (bench_throughput.cpp:61-107). */
typedef struct zp_ntt_batch zp_ntt_batch;
int zp_ntt_batch_create(uint64_t *const *d_ptrs, const unsigned *log_ns, const int *dirs, size_t count,
                        zp_ntt_batch **out);
int zp_ntt_batch_launch(zp_ntt_batch *batch, int use_graph, void *stream);
int zp_ntt_batch_destroy(zp_ntt_batch *batch);
/* create + eager launch + synchronize + destroy. */
int zp_ntt_batch_device(uint64_t *const *d_ptrs, const unsigned *log_ns, const int *dirs, size_t count,
                        void *stream);

/* Tile choices of one size class (log_n in [10, 23]), for single transforms
 * (batched = 0: plans, the buffer API, zp_ntt_inplace_device) or for the
 * groups of a batch (batched = 1: zp_ntt_batch_*, zp_mixed_batch_*): the
 * column length 2^log_n1 (pass 1 sub-DFTs; log_n1 and log_n - log_n1 in
 * [1, 12]; 12 puts pass 1 on the register-resident 4096-point column kernel)
 * and log2 of the sequences per workgroup of pass 1 (log_t1 <= min(13 -
 * log_n1, log_n - log_n1); unused when log_n1 is 12) and pass 2 (log_t2 <=
 * min(13 - (log_n - log_n1), log_n1)).  A field of -1 takes the built-in
 * policy (balanced split up to 2^20, log_n1 = 12 from 2^21); (-1, -1, -1)
 * restores it.  Process-wide; takes
 * effect for transforms launched (or batches captured) afterwards.  No
 * reference counterpart (the FPGA core's layout is fixed): this is the hook
 * of the size-class autotuner (hardcaml_zprize_amd/autotune.py).
 * zp_ntt_get_tiles returns the effective choice (built-in fields resolved). */
int zp_ntt_set_tiles(unsigned log_n, int batched, int log_n1, int log_t1, int log_t2);
int zp_ntt_get_tiles(unsigned log_n, int batched, int *log_n1, int *log_t1, int *log_t2);

/* ------------------------------------------------------------------ MSM --
 * Layouts of zprize/msm_pippenger/host/msm_compute/include/bls12_377_g1/rust_types.h:8-83.
 * Base-field words are little-endian 64-bit limbs in Montgomery form with
 * R = 2^384 (the arkworks 0.3 representation); scalars are plain 256-bit
 * little-endian integers < r.
 */
typedef struct { uint64_t data[4]; } biginteger256_t;
typedef struct { uint64_t data[6]; } biginteger384_t;
typedef struct { biginteger384_t x, y; ZP_BOOL infinity; } g1_affine_t;   /* 104 bytes */
typedef struct { biginteger384_t x, y, z; } g1_projective_t;             /* 144 bytes */

typedef struct zp_msm_context zp_msm_context;

/* extern "C" Driver *msm_init(const char *xclbin, ssize_t xclbin_len,
 *                             g1_affine_t *rust_points, ssize_t npoints)
 * (zprize/msm_pippenger/host/driver/driver.cpp:603-617).  The xclbin string
 * (not NUL terminated) is an optional option string here, e.g. "window=16";
 * it may be NULL/empty.  "coords=canonical" takes x and y as plain integers
 * < q (the form ark-serialize files hold) instead of x R, y R; the default
 * is "coords=montgomery".  Points are converted on the GPU to the
 * twisted-Edwards precomputed form and stay resident in HBM.
 * "precompute=<c>" (c in [8, 24]) or "precompute=auto" also keeps the
 * window-shifted copies 2^(c w) P_i for every window w of the uniform c-bit
 * layout (ceil(253 / c) copies, 144 bytes each: 106 GB for 2^26 points at
 * the automatic c = 23); an MSM over the table, or over a prefix large enough
 * for it to pay, then adds every window's digits into one bucket set.  Fails
 * with ZP_ERR_OUT_OF_MEMORY when the copies do not fit.  Results are the same
 * bytes either way.  Returns NULL on failure (zp_last_error() says why). */
zp_msm_context *msm_init(const char *xclbin, ssize_t xclbin_len, g1_affine_t *rust_points, ssize_t npoints);

/* extern "C" void msm_mult(Driver *driver, g1_projective_t *out,
 *                          uint64_t num_batches, biginteger256_t *ptr_scalars)
 * (driver.cpp:619-636).  scalars[b * npoints + i]; out[b] = (x*R, y*R, R)
 * for an affine result (x, y), or z = 0 for the point at infinity. */
void msm_mult(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches, biginteger256_t *ptr_scalars);

/* Not in the reference (its Driver is never freed): releases the context. */
void msm_free(zp_msm_context *ctx);

/* Status-returning forms of the two entry points above. */
int zp_msm_init(const char *options, size_t options_len, const g1_affine_t *points, size_t npoints,
                zp_msm_context **out);
/* msm_init with the points already resident in HBM (d_points: npoints g1_affine_t). */
int zp_msm_init_device(const char *options, size_t options_len, const g1_affine_t *d_points, size_t npoints,
                       zp_msm_context **out);
int zp_msm_mult(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches, const biginteger256_t *scalars);

/* In place, s -> s * 2^256 mod r for n scalars < r in HBM: the integers the
 * ZPrize harness multiplies by when it loads Fr values from its data files
 * (ark-ff 0.3 keeps them as s R_r and the harness transmutes them to
 * BigInteger256, zprize/msm_pippenger/test_fpga_harness/src/util.rs:115-118,
 * 150-173).  Asynchronous on `stream`. */
int zp_fr_to_montgomery_device(biginteger256_t *d_scalars, uint64_t n, void *stream);
/* Scalars already resident in HBM (num_batches * npoints * 32 bytes). */
int zp_msm_mult_device(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches,
                       const biginteger256_t *d_scalars, void *stream);
/* MSM over the first n points of the context's table (n <= npoints). */
int zp_msm_mult_prefix_device(zp_msm_context *ctx, size_t n, g1_projective_t *out, const biginteger256_t *d_scalars,
                              void *stream);
size_t zp_msm_num_points(const zp_msm_context *ctx);
/* zp_msm_mult_prefix_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The per-call working memory
 * (digits, sort entries, counters, bucket and unit partial sums, window sums,
 * results) lives in a static 64 MiB device array per GPU, held until the
 * process ends; a job needing more (roughly beyond 2^15 points at the default
 * window) fails with ZP_ERR_TOO_LARGE and a message giving the sizes.  Calls
 * on one GPU share that array and are serialized by a per-device lock; the
 * array keeps the last call's contents.  Blocks until *out is written (host
 * memory).  Allocates nothing (no hipMalloc/hipFree). */
int zp_msm_mult_prefix_noalloc(zp_msm_context *ctx, size_t n, g1_projective_t *out,
                               const biginteger256_t *d_scalars, void *stream);
/* Width c of the context's window-shifted table ("precompute="), 0 if it has none. */
int zp_msm_precompute_width(const zp_msm_context *ctx);

/* Batched form: many MSMs of mixed sizes, each over a prefix of a resident
 * table (tables may differ between jobs), run as ONE launch sequence: the
 * jobs' windows and buckets are concatenated, so digits, counting sort,
 * bucket sums and window reductions of every job share each kernel launch.
 * Blocks until every out is written.  One workspace runs one batch at a
 * time; concurrent batches use distinct workspaces and streams.  No
 * reference counterpart (the reference runs one MSM batch after another,
 * driver.cpp:495-600). */
typedef struct {
  zp_msm_context *ctx;               /* resident point table */
  size_t n;                          /* use points [0, n) */
  const biginteger256_t *d_scalars;  /* n scalars in HBM */
  g1_projective_t *out;              /* host output, one point */
  int window;                        /* window bits in [8, 24]; 0 = size-class default */
} zp_msm_job;
typedef struct zp_msm_workspace zp_msm_workspace;
int zp_msm_workspace_create(zp_msm_workspace **out);
int zp_msm_workspace_destroy(zp_msm_workspace *ws);
int zp_msm_batch_device(zp_msm_workspace *ws, const zp_msm_job *jobs, size_t count, const char *options,
                        size_t options_len, void *stream);
/* The batched form as ONE HIP graph: NTT jobs (in place on d_ptrs, as
 * zp_ntt_batch_create) and MSM jobs (as zp_msm_batch_device; options as
 * there, e.g. "window=16"; per-job windows from an autotuner table go in
 * zp_msm_job.window) are planned, allocated and captured once here; the
 * NTT groups and the MSM launch sequence are parallel branches of the graph.
 * zp_mixed_batch_launch replays the whole graph on `stream` (asynchronous:
 * it recomputes every job from the current vectors and scalars);
 * zp_mixed_batch_wait synchronizes `stream` and writes each MSM job's out.
 * The NTT vectors, scalar buffers and MSM contexts must outlive the batch.
 * No reference counterpart (the reference runs one transform / one MSM at a
 * time on its device, zprize/ntt/host/bench_throughput.cpp:61-107,
 * zprize/msm_pippenger/host/driver/driver.cpp:495-600). */
typedef struct zp_mixed_batch zp_mixed_batch;
int zp_mixed_batch_create(uint64_t *const *ntt_ptrs, const unsigned *ntt_log_ns, const int *ntt_dirs,
                          size_t ntt_count, const zp_msm_job *msm_jobs, size_t msm_count, const char *options,
                          size_t options_len, zp_mixed_batch **out);
int zp_mixed_batch_launch(zp_mixed_batch *b, void *stream);
int zp_mixed_batch_wait(zp_mixed_batch *b, void *stream);
int zp_mixed_batch_destroy(zp_mixed_batch *b);

/* Number of loaded points without a twisted-Edwards image (handled by the
 * Weierstrass fallback, driver.cpp:349-372). */
size_t zp_msm_num_nonconvertible(const zp_msm_context *ctx);

/* Seeded base-point generation on the GPU: out[i] = k_i * G (G the G1
 * generator), affine Montgomery form; d_k: n scalars (device), d_out: n
 * g1_affine_t (device).  Test/benchmark input generation for tables too large
 * to build on the host. */
int zp_g1_mul_generator_device(const biginteger256_t *d_k, g1_affine_t *d_out, size_t n, void *stream);

/* -------------------------------------------------------- diagnostics --
 * Component checks through the device code paths (used by the parity tests). */
/* out[i] = a[i] * b[i] for wire-form (Montgomery R = 2^384) Fq values, 6 words each. */
int zp_debug_fq_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
/* Signed window digits of n host scalars under the layout an options string
 * selects ("window=reference" = the reference's 20 windows,
 * config_utils.ml:7-41); digits[i * W + w], W written to *num_windows
 * (digits may be NULL to query W). */
int zp_debug_msm_digits(const char *options, size_t options_len, const biginteger256_t *scalars, size_t n,
                        int32_t *digits, int *num_windows);
/* The plan an MSM over the first n points of ctx's table takes: *shifted = 1
 * when it runs as one window over the window-shifted table (shared buckets),
 * *window_bits = its window width (the first window's for the reference
 * layout), *num_windows = the digit windows per scalar. */
int zp_debug_msm_plan(const zp_msm_context *ctx, size_t n, int *shifted, int *window_bits, int *num_windows);
/* The resident precomputed points [i0, i0 + count) as wire-form triples
 * (y - x, y + x, 2d x y) of the twisted-Edwards image, 18 words each. */
int zp_debug_msm_precomputed(zp_msm_context *ctx, size_t i0, size_t count, uint64_t *out);

/* ------------------------------------------------------------ utilities -- */
/* Measured issue rate, lane-operations per second over the whole device:
 * kind 0 = v_mad_u64_u32 (the Fq multiply's multiply-add), kind 1 = a
 * full-rate 32-bit VALU op.  The MSM roofline peak is kind 0. */
int zp_probe_valu_rate(int kind, double *lane_ops_per_s);
int zp_device_malloc(void **ptr, size_t bytes);
int zp_device_free(void *ptr);
int zp_memcpy_h2d(void *dst, const void *src, size_t bytes);
int zp_memcpy_d2h(void *dst, const void *src, size_t bytes);
int zp_device_synchronize(void);
/* Streams (default priority), pinned host memory and async copies for hosts
 * that drive the library without another GPU runtime. */
int zp_stream_create(void **stream);
int zp_stream_destroy(void *stream);
int zp_stream_synchronize(void *stream);
int zp_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream);
int zp_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream);
int zp_host_alloc(void **ptr, size_t bytes);
/* Large pageable host arrays <-> HBM through pinned staging: the array is
 * split over 8 host threads, each copying through its own pair of 16 MiB
 * pinned slabs while the other slab's DMA runs on its own stream (~2x a
 * plain pageable copy).  Both calls block until the data has arrived;
 * zp_transfer_d2h first waits for `after_stream` (may be NULL).  One
 * transfer at a time per object. */
typedef struct zp_transfer zp_transfer;
int zp_transfer_create(zp_transfer **out);
int zp_transfer_destroy(zp_transfer *t);
int zp_transfer_h2d(zp_transfer *t, void *d_dst, const void *h_src, size_t bytes);
int zp_transfer_d2h(zp_transfer *t, void *h_dst, const void *d_src, size_t bytes, void *after_stream);
int zp_host_free(void *ptr);
int zp_device_memory_info(size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif

/* A C caller of the drop-in MSM boundary, binding it the way the ZPrize Rust
 * harness does (zprize/msm_pippenger/test_fpga_harness/src/lib.rs:19-33,
 * 40-85): msm_init(XCLBIN path and its length, points, npoints), then
 * msm_mult(ctx, out, batch_size = #scalars / npoints, scalars), one output
 * point per batch.  Plain C against include/zpgpu.h.
 *
 *   msm_harness <points.bin> <scalars.bin> <out.bin> <xclbin-string>
 *
 * points.bin: npoints g1_affine_t (104 bytes); scalars.bin: batches * npoints
 * biginteger256_t; out.bin receives batches g1_projective_t (144 bytes). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

#include "zpgpu.h"

static void *read_file(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  void *buf = malloc(n > 0 ? (size_t)n : 1);
  if (buf && n > 0 && fread(buf, 1, (size_t)n, f) != (size_t)n) {
    free(buf);
    buf = NULL;
  }
  fclose(f);
  *len = n > 0 ? (size_t)n : 0;
  return buf;
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s points.bin scalars.bin out.bin xclbin\n", argv[0]);
    return 2;
  }
  if (sizeof(g1_affine_t) != 104 || sizeof(g1_projective_t) != 144 || sizeof(biginteger256_t) != 32) {
    fprintf(stderr, "layout mismatch\n");
    return 3;
  }
  size_t plen = 0, slen = 0;
  g1_affine_t *points = (g1_affine_t *)read_file(argv[1], &plen);
  biginteger256_t *scalars = (biginteger256_t *)read_file(argv[2], &slen);
  if (!points || !scalars || plen % 104 || slen % 32) {
    fprintf(stderr, "bad input files\n");
    return 4;
  }
  const size_t npoints = plen / 104, nscalars = slen / 32;
  if (npoints == 0 || nscalars % npoints) {
    fprintf(stderr, "length mismatch\n"); /* lib.rs:63-66 panics the same way */
    return 5;
  }
  const size_t batches = nscalars / npoints;
  /* The harness passes the XCLBIN variable's bytes and length, no NUL. */
  const size_t xlen = strlen(argv[4]);
  char *xclbin = (char *)malloc(xlen + 1);
  memcpy(xclbin, argv[4], xlen);
  xclbin[xlen] = 'X'; /* not a terminator: the library must use the length */
  zp_msm_context *ctx = msm_init(xclbin, (ssize_t)xlen, points, (ssize_t)npoints);
  if (!ctx) {
    fprintf(stderr, "msm_init returned NULL\n");
    return 6;
  }
  g1_projective_t *out = (g1_projective_t *)calloc(batches, sizeof(g1_projective_t));
  msm_mult(ctx, out, (uint64_t)batches, scalars);
  FILE *f = fopen(argv[3], "wb");
  if (!f || fwrite(out, sizeof(g1_projective_t), batches, f) != batches) {
    fprintf(stderr, "cannot write output\n");
    return 7;
  }
  fclose(f);
  msm_free(ctx);
  free(out);
  free(xclbin);
  free(points);
  free(scalars);
  printf("msm_harness: %zu points, %zu batches\n", npoints, batches);
  return 0;
}

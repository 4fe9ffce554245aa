// Harness of ours that exposes the reference's own Goldilocks class
// (zprize/ntt/host/gf.h:10-63) and OMEGA table (gf.h:65-99) through a C ABI,
// so the tests can check the oracle's field arithmetic against the
// reference's compiled code.  Test infrastructure only; built into
// oracle/_ref/ from the header where it lies in the reference checkout.
#include "gf.h"

extern "C" {
uint64_t ref_gf_add(uint64_t a, uint64_t b) { return (GF(a) + GF(b)).to_uint64(); }
uint64_t ref_gf_sub(uint64_t a, uint64_t b) { return (GF(a) - GF(b)).to_uint64(); }
uint64_t ref_gf_mul(uint64_t a, uint64_t b) { return (GF(a) * GF(b)).to_uint64(); }
uint64_t ref_gf_omega(unsigned i) { return i < 33 ? OMEGA[i].to_uint64() : 0; }
void ref_gf_mul_many(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) out[i] = (GF(a[i]) * GF(b[i])).to_uint64();
}
}

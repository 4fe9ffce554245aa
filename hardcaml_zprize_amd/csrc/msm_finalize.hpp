// MSM finalisation on the GPU: the window combine (Horner over the window
// sums, zprize/msm_pippenger/host/driver/driver.cpp:340-347,
// host/msm_compute/include/bls12_377_g1/pippenger.h:36-45), the points
// without a twisted-Edwards image (driver.cpp:349-372), and the conversion to
// the wire form (x R, y R, R) or z = 0 (bls12_377_g1.h:427-439, 667-690).
//
// One job per 4-lane wave.  The combine is ~250 doublings in sequence; each
// doubling's 8 multiplications fall into two rounds of 4 independent ones
// (dbl-2008-hwcd: X^2, Y^2, Z^2, (X+Y)^2, then E F, G H, E H, F G), so the
// 4 lanes each compute one product of a round and exchange them by lane
// shuffles: a doubling costs two multiplications of latency instead of
// eight.  The additions between rounds are repeated in every lane.
//
// The points without an image are the curve's points of order 2 ((x, 0),
// x^3 = -1: three of them) and 4 (Montgomery u = -1: two of them).  For such
// a point P, S P = (S mod 4) P, so a group of equal exceptional points adds
// M[(S mod 4) - 1] from the table of P, 2P, 3P its context holds, where S is
// the exact sum of the group's scalars (k_nc_sums).  msm_init rejects a point
// without an image whose 4P is not the identity (not on the curve).
#pragma once
#include <hip/hip_runtime.h>

#include "ec.hpp"
#include "msm_kernels.hpp"

namespace zp {

// m P for m = 1, 2, 3 of one exceptional point P (affine Weierstrass, R'
// Montgomery form, canonical); inf[m - 1] = 1 when m P is the identity.
struct NcMultDev {
  F x[3], y[3];
  uint32_t inf[3];
};

// Finalisation of one job: window sums winsum[win0 .. win0 + W) (Horner
// doublings between windows w + 1 and w: wins[win0 + w].width), and ng
// exceptional-point groups with exact scalar sums nc_sums[5 (nc0 + g) ..]
// and multiples mult[g].
struct FinJobDev {
  uint32_t win0, W;
  uint32_t nc0, ng;
  const NcMultDev *mult;
};

__device__ __forceinline__ F shfl4(const F &v, int src) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = (uint32_t)__shfl((int)v.l[i], src, 4);
  return r;
}

// Every lane's product p; then the 4 products in every lane (lane shuffles:
// an exchange through LDS with 16-byte accesses and barriers measured 9 %
// slower per doubling).
__device__ __forceinline__ void swap4(const F &p, F &r0, F &r1, F &r2, F &r3) {
  r0 = shfl4(p, 0);
  r1 = shfl4(p, 1);
  r2 = shfl4(p, 2);
  r3 = shfl4(p, 3);
}

// The operand of lane `sel`, by masks: a select between the structures
// themselves becomes a select between their addresses, i.e. scratch memory.
__device__ __forceinline__ F sel4(int sel, const F &a, const F &b, const F &c, const F &d) {
  const uint32_t m0 = 0u - (uint32_t)(sel == 0), m1 = 0u - (uint32_t)(sel == 1);
  const uint32_t m2 = 0u - (uint32_t)(sel == 2), m3 = 0u - (uint32_t)(sel == 3);
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = (a.l[i] & m0) | (b.l[i] & m1) | (c.l[i] & m2) | (d.l[i] & m3);
  return r;
}

// a = 2a as ec_dbl, the 8 products spread over the 4 lanes.  Four carry
// passes instead of seven: X + Y, E and G stay unnormalized (limbs < 2^30),
// which f_mul takes when the other factor is normalized (C, AB, H, F are),
// and every value stays below 2^384 (E < 10q, G < 6q, F < 14q).
__device__ __forceinline__ void ec_dbl4(Xyzt &a, int sel) {
  F A, B, zz, S;
  swap4(f_sqr(sel4(sel, a.x, a.y, a.z, f_add_nc(a.x, a.y))), A, B, zz, S);
  const F C = f_add(zz, zz);                  // < 4q, normalized
  const F AB = f_add(A, B);                   // < 4q, normalized
  const F E = f_sub_nc_c<8>(S, AB);           // S + 8q - AB < 10q
  const F G = f_sub_nc_c<4>(B, A);            // B - A + 4q < 6q
  const F H = f_sub_kq_c<4>(f_zero(), AB);    // -(A + B) + 4q, normalized
  const F Fm = f_sub_kq_c<8>(G, C);           // G - C + 8q < 14q, normalized
  swap4(f_mul(sel4(sel, E, G, E, Fm), sel4(sel, Fm, H, H, G)), a.x, a.y, a.t, a.z);
}

// a += b as ec_add, the products of its first and last rounds spread over the lanes.
__device__ __forceinline__ void ec_add4(Xyzt &a, const Xyzt &b, int sel, const FConsts &K) {
  const F m1 = sel4(sel, f_sub_kq_c<2>(a.y, a.x), f_add(a.y, a.x), a.t, a.z);
  const F m2 = sel4(sel, f_sub_kq_c<2>(b.y, b.x), f_add(b.y, b.x), b.t, b.z);
  F A, B, tt, zz;
  swap4(f_mul(m1, m2), A, B, tt, zz);
  const F C = f_mul(tt, K.k2d);
  const F D = f_add(zz, zz);
  const F E = f_sub_kq_c<2>(B, A);
  const F H = f_add(B, A);
  const F Fm = f_sub_kq_c<2>(D, C);
  const F Gm = f_add(D, C);
  swap4(f_mul(sel4(sel, E, Gm, E, Fm), sel4(sel, Fm, H, H, Gm)), a.x, a.y, a.t, a.z);
}

__device__ __forceinline__ void store_words(const F &v, uint64_t *out6) {
  const Packed p = f_pack(f_canonical(v));
#pragma unroll
  for (int i = 0; i < 6; i++) out6[i] = (uint64_t)p.w[2 * i] | ((uint64_t)p.w[2 * i + 1] << 32);
}

// blockDim.x = 4: block j finalises job j into out[j] (g1_projective_t words:
// x[6], y[6], z[6]).  The constants are read from device memory (Kp) where
// used: held in registers, the ~1 KB FConsts would spill.
__global__ void __launch_bounds__(4) k_msm_finalize(const FinJobDev *jobs, const WinDev *wins, const Xyzt *winsum,
                                                    const uint64_t *nc_sums, uint64_t *out, const FConsts *Kp) {
  const FConsts &K = *Kp;
  const FinJobDev J = jobs[blockIdx.x];
  const int sel = threadIdx.x & 3;
  Xyzt acc = winsum[J.win0 + J.W - 1];
  for (int w = (int)J.W - 2; w >= 0; w--) {
    const uint32_t width = wins[J.win0 + w].width;
    for (uint32_t i = 0; i < width; i++) ec_dbl4(acc, sel);
    ec_add4(acc, winsum[J.win0 + w], sel, K);
  }
  // Twisted Edwards -> Montgomery -> Weierstrass affine, in R' form: the
  // identity (and the order-2 point with x = 0) has no affine image.
  Jac res;
  res.X = f_zero();
  res.Y = K.one;
  res.Z = f_zero();
  const F den = f_mul(f_sub_kq(acc.z, acc.y, K.q2), acc.x);
  if (!f_is_zero_canon(acc.z) && !f_is_zero_canon(den)) {
    const F inv = f_inv_fast(den, K);
    const F zpy = f_add(acc.z, acc.y);
    const F u = f_mul(f_mul(zpy, acc.x), inv);
    const F v = f_mul(f_mul(f_mul(zpy, acc.z), inv), K.ts);
    res.X = f_canonical(f_add(f_mul(u, K.inv_cb), K.ca_3cb));
    res.Y = f_canonical(f_mul(v, K.inv_cb));
    res.Z = K.one;
  }
  bool affine = true;
  for (uint32_t g = 0; g < J.ng; g++) {
    const uint32_t m = (uint32_t)(nc_sums[5 * (uint64_t)(J.nc0 + g)] & 3);
    if (m == 0 || J.mult[g].inf[m - 1]) continue;
    jac_add_aff(res, J.mult[g].x[m - 1], J.mult[g].y[m - 1], K);
    affine = false;
  }
  if (sel != 0) return;
  uint64_t *o = out + 18 * (uint64_t)blockIdx.x;
  if (f_is_zero_canon(res.Z)) {
#pragma unroll
    for (int i = 0; i < 18; i++) o[i] = 0;
    return;
  }
  F x = res.X, y = res.Y;
  if (!affine) {
    const F zi = f_inv_fast(res.Z, K), zi2 = f_sqr(zi);
    x = f_mul(res.X, zi2);
    y = f_mul(res.Y, f_mul(zi2, zi));
  }
  store_words(f_mul(x, K.to_wire), o);
  store_words(f_mul(y, K.to_wire), o + 6);
  store_words(f_mul(K.one, K.to_wire), o + 12);
}

}  // namespace zp

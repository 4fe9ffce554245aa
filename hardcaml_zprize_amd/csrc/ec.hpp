// BLS12-377 G1 point arithmetic on gfx950 (and on the host for the final
// conversion), in the twisted-Edwards model of the reference
// (a = -1, d, k = 2d; zprize/msm_pippenger/host/msm_compute/include/bls12_377_g1/bls12_377_g1.h:31-54,
// libs/twisted_edwards/model/twisted_edwards_curve.ml).
//
// Accumulators are extended coordinates (X : Y : Z : T), x = X/Z, y = Y/Z,
// T = XY/Z, each coordinate a field value < 2q.  Base points are stored as
// the precomputed affine triple (y - x, y + x, 2d x y), canonical, packed
// (144 bytes): the mixed addition then needs 7 multiplications (hwcd-3 with
// Z2 = 1, the operation of twisted_edwards_curve.ml:145-170 without the
// reference's 1/2 and 4 scalings), and subtracting a point swaps the first
// two entries and negates the third (twisted_edwards_curve.ml:151-152).
#pragma once
#include "fq377.hpp"

namespace zp {

struct Xyzt {
  F x, y, z, t;
};

struct PrePt {
  F ymx, ypx, t2d;
};

struct PackedPre {
  Packed ymx, ypx, t2d;
};

ZP_HD Xyzt ec_identity(const FConsts &K) {
  Xyzt r;
  r.x = f_zero();
  r.y = K.one;
  r.z = K.one;
  r.t = f_zero();
  return r;
}

// acc += (neg ? -P : P), P in precomputed form.  7 multiplications.
// Sums feeding a multiply stay un-normalized where the other factor is
// normalized (f_mul bound), so only E, H and G take a carry pass.
ZP_HD void ec_madd(Xyzt &a, const PrePt &p, bool neg, const FConsts &K) {
  const F &p0 = neg ? p.ypx : p.ymx;
  const F &p1 = neg ? p.ymx : p.ypx;
  const F A = f_mul(f_sub_nc(a.y, a.x, K.q4), p0);  // (Y - X)(y2 - x2) ; Y + 4q - X < 6q
  const F B = f_mul(f_add_nc(a.y, a.x), p1);        // (Y + X)(y2 + x2) ; < 4q
  const F C = f_mul(a.t, p.t2d);                    // T 2d t2
  const F D = f_add_nc(a.z, a.z);                   // 2 Z ; < 4q, limbs < 2^29
  F E = f_sub_nc(B, A, K.q4);                       // < 6q
  F H = f_add_nc(B, A);                             // < 4q
  f_carry(E);
  f_carry(H);
  // -P negates t2: F and G swap.  D + 4q - C < 8q, D + C < 6q.
  F Fm = neg ? f_add_nc(D, C) : f_sub_nc(D, C, K.q4);
  F Gm = neg ? f_sub_nc(D, C, K.q4) : f_add_nc(D, C);
  f_carry(Gm);
  a.x = f_mul(E, Fm);
  a.y = f_mul(Gm, H);
  a.t = f_mul(E, H);
  a.z = f_mul(Fm, Gm);
}

// a += b, both extended (hwcd-3 unified addition, bls12_377_g1.h:449-491).
// 9 multiplications.  Complete for the points this code adds.
ZP_HD void ec_add(Xyzt &a, const Xyzt &b, const FConsts &K) {
  const F A = f_mul(f_sub_kq(a.y, a.x, K.q2), f_sub_kq(b.y, b.x, K.q2));
  const F B = f_mul(f_add(a.y, a.x), f_add(b.y, b.x));
  const F C = f_mul(f_mul(a.t, b.t), K.k2d);
  const F zz = f_mul(a.z, b.z);
  const F D = f_add(zz, zz);
  const F E = f_sub_kq(B, A, K.q2);
  const F H = f_add(B, A);
  const F Fm = f_sub_kq(D, C, K.q2);
  const F Gm = f_add(D, C);
  a.x = f_mul(E, Fm);
  a.y = f_mul(Gm, H);
  a.t = f_mul(E, H);
  a.z = f_mul(Fm, Gm);
}

// a = 2a (dbl-2008-hwcd with a = -1).  8 multiplications.
ZP_HD void ec_dbl(Xyzt &a, const FConsts &K) {
  const F A = f_sqr(a.x);
  const F B = f_sqr(a.y);
  const F zz = f_sqr(a.z);
  const F C = f_add(zz, zz);                          // < 4q
  const F S = f_sqr(f_add(a.x, a.y));                 // (X+Y)^2
  const F AB = f_add(A, B);                           // < 4q
  const F E = f_sub_kq(S, AB, K.q4);                  // < 6q
  const F G = f_sub_kq(B, A, K.q2);                   // D + B = B - A ; < 4q
  const F Fm = f_sub_kq(G, C, K.q4);                  // < 8q
  const F H = f_sub_kq(f_zero(), AB, K.q4);           // D - B = -(A + B) ; < 4q
  a.x = f_mul(E, Fm);
  a.y = f_mul(G, H);
  a.t = f_mul(E, H);
  a.z = f_mul(Fm, G);
}

ZP_HD PrePt pre_unpack(const PackedPre &p) {
  PrePt r;
  r.ymx = f_unpack(p.ymx);
  r.ypx = f_unpack(p.ypx);
  r.t2d = f_unpack(p.t2d);
  return r;
}

ZP_HD PackedPre pre_pack(const PrePt &p) {
  PackedPre r;
  r.ymx = f_pack(f_canonical(p.ymx));
  r.ypx = f_pack(f_canonical(p.ypx));
  r.t2d = f_pack(f_canonical(p.t2d));
  return r;
}

// Affine twisted-Edwards (x, y) to the precomputed triple.
ZP_HD PrePt pre_from_affine(const F &x, const F &y, const FConsts &K) {
  PrePt r;
  r.ymx = f_sub_kq(y, x, K.q2);
  r.ypx = f_add(y, x);
  r.t2d = f_mul(f_mul(x, y), K.d2);
  return r;
}

// Field halving: a / 2 for normalized a < 8q (adds q when odd, then shifts).
ZP_HD F f_half(const F &a) {
  F t = a;
  if (t.l[0] & 1) {
#pragma unroll
    for (int i = 0; i < 14; i++) t.l[i] += q28(i);
    f_carry(t);
  }
  F r;
#pragma unroll
  for (int i = 0; i < 13; i++) r.l[i] = (t.l[i] >> 1) | ((t.l[i + 1] & 1) << 27);
  r.l[13] = t.l[13] >> 1;
  return r;
}

// Precomputed triple back to extended coordinates (x = (p1 - p0)/2, y = (p1 + p0)/2, Z = 1).
ZP_HD Xyzt ext_from_pre(const PrePt &p, bool neg, const FConsts &K) {
  Xyzt r;
  F x = f_half(f_sub_kq(p.ypx, p.ymx, K.q2));
  const F y = f_half(f_add(p.ypx, p.ymx));
  if (neg) x = f_sub_kq(f_zero(), x, K.q4);
  r.x = x;
  r.y = y;
  r.z = K.one;
  r.t = f_mul(x, y);
  return r;
}

// ---- Weierstrass (y^2 = x^3 + 1) Jacobian arithmetic: the points without a
// twisted-Edwards image and the final result (bls12_377_g1.h:493-566) ----
struct Jac {
  F X, Y, Z;  // Z == 0: infinity
};

ZP_HD bool f_zero_c(const F &a) { return f_is_zero_canon(a); }

ZP_HD void jac_dbl(Jac &p, const FConsts &K) {
  if (f_zero_c(p.Z)) return;
  F A = f_sqr(p.X), B = f_sqr(p.Y), C = f_sqr(B);
  F t = f_sqr(f_add(p.X, B));
  t = f_sub_kq(f_sub_kq(t, A, K.q2), C, K.q2);
  F D = f_canonical(f_add(t, t));
  F E = f_canonical(f_add(f_add(A, A), A));
  F Fv = f_sqr(E);
  F X3 = f_canonical(f_sub_kq(f_sub_kq(Fv, D, K.q2), D, K.q2));
  F C8 = f_canonical(f_add(C, C));
  C8 = f_canonical(f_add(C8, C8));
  C8 = f_canonical(f_add(C8, C8));
  F Y3 = f_sub_kq(f_mul(E, f_sub_kq(D, X3, K.q2)), C8, K.q2);
  F Z3 = f_mul(p.Y, p.Z);
  Z3 = f_add(Z3, Z3);
  p.X = f_canonical(X3);
  p.Y = f_canonical(Y3);
  p.Z = f_canonical(Z3);
}

// p += (x2, y2) affine.
ZP_HD void jac_add_aff(Jac &p, const F &x2, const F &y2, const FConsts &K) {
  if (f_zero_c(p.Z)) {
    p.X = x2;
    p.Y = y2;
    p.Z = K.one;
    return;
  }
  F Z1Z1 = f_sqr(p.Z);
  F U2 = f_mul(x2, Z1Z1);
  F S2 = f_mul(f_mul(y2, p.Z), Z1Z1);
  F H = f_canonical(f_sub_kq(U2, p.X, K.q2));
  F r = f_canonical(f_sub_kq(S2, p.Y, K.q2));
  if (f_zero_c(H)) {
    if (f_zero_c(r))
      jac_dbl(p, K);
    else
      p.Z = f_zero();
    return;
  }
  F HH = f_sqr(H), HHH = f_mul(HH, H), V = f_canonical(f_mul(p.X, HH));
  F X3 = f_sqr(r);
  X3 = f_canonical(f_sub_kq(X3, HHH, K.q2));
  X3 = f_canonical(f_sub_kq(X3, V, K.q2));
  X3 = f_canonical(f_sub_kq(X3, V, K.q2));
  F Y3 = f_sub_kq(f_mul(r, f_sub_kq(V, X3, K.q2)), f_mul(p.Y, HHH), K.q2);
  p.X = X3;
  p.Y = f_canonical(Y3);
  p.Z = f_canonical(f_mul(p.Z, H));
}

ZP_HD void jac_add_jac(Jac &p, const Jac &q, const FConsts &K) {
  if (f_zero_c(q.Z)) return;
  if (f_zero_c(p.Z)) {
    p = q;
    return;
  }
  // Normalise q to affine (rare path: a handful of points per call).
  F zi = f_inv(q.Z, K.one), zi2 = f_sqr(zi);
  F x = f_canonical(f_mul(q.X, zi2));
  F y = f_canonical(f_mul(q.Y, f_mul(zi2, zi)));
  jac_add_aff(p, x, y, K);
}

}  // namespace zp

"""Test-data directories of the ZPrize MSM harness: points.bin, scalars.bin and
arkworks_results.bin (zprize/msm_pippenger/test_fpga_harness/src/util.rs:123-173,
`TEST_WRITE_DATA_TO` / `TEST_LOAD_DATA_FROM`).

The harness writes them with ark-serialize 0.3.0 `serialize_unchecked` (pinned in
test_fpga_harness/Cargo.lock; the crate's source is not part of the reference).
Its format, restated:

  Vec<T>       u64 little-endian element count, then the elements;
  Fq (377-bit) the canonical integer (not Montgomery), 48 bytes little endian;
  Fr (253-bit) the canonical integer, 32 bytes little endian;
  G1Affine     uncompressed: x (48 bytes), then y (48 bytes) whose last byte
               carries the short-Weierstrass flags: bit 6 = point at infinity,
               bit 7 = "positive y" (only the compressed form sets it).

Scalar semantics: the harness deserializes Fr values (stored in Montgomery
form, s R_r with R_r = 2^256 mod r) and transmutes the slice to BigInteger256
(util.rs:115-118), so the integers arkworks and the FPGA driver multiply by are
s R_r mod r, not s.  `harness_msm` converts the file's scalars that way on the
GPU (zp_fr_to_montgomery_device) before the MSM.  (Scaling the result by R_r
instead would be wrong for points outside the order-r subgroup, such as the
2-torsion point (-1, 0) of the harness's unit tests.)

Points are handed to msm_init with "coords=canonical", so a 2^26-point file
needs no host-side Montgomery conversion.

    python -m hardcaml_zprize_amd.arkdata verify DIR
    python -m hardcaml_zprize_amd.arkdata write DIR --log-n 10 --batches 2 --seed 1
"""
import argparse
import os
import random
import sys

import numpy as np

Q = 0x01AE3A4617C510EAC63B05C06CA1493B1A22D9F300F5138F1EF3622FBA094800170B5D44300000008508C00000000001
R = 0x12AB655E9A2CA55660B44D1E5C37B00159AA76FED00000010A11800000000001
R_FR = pow(2, 256, R)        # Montgomery radix of Fr
R_FQ = pow(2, 384, Q)        # Montgomery radix of Fq (the wire form of g1_affine_t)
FQ_BYTES, FR_BYTES = 48, 32
FLAG_INFINITY, FLAG_POSITIVE_Y = 0x40, 0x80

INFINITY = None  # affine point at infinity


class FormatError(ValueError):
    pass


def _vec(path, item_bytes):
    data = np.fromfile(path, dtype=np.uint8)
    if data.size < 8:
        raise FormatError(f"{path}: shorter than the u64 length prefix")
    n = int(data[:8].view("<u8")[0])
    if data.size != 8 + n * item_bytes:
        raise FormatError(f"{path}: {data.size} bytes, expected 8 + {n} x {item_bytes}")
    return n, data[8:].reshape(n, item_bytes)


def _words(b, nwords):
    """(n, 8 * nwords) little-endian bytes -> (n, nwords) uint64."""
    return np.ascontiguousarray(b).view("<u8").reshape(-1, nwords)


def read_points(path):
    """points.bin -> (x, y, infinity): x, y (n, 6) uint64 canonical words, infinity (n,) bool."""
    n, raw = _vec(path, 2 * FQ_BYTES)
    ybytes = raw[:, FQ_BYTES:].copy()
    flags = ybytes[:, -1] & 0xC0
    if np.any(flags == (FLAG_INFINITY | FLAG_POSITIVE_Y)):
        raise FormatError(f"{path}: invalid flag combination")
    ybytes[:, -1] &= 0x3F
    x = _words(raw[:, :FQ_BYTES], 6)
    y = _words(ybytes, 6)
    return x, y, (flags & FLAG_INFINITY) != 0


def read_scalars(path):
    """scalars.bin -> (n, 4) uint64, the canonical Fr integers of the file."""
    n, raw = _vec(path, FR_BYTES)
    return _words(raw, 4)


def _int(words):
    return int.from_bytes(np.ascontiguousarray(words, dtype="<u8").tobytes(), "little")


def read_results(path):
    """arkworks_results.bin -> list of affine points (x, y) as ints, or INFINITY."""
    x, y, inf = read_points(path)
    return [INFINITY if inf[i] else (_int(x[i]), _int(y[i])) for i in range(len(inf))]


def _header(n):
    return np.array([n], dtype="<u8").tobytes()


def write_points(path, x, y, infinity):
    """x, y: (n, 6) uint64 canonical words; infinity: (n,) bool."""
    x = np.ascontiguousarray(x, dtype="<u8").reshape(-1, 6)
    y = np.ascontiguousarray(y, dtype="<u8").reshape(-1, 6)
    inf = np.asarray(infinity, dtype=bool).reshape(-1)
    n = len(inf)
    rec = np.zeros((n, 2 * FQ_BYTES), dtype=np.uint8)
    rec[:, :FQ_BYTES] = x.view(np.uint8).reshape(n, FQ_BYTES)
    rec[:, FQ_BYTES:] = y.view(np.uint8).reshape(n, FQ_BYTES)
    rec[inf, -1] |= FLAG_INFINITY
    with open(path, "wb") as f:
        f.write(_header(n))
        f.write(rec.tobytes())


def write_scalars(path, scalars):
    s = np.ascontiguousarray(scalars, dtype="<u8").reshape(-1, 4)
    with open(path, "wb") as f:
        f.write(_header(len(s)))
        f.write(s.tobytes())


def _to_words(v, nwords):
    return np.frombuffer(v.to_bytes(8 * nwords, "little"), dtype="<u8")


def write_results(path, points):
    """points: list of affine (x, y) ints or INFINITY (written as x = 0, y = 1)."""
    n = len(points)
    x = np.zeros((n, 6), dtype="<u8")
    y = np.zeros((n, 6), dtype="<u8")
    inf = np.zeros(n, dtype=bool)
    for i, p in enumerate(points):
        if p is INFINITY:
            y[i] = _to_words(1, 6)
            inf[i] = True
        else:
            x[i] = _to_words(p[0], 6)
            y[i] = _to_words(p[1], 6)
    write_points(path, x, y, inf)


def to_affine_records(x, y, infinity):
    """Canonical coordinates -> g1_affine_t records for msm_init with coords=canonical."""
    from .msm import G1_AFFINE_DTYPE

    rec = np.zeros(len(infinity), dtype=G1_AFFINE_DTYPE)
    rec["x"] = x
    rec["y"] = y
    rec["infinity"] = np.asarray(infinity, dtype=np.uint8)
    return rec


# ---- affine short-Weierstrass arithmetic over Fq (y^2 = x^3 + 1), host side
def _add(p1, p2):
    if p1 is INFINITY:
        return p2
    if p2 is INFINITY:
        return p1
    x1, y1 = p1
    x2, y2 = p2
    if x1 == x2:
        if (y1 + y2) % Q == 0:
            return INFINITY
        lam = 3 * x1 * x1 * pow(2 * y1, Q - 2, Q) % Q
    else:
        lam = (y2 - y1) * pow(x2 - x1, Q - 2, Q) % Q
    x3 = (lam * lam - x1 - x2) % Q
    return x3, (lam * (x1 - x3) - y1) % Q


def scalar_mul(k, p):
    acc, run = INFINITY, p
    while k:
        if k & 1:
            acc = _add(acc, run)
        run = _add(run, run)
        k >>= 1
    return acc


def wire_to_affine(rec):
    """A g1_projective_t record from msm_mult ((x R, y R, R), or z = 0) -> affine ints."""
    z = _int(rec["z"])
    if z == 0:
        return INFINITY
    zi = pow(z, Q - 2, Q)
    return _int(rec["x"]) * zi % Q, _int(rec["y"]) * zi % Q


def check_scalars(scalars):
    """ark-ff's deserializer rejects Fr values >= r; so does the loader."""
    w = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    rw = _to_words(R, 4)
    lt = np.zeros(len(w), dtype=bool)
    eq = np.ones(len(w), dtype=bool)
    for i in (3, 2, 1, 0):
        lt |= eq & (w[:, i] < rw[i])
        eq &= w[:, i] == rw[i]
    if not lt.all():
        raise FormatError(f"scalar {int(np.argmin(lt))} is not below r")


def harness_msm(ctx, scalars, batches):
    """The harness's result for each batch, sum_i (s_i R_r mod r) P_i, as affine ints."""
    import ctypes

    from ._lib import call
    from .msm import DeviceBuffer

    sc = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    check_scalars(sc)
    d = DeviceBuffer(max(sc.nbytes, 32)).upload(sc)
    try:
        call("zp_fr_to_montgomery_device", ctypes.c_void_p(d.ptr), ctypes.c_uint64(len(sc)), ctypes.c_void_p(0))
        # the MSM runs on the context's own (non-blocking) stream
        call("zp_device_synchronize")
        out = ctx.mult_device(d.ptr, batches)
    finally:
        d.free()
    return [wire_to_affine(o) for o in out]


def load(dirname):
    x, y, inf = read_points(os.path.join(dirname, "points.bin"))
    scalars = read_scalars(os.path.join(dirname, "scalars.bin"))
    results = read_results(os.path.join(dirname, "arkworks_results.bin"))
    n = len(inf)
    # util.rs:167-170
    if n == 0 or len(scalars) % n != 0 or len(scalars) < n:
        raise FormatError(f"{len(scalars)} scalars for {n} points")
    return x, y, inf, scalars, results


def verify(dirname, options="", log=print):
    """Runs the GPU MSM on a harness test-data directory and compares every batch
    with arkworks_results.bin (tests/msm.rs:30-40).  Returns the list of per-batch
    booleans."""
    from .msm import MsmContext

    x, y, inf, scalars, expected = load(dirname)
    batches = len(scalars) // len(inf)
    if len(expected) != batches:
        raise FormatError(f"{len(expected)} results for {batches} batches")
    opt = "coords=canonical" + ("," + options if options else "")
    with MsmContext(to_affine_records(x, y, inf), opt) as ctx:
        got = harness_msm(ctx, scalars, batches)
    ok = [g == e for g, e in zip(got, expected)]
    for b, v in enumerate(ok):
        log(f"batch {b}: {'ok' if v else 'MISMATCH'}")
    return ok


def write_generated(dirname, log_n, batches, seed=1):
    """A harness test-data directory from seeded inputs: points k_i G (k_i uniform
    in [1, r)), scalars uniform in [0, r), results computed by the GPU MSM.  The
    points are converted to canonical form with host big integers, so this is
    meant for test sizes (up to ~2^20 points)."""
    from .msm import MsmContext, generate_points

    rng = random.Random(seed)
    n = 1 << log_n

    def uniform(count, lo):
        """count integers uniform in [lo, r) by rejection of 253-bit draws."""
        out = np.zeros((count, 4), dtype=np.uint64)
        for i in range(count):
            v = rng.getrandbits(253)
            while v >= R or v < lo:
                v = rng.getrandbits(253)
            out[i] = _to_words(v, 4)
        return out

    ks = uniform(n, 1)
    wire = generate_points(ks)
    rinv = pow(R_FQ, Q - 2, Q)
    x = np.zeros((n, 6), dtype="<u8")
    y = np.zeros((n, 6), dtype="<u8")
    for i in range(n):
        x[i] = _to_words(_int(wire["x"][i]) * rinv % Q, 6)
        y[i] = _to_words(_int(wire["y"][i]) * rinv % Q, 6)
    inf = wire["infinity"].astype(bool)
    scalars = uniform(n * batches, 0)
    with MsmContext(to_affine_records(x, y, inf), "coords=canonical") as ctx:
        results = harness_msm(ctx, scalars, batches)
    os.makedirs(dirname, exist_ok=True)
    write_points(os.path.join(dirname, "points.bin"), x, y, inf)
    write_scalars(os.path.join(dirname, "scalars.bin"), scalars)
    write_results(os.path.join(dirname, "arkworks_results.bin"), results)
    return ks, scalars, results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    v = sub.add_parser("verify")
    v.add_argument("dir")
    v.add_argument("--options", default="")
    w = sub.add_parser("write")
    w.add_argument("dir")
    w.add_argument("--log-n", type=int, default=10)
    w.add_argument("--batches", type=int, default=1)
    w.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)
    if args.cmd == "verify":
        ok = verify(args.dir, args.options)
        print("all batches match" if all(ok) else "MISMATCH")
        return 0 if all(ok) else 1
    write_generated(args.dir, args.log_n, args.batches, args.seed)
    print(f"wrote {args.dir}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

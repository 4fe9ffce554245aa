// MSM device kernels.  One launch sequence serves a batch of jobs (one job =
// one MSM over a prefix of a resident point table, with its own scalars and
// window layout); the single msm_mult is a batch of one.
//
//   k_digits_count   signed-digit window decomposition of every scalar
//                    (zprize/msm_pippenger/hardcaml/test/top/utils.ml:243-288,
//                    generalised to any window widths) and a histogram of
//                    global bucket keys; zero digits are skipped
//                    (libs/pippenger/src/controller.ml:128-132);
//   scan             bucket offsets and work-unit offsets;
//   k_digits_scatter counting-sort placement: vals[pos] = point | sign << 31;
//   k_fill_units     unit -> bucket map (a unit is <= U entries of one bucket);
//   k_accumulate     one thread per unit: mixed twisted-Edwards additions of
//                    the unit's points into an extended accumulator
//                    (the FPGA bucket pipeline, top.ml / mixed_add_precompute.ml);
//   k_unit_tree      per bucket: log-depth pairwise sum of its unit partials;
//   k_bucket_sum     per bucket: the reduced partial (identity if empty);
//   k_window_groups  per group of g buckets: running-sum reduction
//                    (driver.cpp:318-339, pippenger.h:25-33) done as
//                    sum_b (b - lo + 1) S_b + lo * sum_b S_b;
//   k_window_sum     one workgroup per window: tree sum of the group results.
// The window combine and the affine conversion run on the host (a few
// thousand field operations per job, the part of driver.cpp:340-375 left).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ec.hpp"

namespace zp {

static constexpr int MSM_THREADS = 256;

struct MsmJobDev {
  const PackedPre *pts;     // resident table of this job
  const uint8_t *conv;      // 0 = point without a twisted-Edwards image
  const uint64_t *scalars;  // n scalars, 4 words each
  uint32_t n;               // points used (a prefix of the table)
  uint32_t pt0;             // first global point slot of this job
  uint32_t win0, W;         // windows [win0, win0 + W)
  uint32_t bucket0;         // first global bucket of this job
  uint32_t chunks;          // point chunks of the sorted path (C)
  uint32_t chunk_len;       // points per chunk
  uint64_t dig0;            // first digit slot: digit(w, i) at dig0 + w * n + i
  uint64_t cnt0;            // first per-(bucket, chunk) counter: cnt0 + b_local * C + c
                            // (two-level path: per-(group, chunk), cnt0 + (g - g0) * C + c)
  uint32_t g0;              // two-level path: global index of the job's first bucket group
  // Digit generation: windows [dwin0, dwin0 + dW) of the scalar layout over
  // the dn real points.  Equal to (win0, W, n) except for a job on a
  // window-shifted table (pre = 1): there the dW digits of point i are the
  // entries i * dW + w of ONE sort window over n = dn * dW virtual points, and
  // virtual point i * dW + w is table entry 2^(off_w) P_i.
  uint32_t dwin0, dW, dn;
  uint32_t pre;
};

// One block of the sorted path: window w (global) of job j, chunk c.
struct SegDev {
  uint32_t job, win, chunk;
};

struct WinDev {
  uint32_t job;
  uint32_t off, width;  // bit offset and width of the window
  uint32_t nb, base;    // buckets, first global bucket
  uint32_t top;         // 1 = unsigned top window (slice + carry)
  uint32_t g0;          // two-level path: global index of the window's first bucket group
  uint32_t lbits;       // two-level path: group = bucket >> lbits
};

// Largest i in [0, n) with key[i] <= v (keys ascending, key[0] == 0).
__device__ __forceinline__ uint32_t upper_index(const uint32_t *key, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (key[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Bits [off, off + len) of a 253-bit scalar (bits >= 253 read as 0,
// config.ml:5-9 scalar_bits = 253).
__device__ __forceinline__ uint32_t scalar_slice(const uint64_t s[4], int off, int len) {
  const int w = off >> 6, b = off & 63;
  uint64_t lo = s[w] >> b;
  if (b && w + 1 < 4) lo |= s[w + 1] << (64 - b);
  if (off + len > 253) len = 253 - off;
  if (len <= 0) return 0;
  return (uint32_t)lo & ((len >= 32) ? 0xffffffffu : ((1u << len) - 1));
}

// Signed digit of one window given the incoming carry.
__device__ __forceinline__ int32_t window_digit(const uint64_t s[4], const WinDev &wd, uint32_t &carry) {
  const uint32_t slice = scalar_slice(s, wd.off, wd.width) + carry;
  if (wd.top) {
    carry = 0;
    return (int32_t)slice;
  }
  if (slice >= (1u << wd.width)) {
    carry = 1;
    return 0;
  }
  if (slice >= (1u << (wd.width - 1))) {
    carry = 1;
    return (int32_t)slice - (int32_t)(1u << wd.width);
  }
  carry = 0;
  return (int32_t)slice;
}

__device__ __forceinline__ void load_scalar(const uint64_t *scalars, uint64_t i, uint64_t s[4]) {
  const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(scalars + 4 * i);
  const ulonglong2 a = p[0], b = p[1];
  s[0] = a.x;
  s[1] = a.y;
  s[2] = b.x;
  s[3] = b.y;
}

struct BatchDev {
  const MsmJobDev *jobs;
  const uint32_t *job_pt0;  // J entries (ascending)
  const uint32_t *job_b0;   // J entries: first bucket per job (ascending)
  const WinDev *wins;
  uint32_t J;
  uint32_t total_points;
  uint32_t K;  // total buckets
};

// MODE 0: histogram; MODE 1: scatter; MODE 2: dump digits (diagnostics).
template <int MODE>
__global__ void __launch_bounds__(MSM_THREADS) k_digits(BatchDev B, uint32_t *cnt, uint32_t *cursor, uint32_t *vals,
                                                        int32_t *dump) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < B.total_points; g += gridDim.x * blockDim.x) {
    const uint32_t j = upper_index(B.job_pt0, B.J, g);
    const MsmJobDev jb = B.jobs[j];
    const uint32_t i = g - jb.pt0;
    if (MODE != 2 && !jb.conv[i]) continue;  // scalar zeroed, driver.cpp:398-406
    uint64_t s[4];
    load_scalar(jb.scalars, i, s);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < jb.W; w++) {
      const WinDev wd = B.wins[jb.win0 + w];
      const int32_t d = window_digit(s, wd, carry);
      if (MODE == 2) {
        dump[(uint64_t)i * jb.W + w] = d;
        continue;
      }
      if (d == 0) continue;
      const uint32_t key = wd.base + (uint32_t)(d < 0 ? -d : d) - 1;
      if (MODE == 0) {
        atomicAdd(&cnt[key], 1u);
      } else {
        const uint32_t pos = atomicAdd(&cursor[key], 1u);
        vals[pos] = i | (d < 0 ? 0x80000000u : 0u);
      }
    }
  }
}

// ---- sorted path: digits once, then per-(window, chunk) LDS histograms ----
// Signed digits of every (job, point, window), zero for points without a
// twisted-Edwards image (their scalars are handled by the host fallback).
__global__ void __launch_bounds__(MSM_THREADS) k_digits_make(BatchDev B, int32_t *digits) {
  // A block whose points all belong to one shifted-table job writes their
  // digits (point-major there) through LDS, as one contiguous run.
  __shared__ int32_t stage[MSM_THREADS * 32];
  for (uint32_t base = blockIdx.x * blockDim.x; base < B.total_points; base += gridDim.x * blockDim.x) {
    const uint32_t last = min(B.total_points, base + blockDim.x) - 1;
    const uint32_t jf = upper_index(B.job_pt0, B.J, base);
    const bool staged = jf == upper_index(B.job_pt0, B.J, last) && B.jobs[jf].pre;  // block-uniform
    const uint32_t g = base + threadIdx.x;
    if (g <= last) {
      const uint32_t j = upper_index(B.job_pt0, B.J, g);
      const MsmJobDev jb = B.jobs[j];
      const uint32_t i = g - jb.pt0;
      // digit (w, i) at dig0 + w * n + i; on a window-shifted table at dig0 + i * dW + w
      int32_t *dst = staged ? stage + threadIdx.x * jb.dW
                            : digits + jb.dig0 + (jb.pre ? (uint64_t)i * jb.dW : (uint64_t)i);
      const uint64_t step = jb.pre ? 1 : jb.dn;
      if (!jb.conv[i]) {
        for (uint32_t w = 0; w < jb.dW; w++) dst[w * step] = 0;
      } else {
        uint64_t s[4];
        load_scalar(jb.scalars, i, s);
        uint32_t carry = 0;
        for (uint32_t w = 0; w < jb.dW; w++) dst[w * step] = window_digit(s, B.wins[jb.dwin0 + w], carry);
      }
    }
    if (staged) {
      __syncthreads();
      const MsmJobDev jb = B.jobs[jf];
      const uint32_t cnt = (last - base + 1) * jb.dW;
      int32_t *dst = digits + jb.dig0 + (uint64_t)(base - jb.pt0) * jb.dW;
      for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) dst[k] = stage[k];
      __syncthreads();
    }
  }
}

// Per segment (job, window, chunk): LDS histogram of the chunk's digits,
// written to counts[cnt0 + b * C + c] (every bucket, zeros included).
__global__ void __launch_bounds__(1024) k_seg_count(BatchDev B, const SegDev *segs, const int32_t *digits,
                                                    uint32_t *counts) {
  extern __shared__ uint32_t hist[];
  const SegDev sg = segs[blockIdx.x];
  const MsmJobDev jb = B.jobs[sg.job];
  const WinDev wd = B.wins[sg.win];
  for (uint32_t b = threadIdx.x; b < wd.nb; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const uint32_t i0 = sg.chunk * jb.chunk_len, i1 = min(jb.n, i0 + jb.chunk_len);
  const int32_t *dig = digits + jb.dig0 + (uint64_t)(sg.win - jb.win0) * jb.n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int32_t d = dig[i];
    if (d != 0) atomicAdd(&hist[(uint32_t)(d < 0 ? -d : d) - 1], 1u);
  }
  __syncthreads();
  uint32_t *dst = counts + jb.cnt0 + (uint64_t)(wd.base - jb.bucket0) * jb.chunks + sg.chunk;
  for (uint32_t b = threadIdx.x; b < wd.nb; b += blockDim.x) dst[(uint64_t)b * jb.chunks] = hist[b];
}

// Per segment: cursors from the scanned counts, then LDS-atomic placement.
__global__ void __launch_bounds__(1024) k_seg_scatter(BatchDev B, const SegDev *segs, const int32_t *digits,
                                                      const uint32_t *scanned, uint32_t *vals) {
  extern __shared__ uint32_t cursor[];
  const SegDev sg = segs[blockIdx.x];
  const MsmJobDev jb = B.jobs[sg.job];
  const WinDev wd = B.wins[sg.win];
  const uint32_t *src = scanned + jb.cnt0 + (uint64_t)(wd.base - jb.bucket0) * jb.chunks + sg.chunk;
  for (uint32_t b = threadIdx.x; b < wd.nb; b += blockDim.x) cursor[b] = src[(uint64_t)b * jb.chunks];
  __syncthreads();
  const uint32_t i0 = sg.chunk * jb.chunk_len, i1 = min(jb.n, i0 + jb.chunk_len);
  const int32_t *dig = digits + jb.dig0 + (uint64_t)(sg.win - jb.win0) * jb.n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int32_t d = dig[i];
    if (d == 0) continue;
    const uint32_t pos = atomicAdd(&cursor[(uint32_t)(d < 0 ? -d : d) - 1], 1u);
    vals[pos] = i | (d < 0 ? 0x80000000u : 0u);
  }
}

// ---- two-level sorted path (windows of any width up to 2^21 buckets) ----
// Pass A partitions each window's entries by the high part of the bucket,
// group h = b >> lbits (lbits per window), with per-(window, chunk) LDS
// histograms as above; pass B sorts each group by the low lbits in a single
// block and writes the
// final bucket offsets.  Every write then lands in a run of consecutive
// slots -- a chunk's entries of one group, a group's entries of one bucket --
// where the single-level scatter writes about two entries per bucket per
// chunk.  Group sizes are data dependent, so pass B finds its range in the
// scanned pass-A counters on the device: group g's entries are
// [scanned[i], scanned[i + C]) with i its first counter.

__host__ __device__ __forceinline__ uint32_t bucket_groups(uint32_t nb, uint32_t lbits) { return ((nb - 1) >> lbits) + 1; }

// *ctr += 1 for each active lane; returns the value before this lane's
// increment.  When every active lane of the wave has the same counter, one
// lane adds the wave's count (hot buckets would otherwise serialize the LDS
// atomics 64 ways).
__device__ __forceinline__ uint32_t lds_claim(uint32_t *base, uint32_t key) {
  const uint64_t active = __ballot(1);
  const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
  if (__ballot(key == k0) == active) {
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(active >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)active, 0u));
    uint32_t first = 0;
    if (rank == 0) first = atomicAdd(&base[k0], (uint32_t)__popcll(active));
    // the leader is the lowest active lane
    first = __shfl(first, __ffsll((long long)active) - 1);
    return first + rank;
  }
  return atomicAdd(&base[key], 1u);
}

// Pass A count: per segment (job, window, chunk), the chunk's nonzero digits
// by group, to counts[cnt0 + (g - job g0) * C + c] (every group, zeros included).
__global__ void __launch_bounds__(1024) k_msd_count(BatchDev B, const SegDev *segs, const int32_t *digits,
                                                    uint32_t *counts) {
  extern __shared__ uint32_t hist[];
  const SegDev sg = segs[blockIdx.x];
  const MsmJobDev jb = B.jobs[sg.job];
  const WinDev wd = B.wins[sg.win];
  const uint32_t lbits = wd.lbits;
  const uint32_t H = bucket_groups(wd.nb, lbits);
  for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) hist[h] = 0;
  __syncthreads();
  const uint32_t i0 = sg.chunk * jb.chunk_len, i1 = min(jb.n, i0 + jb.chunk_len);
  const int32_t *dig = digits + jb.dig0 + (uint64_t)(sg.win - jb.win0) * jb.n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int32_t d = dig[i];
    if (d != 0) lds_claim(hist, ((uint32_t)(d < 0 ? -d : d) - 1) >> lbits);
  }
  __syncthreads();
  uint32_t *dst = counts + jb.cnt0 + (uint64_t)(wd.g0 - jb.g0) * jb.chunks + sg.chunk;
  for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) dst[(uint64_t)h * jb.chunks] = hist[h];
}

// Pass A scatter, grouped by (window, group), chunks in order within a
// group.  COMPACT (batches with a shifted-table job): 4-byte entries
// tmp[pos] = (i - chunk start) << (lbits + 1) | low lbits of the bucket << 1
// | sign (pass B recovers the chunk from the group's chunk offsets; the host
// keeps chunk_len << (lbits + 1) within 32 bits).  Otherwise 8-byte entries
// (point | sign << 31, bucket), which the per-window buckets' scatter of 2^11
// runs per chunk writes faster.  A window whose groups are single buckets
// (lbits = 0) is sorted by now: its entries go straight to vals (same slots,
// point | sign << 31), and pass B only writes their offsets.
template <bool COMPACT>
__global__ void __launch_bounds__(1024) k_msd_scatter(BatchDev B, const SegDev *segs, const int32_t *digits,
                                                      const uint32_t *scanned, uint32_t *tmp, uint32_t *vals) {
  extern __shared__ uint32_t cursor[];
  const SegDev sg = segs[blockIdx.x];
  const MsmJobDev jb = B.jobs[sg.job];
  const WinDev wd = B.wins[sg.win];
  const uint32_t lbits = wd.lbits;
  const uint32_t H = bucket_groups(wd.nb, lbits);
  const uint32_t *src = scanned + jb.cnt0 + (uint64_t)(wd.g0 - jb.g0) * jb.chunks + sg.chunk;
  for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) cursor[h] = src[(uint64_t)h * jb.chunks];
  __syncthreads();
  const uint32_t i0 = sg.chunk * jb.chunk_len, i1 = min(jb.n, i0 + jb.chunk_len);
  const int32_t *dig = digits + jb.dig0 + (uint64_t)(sg.win - jb.win0) * jb.n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int32_t d = dig[i];
    if (d == 0) continue;
    const uint32_t b = (uint32_t)(d < 0 ? -d : d) - 1;
    const uint32_t pos = lds_claim(cursor, b >> lbits);
    const uint32_t sign = d < 0 ? 1u : 0u;
    if (lbits == 0)
      vals[pos] = i | sign << 31;
    else if (COMPACT)
      tmp[pos] = (i - i0) << (lbits + 1) | (b & ((1u << lbits) - 1)) << 1 | sign;
    else
      reinterpret_cast<uint2 *>(tmp)[pos] = make_uint2(i | sign << 31, b);
  }
}

// Pass B: block g sorts bucket group g (window w, high part h) by the low
// lbits of its window: histogram, exclusive scan, off/cnt of the group's
// buckets, then placement into vals (point | sign << 31, the point being its
// chunk's first point plus the offset pass A stored).  The last group also
// writes off[K] (the total).  LDS: 2 * 2^max_lbits + 1024 + max_chunks + 1 words.
template <bool COMPACT>
__global__ void __launch_bounds__(1024) k_msd_group(BatchDev B, const uint32_t *win_g0, uint32_t Wt, uint32_t GT,
                                                    uint32_t max_lbits, const uint32_t *scanned, const uint32_t *tmp,
                                                    uint32_t *vals, uint32_t *off, uint32_t *cnt) {
  extern __shared__ uint32_t lds[];
  const uint32_t g = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const uint32_t w = upper_index(win_g0, Wt, g);
  const WinDev wd = B.wins[w];
  const MsmJobDev jb = B.jobs[wd.job];
  const uint64_t ci = jb.cnt0 + (uint64_t)(g - jb.g0) * jb.chunks;
  const uint32_t start = scanned[ci], end = scanned[ci + jb.chunks];
  const uint32_t b0 = (g - wd.g0) << wd.lbits;
  const uint32_t nl = min(1u << wd.lbits, wd.nb - b0);  // buckets in this group
  uint32_t *hist = lds, *cur = lds + (1u << max_lbits), *part = cur + (1u << max_lbits);
  uint32_t *cs = part + nt;  // the group's chunk offsets: chunk c's entries are [cs[c], cs[c + 1])
  if (wd.lbits == 0) {  // one bucket, placed by pass A
    if (tid == 0) {
      off[wd.base + b0] = start;
      cnt[wd.base + b0] = end - start;
      if (g == GT - 1) off[B.K] = end;
    }
    return;
  }
  const uint32_t lmask = (1u << wd.lbits) - 1;
  for (uint32_t l = tid; l < nl; l += nt) hist[l] = 0;
  for (uint32_t c = tid; c <= jb.chunks; c += nt) cs[c] = scanned[ci + c];
  __syncthreads();
  const uint2 *tmp2 = reinterpret_cast<const uint2 *>(tmp);
  for (uint32_t p = start + tid; p < end; p += nt)
    lds_claim(hist, COMPACT ? (tmp[p] >> 1) & lmask : tmp2[p].y & lmask);
  __syncthreads();
  // exclusive scan of hist[0, nl): each thread sums a run of `per`, a
  // Hillis-Steele scan over the 1024 run sums, then each run is expanded.
  const uint32_t per = (nl + nt - 1) / nt;
  const uint32_t l0 = min(nl, tid * per), l1 = min(nl, l0 + per);
  uint32_t sum = 0;
  for (uint32_t l = l0; l < l1; l++) sum += hist[l];
  part[tid] = sum;
  __syncthreads();
  for (uint32_t st = 1; st < nt; st <<= 1) {
    const uint32_t v = part[tid] + (tid >= st ? part[tid - st] : 0u);
    __syncthreads();
    part[tid] = v;
    __syncthreads();
  }
  uint32_t run = start + part[tid] - sum;  // exclusive prefix of this thread's run
  for (uint32_t l = l0; l < l1; l++) {
    const uint32_t c = hist[l];
    off[wd.base + b0 + l] = run;
    cnt[wd.base + b0 + l] = c;
    cur[l] = run;
    run += c;
  }
  if (g == GT - 1 && tid == 0) off[B.K] = end;
  __syncthreads();
  for (uint32_t p = start + tid; p < end && !COMPACT; p += nt) {
    const uint2 e = tmp2[p];
    vals[lds_claim(cur, e.y & lmask)] = e.x;
  }
  for (uint32_t p = start + tid; p < end && COMPACT; p += nt) {
    const uint32_t e = tmp[p];
    uint32_t lo = 0, hi = jb.chunks;  // cs[lo] <= p < cs[hi]: find p's chunk
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cs[mid] <= p)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t i = lo * jb.chunk_len + (e >> (wd.lbits + 1));
    vals[lds_claim(cur, (e >> 1) & lmask)] = i | (e & 1u) << 31;
  }
}

// off[k] = scanned[cnt0 + b * C] for bucket k (local b of its job); off[K] = total.
__global__ void k_bucket_offsets(BatchDev B, const uint32_t *scanned, uint64_t total_counters, uint32_t *off) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= B.K; k += gridDim.x * blockDim.x) {
    if (k == B.K) {
      off[k] = scanned[total_counters];
      continue;
    }
    const MsmJobDev jb = B.jobs[upper_index(B.job_b0, B.J, k)];
    off[k] = scanned[jb.cnt0 + (uint64_t)(k - jb.bucket0) * jb.chunks];
  }
}

__global__ void k_bucket_counts_from_off(const uint32_t *off, uint32_t K, uint32_t *cnt) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x)
    cnt[k] = off[k + 1] - off[k];
}

// nunits[k] = ceil(cnt[k] / U); k == K gets 0 (so the scan's last entry is the total).
// *umax (zeroed by the caller) becomes the largest unit count of any bucket
// when that exceeds 1 (a wave maximum, then one atomic per wave).
__global__ void k_units_per_bucket(const uint32_t *cnt, uint32_t K, uint32_t log_u, uint32_t *nunits,
                                   uint32_t *umax) {
  for (uint32_t base = blockIdx.x * blockDim.x; base <= K; base += gridDim.x * blockDim.x) {  // block-uniform
    const uint32_t k = base + threadIdx.x;
    uint32_t v = 0;
    if (k <= K) {
      v = k < K ? (cnt[k] + (1u << log_u) - 1) >> log_u : 0;
      nunits[k] = v;
    }
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    if ((threadIdx.x & 63) == 0 && v > 1) atomicMax(umax, v);
  }
}

// unit -> bucket: the bucket k with ustart[k] <= u < ustart[k + 1] (ustart is
// the exclusive scan of the per-bucket unit counts, K + 1 entries; empty
// buckets share their start with the next bucket, and upper_bound - 1 picks
// the last of them, the non-empty one).  One thread per unit, so a bucket
// with very many units is not written by one thread.
__global__ void k_fill_units(const uint32_t *ustart, uint32_t K, const uint32_t *total_units, uint32_t *ubucket) {
  const uint32_t total = *total_units;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = K + 1;  // first index with ustart > u lies in (lo, hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ustart[mid] <= u)
        lo = mid;
      else
        hi = mid;
    }
    ubucket[u] = lo;
  }
}

__global__ void __launch_bounds__(MSM_THREADS) k_accumulate(BatchDev B, const uint32_t *vals, const uint32_t *off,
                                                            const uint32_t *cnt, const uint32_t *ustart,
                                                            const uint32_t *ubucket, const uint32_t *total_units,
                                                            uint32_t log_u, Xyzt *partial, FConsts K) {
  const uint32_t total = *total_units;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    const uint32_t k = ubucket[u];
    const PackedPre *pts = B.jobs[upper_index(B.job_b0, B.J, k)].pts;
    const uint32_t s = u - ustart[k];
    const uint32_t b0 = off[k] + (s << log_u);
    const uint32_t b1 = min(off[k] + cnt[k], b0 + (1u << log_u));
    Xyzt acc = ec_identity(K);
    for (uint32_t e = b0; e < b1; e++) {
      const uint32_t v = vals[e];
      const PrePt p = pre_unpack(pts[v & 0x7fffffffu]);
      ec_madd(acc, p, (v >> 31) != 0, K);
    }
    partial[u] = acc;
  }
}

// One level of a segmented pairwise reduction of each bucket's unit partials:
// unit u (rank r within its bucket) absorbs unit u + stride when r is a
// multiple of 2 stride.  After ceil(log2(max units per bucket)) levels the
// bucket's first unit holds its sum, so a bucket whose entries fill many units
// (equal or structured scalars put every point of a window in one bucket)
// costs log depth instead of a serial walk by one thread.  Within a level no
// unit is both read and written.
__global__ void __launch_bounds__(MSM_THREADS) k_unit_tree(Xyzt *partial, const uint32_t *ubucket,
                                                           const uint32_t *ustart, const uint32_t *nunits,
                                                           const uint32_t *total_units, const uint32_t *umax,
                                                           uint32_t stride, FConsts K) {
  // The host launches enough levels for the worst case (one bucket holding a
  // window's every point); a level is needed only if some bucket has more
  // than `stride` units.
  if (stride >= *umax) return;
  const uint32_t total = *total_units;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    const uint32_t b = ubucket[u];
    const uint32_t r = u - ustart[b];
    if ((r & (2 * stride - 1)) == 0 && r + stride < nunits[b]) {
      Xyzt acc = partial[u];
      ec_add(acc, partial[u + stride], K);
      partial[u] = acc;
    }
  }
}

// Bucket value = its first unit's partial after the tree levels (identity if empty).
__global__ void __launch_bounds__(MSM_THREADS) k_bucket_sum(const Xyzt *partial, const uint32_t *ustart,
                                                            const uint32_t *nunits, uint32_t K_total,
                                                            Xyzt *buckets, FConsts K) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K_total; k += gridDim.x * blockDim.x)
    buckets[k] = nunits[k] ? partial[ustart[k]] : ec_identity(K);
}

// acc = m * p for a small non-negative integer m (double-and-add, MSB first).
__device__ __forceinline__ Xyzt ec_mul_small(const Xyzt &p, uint32_t m, const FConsts &K) {
  Xyzt r = ec_identity(K);
  if (m == 0) return r;
  const int top = 31 - __clz(m);
  r = p;
  for (int b = top - 1; b >= 0; b--) {
    ec_dbl(r, K);
    if ((m >> b) & 1) ec_add(r, p, K);
  }
  return r;
}

// Group j of window w covers buckets [j g, j g + g) clipped to nb[w];
// gstart[w] = first group of window w (W_total + 1 entries).  Empty buckets
// (cnt = 0) above a group's highest non-empty one are skipped: the running
// sum is still the identity there (a shifted-table job's shared buckets above
// the largest digit are all empty).
__global__ void __launch_bounds__(MSM_THREADS) k_window_groups(const Xyzt *buckets, const uint32_t *cnt,
                                                               const WinDev *wins, const uint32_t *gstart,
                                                               uint32_t W_total, uint32_t log_g, Xyzt *groups,
                                                               FConsts K) {
  const uint32_t total = gstart[W_total];
  for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < total; gi += gridDim.x * blockDim.x) {
    const uint32_t w = upper_index(gstart, W_total, gi);
    const WinDev wd = wins[w];
    const uint32_t lo = (gi - gstart[w]) << log_g;
    uint32_t hi = min(wd.nb, lo + (1u << log_g));
    while (hi > lo && cnt[wd.base + hi - 1] == 0) hi--;
    Xyzt run = ec_identity(K), acc = ec_identity(K);
    for (int b = (int)hi - 1; b >= (int)lo; b--) {
      ec_add(run, buckets[wd.base + b], K);
      ec_add(acc, run, K);
    }
    // bucket b holds digit magnitude b + 1: weight (b + 1) = (b - lo + 1) + lo
    if (lo && hi > lo) {
      const Xyzt t = ec_mul_small(run, lo, K);
      ec_add(acc, t, K);
    }
    groups[gi] = acc;
  }
}

// Workgroup (w, s) of a (windows x S) grid: tree sum of slice s of window
// w's groups into out[w * S + s] (S = 1: the window's sum).  A window of
// 2^23 buckets has ~2^17 groups; with S slices no thread adds more than a
// few hundred of them in sequence.
__global__ void __launch_bounds__(MSM_THREADS) k_window_sum(const Xyzt *groups, const uint32_t *gstart, Xyzt *out,
                                                            FConsts K) {
  extern __shared__ Xyzt sh[];
  const uint32_t w = blockIdx.x, S = gridDim.y, sl = blockIdx.y;
  const uint32_t a = gstart[w], b = gstart[w + 1];
  const uint32_t per = (b - a + S - 1) / S;
  const uint32_t g0 = min(b, a + sl * per), g1 = min(b, g0 + per);
  Xyzt acc = ec_identity(K);
  for (uint32_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) ec_add(acc, groups[g], K);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      Xyzt a = sh[threadIdx.x];
      ec_add(a, sh[threadIdx.x + s], K);
      sh[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(uint64_t)w * S + sl] = sh[0];
}

// out[w] = sum of the S slice sums part[w * S .. w * S + S) (one thread per window).
__global__ void k_window_slices(const Xyzt *part, uint32_t W_total, uint32_t S, Xyzt *out, FConsts K) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W_total; w += gridDim.x * blockDim.x) {
    Xyzt acc = part[(uint64_t)w * S];
    for (uint32_t s = 1; s < S; s++) ec_add(acc, part[(uint64_t)w * S + s], K);
    out[w] = acc;
  }
}

}  // namespace zp

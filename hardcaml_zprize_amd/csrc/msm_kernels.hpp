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
//   k_fill_units     unit -> bucket map (a unit is the part of one bucket
//                    inside one slice of U sorted entries);
//   k_accumulate     one thread per slice: mixed twisted-Edwards additions
//                    of the slice's points into an extended accumulator per
//                    unit (the FPGA bucket pipeline, top.ml /
//                    mixed_add_precompute.ml);
//   k_unit_tree      per bucket: log-depth pairwise sum of its unit partials;
//   k_window_groups  per group of g buckets: running-sum reduction
//                    (driver.cpp:318-339, pippenger.h:25-33) done as
//                    sum_b (b - lo + 1) S_b + lo * sum_b S_b;
//   k_window_sum     one workgroup per window: tree sum of the group results.
// The window combine and the affine conversion follow in k_msm_finalize
// (msm_finalize.hpp; driver.cpp:340-375 does them on the FPGA's host).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ec.hpp"

namespace zp {

static constexpr int MSM_THREADS = 256;
// Entries each thread of the sort kernels handles per loop round (their
// global loads are issued together, before the dependent LDS work).
static constexpr int SORT_ILP = 4;

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

// ---- two-level sorted path (windows of any width up to 2^23 buckets) ----
// Pass A partitions each window's entries by the high part of the bucket,
// group h = b >> lbits (lbits per window, about 2^11 groups), with
// per-(window, chunk) LDS histograms as above; pass B sorts each group by
// the low lbits in a single block and writes the final bucket offsets.
// Both passes stage tiles of entries in LDS and sort them there, so their
// stores come out as runs of consecutive slots (a tile's entries of one
// group, of one bucket).  Group sizes are data dependent, so pass B finds
// its range in the scanned pass-A counters on the device: group g's entries
// are [scanned[i], scanned[i + C]) with i its first counter.

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
  // SORT_ILP digits per thread and round: their loads are all in flight
  // before the first LDS atomic.
  for (uint32_t base = i0 + threadIdx.x; base < i1; base += SORT_ILP * blockDim.x) {
    int32_t d[SORT_ILP];
#pragma unroll
    for (int k = 0; k < SORT_ILP; k++) {
      const uint32_t i = base + k * blockDim.x;
      d[k] = i < i1 ? dig[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < SORT_ILP; k++)
      if (d[k] != 0) lds_claim(hist, ((uint32_t)(d[k] < 0 ? -d[k] : d[k]) - 1) >> lbits);
  }
  __syncthreads();
  uint32_t *dst = counts + jb.cnt0 + (uint64_t)(wd.g0 - jb.g0) * jb.chunks + sg.chunk;
  for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) dst[(uint64_t)h * jb.chunks] = hist[h];
}

// Block-wide exclusive scan of a[0, L) in LDS, in place, for a block of
// 1024 threads (L <= 4096: each thread takes a run of <= 4 words); ws is 16
// words of LDS.  Returns the total.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t *a, uint32_t L, uint32_t *ws) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x, nw = nt >> 6;
  const uint32_t per = (L + nt - 1) / nt;
  const uint32_t l0 = min(L, tid * per), l1 = min(L, l0 + per);
  uint32_t sum = 0;
  for (uint32_t l = l0; l < l1; l++) sum += a[l];
  const uint32_t lane = tid & 63, wv = tid >> 6;
  uint32_t x = sum;  // inclusive scan of the run sums within the wave
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wv] = x;
  __syncthreads();
  if (wv == 0) {
    uint32_t t = lane < nw ? ws[lane] : 0u;
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(t, o);
      if (lane >= o) t += y;
    }
    if (lane < nw) ws[lane] = t;  // inclusive prefix of the wave totals
  }
  __syncthreads();
  const uint32_t total = ws[nw - 1];
  uint32_t run = x - sum + (wv ? ws[wv - 1] : 0u);
  for (uint32_t l = l0; l < l1; l++) {
    const uint32_t c = a[l];
    a[l] = run;
    run += c;
  }
  __syncthreads();
  return total;
}

// Number of indices k in [0, L) with ends[k] <= q (ends ascending), i.e.
// the bin holding slot q of a bin-sorted tile whose bin ends are ends[];
// top = the largest power of two <= L.  Branch-free binary lifting.
__device__ __forceinline__ uint32_t bin_of_slot(const uint32_t *ends, uint32_t L, uint32_t top, uint32_t q) {
  uint32_t k = 0;
  for (uint32_t st = top; st; st >>= 1)
    if (k + st <= L && ends[k + st - 1] <= q) k += st;
  return k;
}

// bin_of_slot restricted to bins [lo, hi] (the caller knows bin(q) lies
// there): ceil(log2(hi - lo + 1)) lifting steps.
__device__ __forceinline__ uint32_t bin_in_range(const uint32_t *ends, uint32_t lo, uint32_t hi, uint32_t q) {
  uint32_t st = 1;
  while (st * 2 <= hi - lo) st <<= 1;
  uint32_t k = lo;
  for (; st; st >>= 1)
    if (k + st <= hi && ends[k + st - 1] <= q) k += st;
  return k;
}

// For the slots base .. base + count - 1 (count > 0) in 64-slot blocks:
// first[b] = the bin of slot base + 64 b, and first[nblk] = the bin of the
// last slot, so that slot base + i lies in a bin of
// [first[i >> 6], first[(i >> 6) + 1]] and the 64 lanes of a wave
// (consecutive slots) read the same two words.
__device__ __forceinline__ void block_first_bins(const uint32_t *ends, uint32_t L, uint32_t top, uint32_t base,
                                                 uint32_t count, uint32_t *first) {
  const uint32_t nblk = (count + 63) >> 6;
  for (uint32_t b = threadIdx.x; b <= nblk; b += blockDim.x)
    first[b] = bin_of_slot(ends, L, top, base + (b < nblk ? b << 6 : count - 1));
}

// Pass A scatter, grouped by (window, group), chunks in order within a
// group, as 4-byte entries tmp[pos] = (i - chunk start) << (lbits + 1) |
// low lbits of the bucket << 1 | sign (pass B recovers the chunk from the
// group's chunk offsets; the host keeps chunk_len << (lbits + 1) within 32
// bits).  A window whose groups are single buckets (lbits = 0) is sorted by
// now: its entries go straight to vals (same slots, point | sign << 31), and
// pass B only writes their offsets.
//
// The chunk is processed in tiles of MSD_A_TILE entries.  Each tile is
// sorted by group in LDS (count, scan, place) and then written out slot by
// slot, so that consecutive lanes store a group's run of the tile to
// consecutive addresses: a few wide writes per run, where one 4-byte store
// per entry into 2^11 places at once left the L2 nothing to merge (every
// entry became its own 32-byte memory write).  The group's cursor in the
// chunk's run advances by the tile's count.
static constexpr int MSD_A_PER = 30;                           // entries per thread
static constexpr uint32_t MSD_A_TILE = (uint32_t)MSD_A_PER * 1024;  // 120 KiB of LDS
// LDS words of k_msd_scatter for at most H <= 4096 groups per window.
__host__ __device__ constexpr size_t msd_scatter_lds_words(uint32_t H) {
  return MSD_A_TILE + 2 * (size_t)H + 16 + MSD_A_TILE / 64 + 1;
}

__global__ void __launch_bounds__(1024) k_msd_scatter(BatchDev B, const SegDev *segs, const int32_t *digits,
                                                      const uint32_t *scanned, uint32_t *tmp, uint32_t *vals) {
  extern __shared__ uint32_t lds[];
  constexpr uint32_t NONE = 0xffffffffu;
  const SegDev sg = segs[blockIdx.x];
  const MsmJobDev jb = B.jobs[sg.job];
  const WinDev wd = B.wins[sg.win];
  const uint32_t lbits = wd.lbits;
  const uint32_t H = bucket_groups(wd.nb, lbits);
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  uint32_t *ent = lds;                     // the tile sorted by group
  uint32_t *cursor = lds + MSD_A_TILE;     // per group: next slot of the chunk's run
  uint32_t *lend = cursor + H;             // per group: tile count -> start -> end
  uint32_t *ws = lend + H;                 // scan scratch
  uint32_t *first = ws + 16;               // first group of each 64-slot block
  const uint32_t *src = scanned + jb.cnt0 + (uint64_t)(wd.g0 - jb.g0) * jb.chunks + sg.chunk;
  for (uint32_t h = tid; h < H; h += nt) cursor[h] = src[(uint64_t)h * jb.chunks];
  uint32_t top = 1;
  while (top * 2 <= H) top <<= 1;
  const uint32_t i0 = sg.chunk * jb.chunk_len, i1 = min(jb.n, i0 + jb.chunk_len);
  const int32_t *dig = digits + jb.dig0 + (uint64_t)(sg.win - jb.win0) * jb.n;
  for (uint32_t t0 = i0; t0 < i1; t0 += MSD_A_TILE) {
    for (uint32_t h = tid; h < H; h += nt) lend[h] = 0;
    __syncthreads();
    uint32_t key[MSD_A_PER], e[MSD_A_PER];
#pragma unroll
    for (int k = 0; k < MSD_A_PER; k++) {
      const uint32_t i = t0 + tid + k * nt;
      const int32_t d = i < i1 ? dig[i] : 0;
      key[k] = NONE;
      e[k] = 0;
      if (d != 0) {
        const uint32_t b = (uint32_t)(d < 0 ? -d : d) - 1;
        const uint32_t sign = d < 0 ? 1u : 0u;
        key[k] = b >> lbits;
        e[k] = lbits == 0 ? i | sign << 31 : (i - i0) << (lbits + 1) | (b & ((1u << lbits) - 1)) << 1 | sign;
      }
    }
#pragma unroll
    for (int k = 0; k < MSD_A_PER; k++)
      if (key[k] != NONE) lds_claim(lend, key[k]);
    __syncthreads();
    const uint32_t tot = block_excl_scan(lend, H, ws);  // lend: each group's first slot in the tile
#pragma unroll
    for (int k = 0; k < MSD_A_PER; k++)
      if (key[k] != NONE) ent[lds_claim(lend, key[k])] = e[k];  // lend ends as each group's end
    __syncthreads();
    if (tot) block_first_bins(lend, H, top, 0, tot, first);
    __syncthreads();
    uint32_t *dst = lbits == 0 ? vals : tmp;
    for (uint32_t q = tid; q < tot; q += nt) {
      const uint32_t h = bin_in_range(lend, first[q >> 6], first[(q >> 6) + 1], q);
      dst[cursor[h] + q - (h ? lend[h - 1] : 0u)] = ent[q];
    }
    __syncthreads();
    for (uint32_t h = tid; h < H; h += nt) cursor[h] += lend[h] - (h ? lend[h - 1] : 0u);
    __syncthreads();
  }
}

// Pass B: block g sorts bucket group g (window w, high part h) by the low
// lbits of its window.  A histogram over the whole group gives the bucket
// offsets and counts; then the group is placed tile by tile (tile_n entries,
// at most MSD_B_PER per thread): each tile is sorted by bucket in LDS and
// written out slot by slot, so a bucket's run of the tile is stored by
// consecutive lanes (vals = point | sign << 31, the point being its chunk's
// first point plus the offset pass A stored).  The last group also writes
// off[K] (the total).  LDS words: msd_group_fixed_words + tile_n.
static constexpr int MSD_B_PER = 32;
static constexpr size_t MSD_LDS_WORDS = 160 * 1024 / 4;  // LDS of one gfx950 workgroup
__host__ __device__ constexpr size_t msd_group_fixed_words(uint32_t max_lbits, uint32_t max_chunks) {
  return (2u << max_lbits) + 16 + 2 * (MSD_B_PER * 1024 / 64 + 1) + max_chunks + 1;
}
__global__ void __launch_bounds__(1024) k_msd_group(BatchDev B, const uint32_t *win_g0, uint32_t Wt, uint32_t GT,
                                                    uint32_t max_lbits, uint32_t max_chunks, uint32_t tile_n,
                                                    const uint32_t *scanned, const uint32_t *tmp, uint32_t *vals,
                                                    uint32_t *off, uint32_t *cnt) {
  extern __shared__ uint32_t lds[];
  constexpr uint32_t NONE = 0xffffffffu;
  const uint32_t g = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const uint32_t w = upper_index(win_g0, Wt, g);
  const WinDev wd = B.wins[w];
  const MsmJobDev jb = B.jobs[wd.job];
  const uint64_t ci = jb.cnt0 + (uint64_t)(g - jb.g0) * jb.chunks;
  const uint32_t start = scanned[ci], end = scanned[ci + jb.chunks];
  const uint32_t b0 = (g - wd.g0) << wd.lbits;
  const uint32_t nl = min(1u << wd.lbits, wd.nb - b0);  // buckets in this group
  if (wd.lbits == 0) {  // one bucket, placed by pass A
    if (tid == 0) {
      off[wd.base + b0] = start;
      cnt[wd.base + b0] = end - start;
      if (g == GT - 1) off[B.K] = end;
    }
    return;
  }
  uint32_t *cur = lds, *lend = cur + (1u << max_lbits), *ws = lend + (1u << max_lbits);
  uint32_t *cfirst = ws + 16;                               // chunk of each 64-position block of the tile
  uint32_t *bfirst = cfirst + (MSD_B_PER * 1024 / 64 + 1);  // bucket of each 64-slot block
  uint32_t *cs = bfirst + (MSD_B_PER * 1024 / 64 + 1);      // chunk c's entries are [cs[c], cs[c + 1])
  uint32_t *ent = cs + max_chunks + 1;
  const uint32_t lmask = (1u << wd.lbits) - 1;
  for (uint32_t l = tid; l < nl; l += nt) lend[l] = 0;
  for (uint32_t c = tid; c <= jb.chunks; c += nt) cs[c] = scanned[ci + c];
  __syncthreads();
  for (uint32_t base = start + tid; base < end; base += SORT_ILP * nt) {
    uint32_t key[SORT_ILP];
#pragma unroll
    for (int k = 0; k < SORT_ILP; k++) {
      const uint32_t p = base + k * nt;
      key[k] = p < end ? (tmp[p] >> 1) & lmask : NONE;
    }
#pragma unroll
    for (int k = 0; k < SORT_ILP; k++)
      if (key[k] != NONE) lds_claim(lend, key[k]);
  }
  __syncthreads();
  for (uint32_t l = tid; l < nl; l += nt) cnt[wd.base + b0 + l] = lend[l];
  block_excl_scan(lend, nl, ws);
  for (uint32_t l = tid; l < nl; l += nt) {
    off[wd.base + b0 + l] = start + lend[l];
    cur[l] = start + lend[l];
  }
  if (g == GT - 1 && tid == 0) off[B.K] = end;
  uint32_t top = 1;
  while (top * 2 <= nl) top <<= 1;
  uint32_t ctop = 1;  // the largest power of two <= C - 1
  while (ctop * 2 < jb.chunks) ctop <<= 1;
  const uint32_t per = tile_n / nt;
  for (uint32_t t0 = start; t0 < end; t0 += tile_n) {
    const uint32_t t1 = min(end, t0 + tile_n);
    __syncthreads();  // cur (first tile) / the previous tile's reads of lend, bfirst and ent
    for (uint32_t l = tid; l < nl; l += nt) lend[l] = 0;
    // p's chunk = the number of chunk starts cs[1 .. C - 1] at or before p (cs[0] = start <= p)
    if (jb.chunks > 1) block_first_bins(cs + 1, jb.chunks - 1, ctop, t0, t1 - t0, cfirst);
    __syncthreads();
    uint32_t key[MSD_B_PER], val[MSD_B_PER];
#pragma unroll
    for (int k = 0; k < MSD_B_PER; k++) {
      const uint32_t p = t0 + tid + k * nt;
      key[k] = NONE;
      val[k] = 0;
      if (k < (int)per && p < t1) {
        const uint32_t e = tmp[p];
        key[k] = (e >> 1) & lmask;
        uint32_t c = 0;
        if (jb.chunks > 1) {
          const uint32_t bk = (p - t0) >> 6;
          c = bin_in_range(cs + 1, cfirst[bk], cfirst[bk + 1], p);
        }
        val[k] = (c * jb.chunk_len + (e >> (wd.lbits + 1))) | (e & 1u) << 31;
      }
    }
#pragma unroll
    for (int k = 0; k < MSD_B_PER; k++)
      if (key[k] != NONE) lds_claim(lend, key[k]);
    __syncthreads();
    const uint32_t tot = block_excl_scan(lend, nl, ws);
#pragma unroll
    for (int k = 0; k < MSD_B_PER; k++)
      if (key[k] != NONE) ent[lds_claim(lend, key[k])] = val[k];
    __syncthreads();
    if (tot) block_first_bins(lend, nl, top, 0, tot, bfirst);
    __syncthreads();
    for (uint32_t q = tid; q < tot; q += nt) {
      const uint32_t l = bin_in_range(lend, bfirst[q >> 6], bfirst[(q >> 6) + 1], q);
      vals[cur[l] + q - (l ? lend[l - 1] : 0u)] = ent[q];
    }
    __syncthreads();
    for (uint32_t l = tid; l < nl; l += nt) cur[l] += lend[l] - (l ? lend[l - 1] : 0u);
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

// Work units: the sorted entries are cut into slices of U = 2^log_u (one
// thread each, k_accumulate), and a unit is the part of one bucket inside
// one slice.  nunits[k] = the number of slices bucket k's entries
// [off[k], off[k] + cnt[k]) touch (0 when empty); k == K gets 0 (so the
// scan's last entry is the total).  *umax (zeroed by the caller) becomes the
// largest unit count of any bucket when that exceeds 1 (a wave maximum,
// then one atomic per wave).
__global__ void k_units_per_bucket(const uint32_t *off, const uint32_t *cnt, uint32_t K, uint32_t log_u,
                                   uint32_t *nunits, uint32_t *umax) {
  for (uint32_t base = blockIdx.x * blockDim.x; base <= K; base += gridDim.x * blockDim.x) {  // block-uniform
    const uint32_t k = base + threadIdx.x;
    uint32_t v = 0;
    if (k <= K) {
      v = k < K && cnt[k] ? ((off[k] + cnt[k] - 1) >> log_u) - (off[k] >> log_u) + 1 : 0;
      nunits[k] = v;
    }
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    // most buckets touch one or two slices: read first, so that the waves
    // whose maximum is already there skip the atomic (a value read late is
    // only smaller, never larger, than the current one)
    if ((threadIdx.x & 63) == 0 && v > 1 && v > __hip_atomic_load(umax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(umax, v);
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

// One thread per slice of U = 2^log_u consecutive sorted entries (the last
// one shorter), so every lane of a wave makes the same number of mixed
// additions whatever the bucket sizes (with one thread per bucket, lanes
// idle while the wave's largest bucket finishes).  A slice's entries belong
// to one or more consecutive buckets; each (bucket, slice) piece is a work
// unit whose partial goes to partial[u], u = ustart[k] + the slice's rank
// among the slices bucket k touches, and k_unit_tree adds a bucket's
// pieces into its first unit.  The bucket changes inside the one flat loop (a short
// divergent branch), so lanes never wait for each other's piece lengths.
__global__ void __launch_bounds__(MSM_THREADS) k_accumulate(BatchDev B, const uint32_t *vals, const uint32_t *off,
                                                            const uint32_t *ustart, uint32_t K, uint32_t log_u,
                                                            Xyzt *partial, FConsts KC) {
  const uint32_t total = off[K];
  const uint32_t slices = (total + (1u << log_u) - 1) >> log_u;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < slices; t += gridDim.x * blockDim.x) {
    const uint32_t e0 = t << log_u, e1 = min(total, e0 + (1u << log_u));
    uint32_t k = upper_index(off, K, e0);  // the non-empty bucket holding e0 (empty ones share its offset)
    uint32_t u = ustart[k] + (t - (off[k] >> log_u));
    uint32_t seg_end = min(off[k + 1], e1);
    const PackedPre *pts = B.jobs[upper_index(B.job_b0, B.J, k)].pts;
    Xyzt acc = ec_identity(KC);
    for (uint32_t e = e0; e < e1; e++) {
      if (e == seg_end) {  // bucket k ends inside the slice: store its piece, go to the next bucket
        partial[u] = acc;
        acc = ec_identity(KC);
        while (off[k + 1] <= e) k++;  // off[K] = total > e ends the walk
        u = ustart[k];
        seg_end = min(off[k + 1], e1);
        pts = B.jobs[upper_index(B.job_b0, B.J, k)].pts;
      }
      const uint32_t v = vals[e];
      const PrePt p = pre_unpack(pts[v & 0x7fffffffu]);
      ec_madd(acc, p, (v >> 31) != 0, KC);
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
__global__ void __launch_bounds__(MSM_THREADS) k_window_groups(const Xyzt *partial, const uint32_t *ustart,
                                                               const uint32_t *cnt, const WinDev *wins,
                                                               const uint32_t *gstart, uint32_t W_total,
                                                               uint32_t log_g, Xyzt *groups, FConsts K) {
  const uint32_t total = gstart[W_total];
  for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < total; gi += gridDim.x * blockDim.x) {
    const uint32_t w = upper_index(gstart, W_total, gi);
    const WinDev wd = wins[w];
    const uint32_t lo = (gi - gstart[w]) << log_g;
    uint32_t hi = min(wd.nb, lo + (1u << log_g));
    while (hi > lo && cnt[wd.base + hi - 1] == 0) hi--;
    Xyzt run = ec_identity(K), acc = ec_identity(K);
    for (int b = (int)hi - 1; b >= (int)lo; b--) {
      // a non-empty bucket's sum is its first unit's partial after the tree
      // levels; an empty one adds nothing to the running sum
      if (cnt[wd.base + b]) ec_add(run, partial[ustart[wd.base + b]], K);
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

// sortprim.hip -- record sort (bitonic tile sort in LDS + merge-path merge passes),
// multi-task merge-path merges, and the parallel one-safe-event-per-LP selection.
#include "sortprim.h"

namespace gpq {
namespace {

constexpr int kThreads = 256;                       // selection blocks; sort and merge blocks: kT of
                                                    // 128, 256 or 512 threads (tiles of 8 kT records)
constexpr int kPerThread = 8;                       // records per thread in the tile sort and merge
constexpr uint32_t kSelectTile = 1024;              // records per selection block (256 x 4)
constexpr uint32_t kSplitLanes = 8;                 // lanes per tile boundary in the merge partition

__device__ __forceinline__ Event sentinel() { return Event{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}; }

// A = a1[0:n1) ++ a2[...]
struct ASrc {
  const Event* a1;
  const Event* a2;
  uint32_t n1;
  __device__ __forceinline__ Event get(uint32_t i) const { return i < n1 ? load_event(a1 + i) : load_event(a2 + (i - n1)); }
};

struct OutMap {
  Event* out[3];
  uint32_t end[3];
  __device__ __forceinline__ Event* at(uint32_t o) const {
    if (o < end[0]) return out[0] + o;
    if (o < end[1]) return out[1] + (o - end[0]);
    return out[2] + (o - end[1]);
  }
};

// Merge-path split: the number of A records among the first d outputs of the stable merge
// (ties: A first) is the smallest i in [max(0, d - nb), min(d, na)] with B[d-1-i] < A[i].
// Found for the block's two diagonals d0 and d1 at once, each by half of the block: every
// round the 128 threads of a half probe 128 evenly spaced points of the remaining range and
// keep the gap where the predicate turns true, so a range of n shrinks to nothing in
// ceil(log_128 n) rounds of one load each, not log_2 n dependent loads.
template <int kT>
__device__ __forceinline__ void block_merge_splits(const ASrc& A, uint32_t na, const Event* B, uint32_t nb, uint32_t d0, uint32_t d1,
                                   uint32_t* s_split) {
  constexpr uint32_t kProbes = kT / 2;
  __shared__ uint32_t s_lo[2], s_hi[2], s_first[2];
  const uint32_t h = threadIdx.x / kProbes, k = threadIdx.x % kProbes;
  const uint32_t d = h ? d1 : d0;
  if (k == 0) {
    s_lo[h] = d > nb ? d - nb : 0u;
    s_hi[h] = d < na ? d : na;
  }
  __syncthreads();
  while (s_lo[0] < s_hi[0] || s_lo[1] < s_hi[1]) {   // uniform: read after a barrier
    const uint32_t lo = s_lo[h], hi = s_hi[h];
    const uint32_t step = (hi - lo + kProbes - 1) / kProbes;
    if (k == 0) s_first[h] = 0xFFFFFFFFu;
    __syncthreads();
    if (lo < hi) {
      const uint32_t i = lo + k * step;
      if (i < hi && key_less(load_event(B + (d - 1u - i)), A.get(i))) atomicMin(&s_first[h], k);
    }
    __syncthreads();
    if (k == 0 && lo < hi) {
      const uint32_t f = s_first[h];
      if (f == 0xFFFFFFFFu) {
        s_lo[h] = lo + (hi - lo - 1u) / step * step + 1u;   // past the last probe
      } else {
        s_hi[h] = lo + f * step;
        if (f) s_lo[h] = lo + (f - 1u) * step + 1u;
      }
    }
    __syncthreads();
  }
  if (k == 0) s_split[h] = s_lo[h];
  __syncthreads();
}

// Merges output ranks [t0, t0 + kMergeTile) of merge(A, B) through LDS: the block's input
// ranges are located by two merge-path searches, staged in LDS, each thread merges 8
// outputs, and the results leave through LDS in coalesced 16-byte stores.
// LDS slot of tile element i: one pad record after every 8, so the 8-record runs that
// consecutive threads merge start in different banks.
__device__ __forceinline__ uint32_t pad8(uint32_t i) { return i + (i >> 3); }

// The same split for diagonal d by a group of P lanes (P a power of two dividing the wave; every
// lane of a group passes the same d, `active` false for a group with no diagonal): P probes per
// round, ceil(log_P n) rounds.  The loop runs while any group of the wave is searching.
template <int P>
__device__ __forceinline__ uint32_t group_merge_split(const ASrc& A, uint32_t na, const Event* B, uint32_t nb,
                                                      uint32_t d, bool active) {
  const uint32_t lane = (uint32_t)lane_id(), k = lane % P, gbase = lane - k;
  const uint64_t gmask = (P == kWave ? ~0ull : ((1ull << P) - 1ull)) << gbase;
  uint32_t lo = active && d > nb ? d - nb : 0u, hi = active ? (d < na ? d : na) : 0u;
  while (__ballot(lo < hi)) {
    const bool run = lo < hi;                         // uniform within the group
    const uint32_t step = run ? (hi - lo + P - 1u) / P : 1u;
    const uint32_t i = lo + k * step;
    const bool p = run && i < hi && key_less(load_event(B + (d - 1u - i)), A.get(i));
    const uint64_t m = (__ballot(p) & gmask) >> gbase;
    if (run) {
      if (m == 0) {
        lo = lo + (hi - lo - 1u) / step * step + 1u;  // past the last probe
      } else {
        const uint32_t f = (uint32_t)__ffsll((unsigned long long)m) - 1u;
        hi = lo + f * step;
        if (f) lo = lo + (f - 1u) * step + 1u;
      }
    }
  }
  return lo;
}

// With `pre` set, the block's splits were computed before the launch (merge_partition_kernel):
// pre[0] and pre[1] are those of its first and last output rank.
template <int kT>
__device__ __forceinline__ void merge_tile(const ASrc& A, uint32_t na, const Event* B, uint32_t nb, uint32_t t0,
                                           const OutMap& out, const uint32_t* pre = nullptr) {
  constexpr uint32_t kTile = kT * kPerThread;
  __shared__ Event s[kTile + kTile / 8];
  __shared__ uint32_t s_split[2];
  const uint32_t total = na + nb;
  const uint32_t t1 = t0 + kTile < total ? t0 + kTile : total;
  uint32_t i0, i1;
  if (pre) {
    i0 = pre[0];
    i1 = pre[1];
  } else {
    block_merge_splits<kT>(A, na, B, nb, t0, t1, s_split);
    i0 = s_split[0];
    i1 = s_split[1];
  }
  const uint32_t j0 = t0 - i0;
  const uint32_t ca = i1 - i0, cb = (t1 - t0) - ca;
  for (uint32_t x = threadIdx.x; x < ca; x += kT) s[pad8(x)] = A.get(i0 + x);
  for (uint32_t x = threadIdx.x; x < cb; x += kT) s[pad8(ca + x)] = load_event(B + j0 + x);
  __syncthreads();
  const uint32_t len = t1 - t0;
  const uint32_t d = threadIdx.x * kPerThread;
  uint4 res[kPerThread];
  uint32_t got = 0;
  if (d < len) {
    uint32_t lo = d > cb ? d - cb : 0u, hi = d < ca ? d : ca;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (key_less(s[pad8(ca + (d - 1u - mid))], s[pad8(mid)])) hi = mid; else lo = mid + 1u;
    }
    uint32_t ia = lo, ib = d - lo;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      if (d + q < len) {
        const bool take_a = ib >= cb || (ia < ca && !key_less(s[pad8(ca + ib)], s[pad8(ia)]));
        const Event r = take_a ? s[pad8(ia)] : s[pad8(ca + ib)];
        res[q] = make_uint4(r.ts, r.dst, r.src, r.seq);
        ia += take_a ? 1u : 0u;
        ib += take_a ? 0u : 1u;
        ++got;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPerThread; ++q)
    if ((uint32_t)q < got) s[pad8(d + q)] = Event{res[q].x, res[q].y, res[q].z, res[q].w};
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < len; x += kT) store_event(out.at(t0 + x), s[pad8(x)]);
}

__device__ __forceinline__ void compare_swap(Event& a, Event& b, bool up) {
  if (up ? key_less(b, a) : key_less(a, b)) {
    const Event t = a;
    a = b;
    b = t;
  }
}

// Bitonic sort of one kSortTile-record tile.  Thread t holds records 8t .. 8t+7 in registers,
// so every stage with partner distance j < 8 (30 of the 66) runs in registers with no barrier;
// the stages with j >= 8 run in LDS, each thread exchanging 4 of the tile's 1024 pairs.  The
// LDS tile is padded one record per 8 (pad8) so that the 8-record runs of neighbouring threads
// start in different banks.
// seg_pre (optional): `in` is a striped insert area (core.cuh) whose segment k holds records
// [seg_pre[k], seg_pre[k + 1]) of the run (k < kInsertStripes: at k * seg_cap; the spill
// segment k = kInsertStripes after the stripes).
template <int kT>
__global__ __launch_bounds__(kT) void tile_sort_kernel(const Event* in, Event* out, const uint32_t* d_n,
                                                       const uint32_t* seg_pre, uint32_t seg_cap) {
  constexpr uint32_t kTile = kT * kPerThread;
  __shared__ Event s[kTile + kTile / 8];
  __shared__ uint32_t s_pre[kInsertStripes + 2];
  const uint32_t n = *d_n;
  const uint32_t base = blockIdx.x * kTile;
  if (base >= n) return;
  if (seg_pre) {
    for (uint32_t k = threadIdx.x; k < kInsertStripes + 2u; k += kT) s_pre[k] = seg_pre[k];
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < kTile; i += kT) {
    const uint32_t g = base + i;
    if (g >= n) {
      s[pad8(i)] = sentinel();
    } else if (!seg_pre) {
      s[pad8(i)] = load_event(in + g);
    } else {
      uint32_t lo = 0, hi = kInsertStripes;          // the last segment k with s_pre[k] <= g
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (s_pre[mid] <= g) lo = mid; else hi = mid - 1u;
      }
      s[pad8(i)] = load_event(in + (size_t)lo * seg_cap + (g - s_pre[lo]));
    }
  }
  __syncthreads();
  const uint32_t t0 = threadIdx.x * 8u;
  Event r[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) r[q] = s[pad8(t0 + q)];
#pragma unroll
  for (uint32_t k = 2; k <= 8; k <<= 1)
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (uint32_t a = 0; a < 8; ++a)
        if ((a & j) == 0) compare_swap(r[a], r[a | j], ((t0 + a) & k) == 0);
  for (uint32_t k = 16; k <= kTile; k <<= 1) {
#pragma unroll
    for (int q = 0; q < 8; ++q) s[pad8(t0 + q)] = r[q];
    __syncthreads();
    for (uint32_t j = k >> 1; j >= 8; j >>= 1) {
#pragma unroll
      for (uint32_t q = 0; q < kTile / 2 / kT; ++q) {
        const uint32_t pr = threadIdx.x + q * kT;
        const uint32_t i = ((pr & ~(j - 1u)) << 1) | (pr & (j - 1u));   // pair (i, i + j), bit j of i clear
        Event x = s[pad8(i)], y = s[pad8(i | j)];
        compare_swap(x, y, (i & k) == 0);
        s[pad8(i)] = x;
        s[pad8(i | j)] = y;
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = s[pad8(t0 + q)];
    const bool up = (t0 & k) == 0;   // bit k >= 16 of i is that of t0 for all eight
#pragma unroll
    for (uint32_t j = 4; j > 0; j >>= 1)
#pragma unroll
      for (uint32_t a = 0; a < 8; ++a)
        if ((a & j) == 0) compare_swap(r[a], r[a | j], up);
  }
  __syncthreads();   // every thread's last LDS stage reads are done
#pragma unroll
  for (int q = 0; q < 8; ++q) s[pad8(t0 + q)] = r[q];
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kTile; i += kT)
    if (base + i < n) store_event(out + base + i, s[pad8(i)]);
}

// Merges adjacent sorted runs of width w (2w is a multiple of kMergeTile); pass `pass` of
// `passes`.  Once w >= n every remaining pass would only copy the one run, so those passes
// do nothing when their number is even (the run already sits in the final buffer) and the
// first of them copies when it is odd.
template <int kT>
__global__ __launch_bounds__(kT) void merge_pass_kernel(const Event* in, Event* out, const uint32_t* d_n,
                                                        uint32_t w, uint32_t pass, uint32_t passes) {
  constexpr uint32_t kTile = kT * kPerThread;
  const uint32_t n = *d_n;
  const uint32_t t0 = blockIdx.x * kTile;
  if (t0 >= n) return;
  if (w >= n) {
    uint32_t first = 0;   // the first pass whose width covers n
    while (((uint64_t)kTile << first) < n) ++first;
    if (((passes - first) & 1u) == 0 || pass != first) return;
  }
  const uint32_t base = t0 / (2u * w) * (2u * w);
  const uint32_t na = n - base < w ? n - base : w;
  const uint32_t rest = n - base - na;
  const uint32_t nb = rest < w ? rest : w;
  const ASrc A{in + base, in + base, na};
  OutMap om{{out + base, out + base, out + base}, {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};
  merge_tile<kT>(A, na, in + base + na, nb, t0 - base, om);
}

// splits (optional): task y's splits at splits[y * stride + x], for tile x's first output rank
// (x <= the task's tile count; the last one is that of the task's total).
template <int kT>
__global__ __launch_bounds__(kT) void merge_tasks_kernel(const MergeTask* tasks, const uint32_t* splits,
                                                         uint32_t stride) {
  constexpr uint32_t kTile = kT * kPerThread;
  const MergeTask t = tasks[blockIdx.y];
  const uint32_t na = t.n_a1 + t.n_a2;
  const uint32_t total = na + t.n_b;
  const uint32_t t0 = blockIdx.x * kTile;
  if (t0 >= total) return;
  const ASrc A{t.a1, t.a2, t.n_a1};
  const OutMap om{{t.out[0], t.out[1], t.out[2]}, {t.end[0], t.end[1], t.end[2]}};
  merge_tile<kT>(A, na, t.b, t.n_b, t0, om, splits ? splits + (size_t)blockIdx.y * stride + blockIdx.x : nullptr);
}

// The merge-path splits of every tile boundary of every task, kSplitLanes lanes per boundary,
// so that the merge blocks start loading at once.  Boundary x of task y is output rank
// min(x * kTile, total); grid (ceil(boundaries / (256 / kSplitLanes)), ntasks) of 256 threads.
// Fewer lanes per boundary take more rounds but far fewer loads, which is what limits this
// launch: thousands of boundaries search at once and their probes are scattered lines.
template <int kT>
__global__ __launch_bounds__(256) void merge_partition_kernel(const MergeTask* tasks, uint32_t boundaries,
                                                              uint32_t* splits, uint32_t stride) {
  constexpr uint32_t kTile = kT * kPerThread;
  constexpr uint32_t kPer = 256u / kSplitLanes;         // boundaries per block
  const uint32_t x = blockIdx.x * kPer + threadIdx.x / kSplitLanes;
  if (blockIdx.x * kPer >= boundaries) return;          // the whole block is past the boundaries
  const MergeTask t = tasks[blockIdx.y];
  const uint32_t na = t.n_a1 + t.n_a2;
  const uint32_t total = na + t.n_b;
  // boundaries past the task's last tile boundary do not search (their groups stay in the loop)
  const bool active = x < boundaries && (uint64_t)x * kTile <= total + (uint64_t)kTile;
  const uint32_t d = (uint64_t)x * kTile < total ? x * kTile : total;
  const ASrc A{t.a1, t.a2, t.n_a1};
  const uint32_t s = group_merge_split<kSplitLanes>(A, na, t.b, t.n_b, d, active);
  if (active && threadIdx.x % kSplitLanes == 0) splits[(size_t)blockIdx.y * stride + x] = s;
}

// One wave: lane k < kInsertStripes + 1 takes segment k's record count (a stripe counter past
// seg_cap also counts the lanes that spilled), the segments' offsets are a wave prefix sum,
// and every offset is capped at max_total.
__global__ __launch_bounds__(64) void gather_inserts_kernel(uint32_t* seg_m, uint32_t seg_cap, uint32_t spill_cap,
                                                            uint32_t max_total, uint32_t* seg_pre, uint32_t* d_m,
                                                            WindowState* ws) {
  const uint32_t k = threadIdx.x;
  uint32_t n = 0;
  if (k <= kInsertStripes) {
    const uint32_t c = seg_m[k * kStripeWords], cap = k < kInsertStripes ? seg_cap : spill_cap;
    n = c < cap ? c : cap;
    seg_m[k * kStripeWords] = 0;
  }
  uint32_t inc = n;
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)inc, off, kWave);
    if (k >= (uint32_t)off) inc += t;
  }
  const uint32_t exc = inc - n;
  const uint32_t total = (uint32_t)__shfl((int)inc, (int)kInsertStripes, kWave);
  const uint32_t m = total < max_total ? total : max_total;
  if (k <= kInsertStripes) seg_pre[k] = exc < max_total ? exc : max_total;
  if (k == 0) {
    seg_pre[kInsertStripes + 1] = m;
    *d_m = m;
    if (total > max_total) atomicOr(&ws->error, kErrQueueOverflow);
  }
}

__global__ void copy_records_kernel(const Event* src, Event* dst, const uint32_t* d_n, uint32_t cap) {
  const uint32_t n = *d_n < cap ? *d_n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    store_event(dst + i, load_event(src + i));
}

// --- selection --------------------------------------------------------------------

__device__ __forceinline__ bool in_prefix(const Event& e, uint32_t wend, bool bounded, const Event& bound) {
  return e.ts < wend && (!bounded || !key_less(bound, e));
}

__global__ void select_prefix_kernel(const Event* const* d_s, const uint32_t* d_n, const uint32_t* d_wend, const Event* d_bound,
                                     const uint32_t* d_bound_valid, uint32_t* d_prefix) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Event* s = *d_s;
  const uint32_t n = *d_n, wend = *d_wend;
  const bool bounded = d_bound_valid && *d_bound_valid;
  const Event bound = bounded ? *d_bound : sentinel();
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (in_prefix(load_event(s + mid), wend, bounded, bound)) lo = mid + 1u; else hi = mid;
  }
  *d_prefix = lo;
}

__global__ void select_mark_kernel(const Event* const* d_s, const uint32_t* d_prefix, uint32_t* first) {
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p; i += gridDim.x * blockDim.x)
    atomicMin(&first[s[i].dst], i);
}

__global__ __launch_bounds__(kThreads) void select_count_kernel(const Event* const* d_s, const uint32_t* d_prefix,
                                                                const uint32_t* first, uint32_t* counts) {
  __shared__ uint32_t s_w[kThreads / kWave];
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  const uint32_t base = blockIdx.x * kSelectTile;
  uint32_t c = 0;
  if (base < p) {
    for (uint32_t i = base + threadIdx.x; i < base + kSelectTile && i < p; i += kThreads)
      c += first[s[i].dst] == i ? 1u : 0u;
  }
  c = wave_sum_u32(c);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Exclusive scan of counts[0:nb) into offsets[0:nb), nb = the blocks that hold the scanned
// prefix (the count and scatter kernels do nothing past it); totals of marked / rest.  The
// counts are read in sweeps of kScanRounds x 1024: element r * 1024 + t of a sweep goes to
// thread t (coalesced, all loads of a sweep in flight together); each (round, wave) pair's
// total goes to LDS and one wave scans those totals in order.
constexpr int kScanRounds = 8;
__global__ __launch_bounds__(1024) void select_scan_kernel(const uint32_t* counts, uint32_t* offsets, uint32_t nblk,
                                                           const uint32_t* d_prefix, uint32_t* d_deq_n,
                                                           uint32_t* d_rest_n) {
  constexpr uint32_t kWaves = 1024 / kWave;
  constexpr uint32_t kParts = kScanRounds * kWaves;          // (round, wave) totals per sweep
  static_assert(kParts % kWave == 0, "whole parts per lane");
  __shared__ uint32_t s_part[kParts];
  __shared__ uint32_t s_total;
  const uint32_t p = *d_prefix;
  const uint32_t want = (uint32_t)(((uint64_t)p + kSelectTile - 1) / kSelectTile);
  const uint32_t nb = want < nblk ? want : nblk;
  const uint32_t lane = (uint32_t)lane_id(), wv = threadIdx.x / kWave;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kScanRounds * 1024u) {
    uint32_t v[kScanRounds];
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
      const uint32_t i = base + (uint32_t)r * 1024u + threadIdx.x;
      v[r] = i < nb ? counts[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {   // v[r] becomes the exclusive prefix inside the wave
      uint32_t x = v[r];
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
        if ((int)lane >= off) x += y;
      }
      if (lane == kWave - 1) s_part[r * kWaves + wv] = x;
      v[r] = x - v[r];
    }
    __syncthreads();
    if (wv == 0) {   // exclusive scan of the parts in (round, wave) order, kParts / 64 per lane
      constexpr uint32_t kPer = kParts / kWave;
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t u = 0; u < kPer; ++u) sum += s_part[lane * kPer + u];
      uint32_t x = sum;
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
        if ((int)lane >= off) x += y;
      }
      uint32_t run = x - sum;
#pragma unroll
      for (uint32_t u = 0; u < kPer; ++u) {
        const uint32_t c = s_part[lane * kPer + u];
        s_part[lane * kPer + u] = run;
        run += c;
      }
      if (lane == kWave - 1) s_total = x;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
      const uint32_t i = base + (uint32_t)r * 1024u + threadIdx.x;
      if (i < nb) offsets[i] = carry + s_part[r * kWaves + wv] + v[r];
    }
    carry += s_total;
    __syncthreads();   // s_part and s_total are rewritten by the next sweep
  }
  if (threadIdx.x == 0) {
    *d_deq_n = carry;
    *d_rest_n = p - carry;
  }
}

__global__ __launch_bounds__(kThreads) void select_scatter_kernel(const Event* const* d_s, const uint32_t* d_prefix,
                                                                  uint32_t* first, const uint32_t* offsets,
                                                                  Event* deq, Event* rest) {
  __shared__ uint32_t s_w[kThreads / kWave];
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  const uint32_t base = blockIdx.x * kSelectTile;
  if (base >= p) return;
  constexpr uint32_t kItems = kSelectTile / kThreads;   // 4 consecutive records per thread
  const uint32_t i0 = base + threadIdx.x * kItems;
  bool mark[kItems];
  Event e[kItems];
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t q = 0; q < kItems; ++q) {
    const uint32_t i = i0 + q;
    mark[q] = false;
    if (i < p) {
      e[q] = load_event(s + i);
      mark[q] = first[e[q].dst] == i;
      mine += mark[q] ? 1u : 0u;
    }
  }
  uint32_t x = mine;   // inclusive scan over the block, in thread order
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
    if (lane_id() >= off) x += y;
  }
  if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = x;
  __syncthreads();
  uint32_t wave_off = 0;
  for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) wave_off += s_w[w];
  uint32_t before = offsets[blockIdx.x] + wave_off + x - mine;   // marked records before i0
#pragma unroll
  for (uint32_t q = 0; q < kItems; ++q) {
    const uint32_t i = i0 + q;
    if (i < p) {
      if (mark[q]) {
        store_event(deq + before, e[q]);
        first[e[q].dst] = 0xFFFFFFFFu;
        ++before;
      } else {
        store_event(rest + (i - before), e[q]);
      }
    }
  }
}

uint32_t grid_of(uint64_t items, uint32_t per_block) {
  const uint64_t g = (items + per_block - 1) / per_block;
  return g ? (uint32_t)g : 1u;
}

// Work accounting (queues created with collect_stats): one small launch after a primitive
// adds the records it handled to a counter, so the primitive's own kernels are unchanged.
__global__ void acct_count_kernel(const uint32_t* d_n, uint32_t cap, unsigned long long* acct) {
  if (threadIdx.x == 0) atomicAdd(acct, (unsigned long long)min(*d_n, cap));
}

__global__ __launch_bounds__(256) void acct_tasks_kernel(const MergeTask* tasks, uint32_t ntasks,
                                                         unsigned long long* acct) {
  unsigned long long s = 0;
  for (uint32_t t = threadIdx.x; t < ntasks; t += blockDim.x)
    s += (unsigned long long)tasks[t].n_a1 + tasks[t].n_a2 + tasks[t].n_b;
  if (s) atomicAdd(acct, s);
}

}  // namespace

namespace {
template <int kT>
gpq_status sort_records_t(const Event* in, Event* buf0, Event* buf1, const uint32_t* d_n, uint32_t cap,
                          hipStream_t stream, Event** sorted, const uint32_t* seg_pre, uint32_t seg_cap) {
  constexpr uint32_t kTile = kT * kPerThread;
  hipLaunchKernelGGL(tile_sort_kernel<kT>, dim3(grid_of(cap, kTile)), dim3(kT), 0, stream, in, buf0, d_n, seg_pre,
                     seg_cap);
  GPQ_HIP(hipGetLastError());
  Event* cur = buf0;
  Event* other = buf1;
  uint32_t passes = 0;
  for (uint64_t w = kTile; w < cap; w *= 2) ++passes;
  uint32_t pass = 0;
  for (uint64_t w = kTile; w < cap; w *= 2, ++pass) {
    hipLaunchKernelGGL(merge_pass_kernel<kT>, dim3(grid_of(cap, kTile)), dim3(kT), 0, stream, cur, other, d_n,
                       (uint32_t)w, pass, passes);
    GPQ_HIP(hipGetLastError());
    Event* t = cur;
    cur = other;
    other = t;
  }
  *sorted = cur;
  return GPQ_OK;
}
}  // namespace

bool valid_sort_tile(uint32_t tile) { return tile == 1024 || tile == 2048 || tile == 4096; }

uint32_t sort_passes(uint32_t cap, uint32_t tile) {
  uint32_t passes = 0;
  for (uint64_t w = tile; w < cap; w *= 2) ++passes;
  return passes;
}

gpq_status sort_records(const Event* in, Event* buf0, Event* buf1, const uint32_t* d_n, uint32_t cap,
                        hipStream_t stream, Event** sorted, uint32_t tile, unsigned long long* acct,
                        const uint32_t* seg_pre, uint32_t seg_cap) {
  gpq_status s;
  switch (tile) {
    case 1024: s = sort_records_t<128>(in, buf0, buf1, d_n, cap, stream, sorted, seg_pre, seg_cap); break;
    case 2048: s = sort_records_t<256>(in, buf0, buf1, d_n, cap, stream, sorted, seg_pre, seg_cap); break;
    case 4096: s = sort_records_t<512>(in, buf0, buf1, d_n, cap, stream, sorted, seg_pre, seg_cap); break;
    default:
      set_error("sort tile must be 1024, 2048 or 4096 records");
      return GPQ_ERR_INVALID;
  }
  if (s != GPQ_OK || !acct) return s;
  hipLaunchKernelGGL(acct_count_kernel, dim3(1), dim3(64), 0, stream, d_n, cap, acct);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

uint32_t merge_split_words(uint64_t max_total, uint32_t tile) {
  return (uint32_t)((max_total + tile - 1) / tile) + 2u;
}

namespace {
template <int kT>
gpq_status merge_tasks_t(const MergeTask* d_tasks, uint32_t ntasks, uint64_t max_total, hipStream_t stream,
                         uint32_t* splits, uint32_t stride) {
  constexpr uint32_t kTile = kT * kPerThread;
  const uint32_t gx = grid_of(max_total, kTile);
  if (splits) {
    GPQ_CHECK_ARG(stride >= gx + 1u, "merge_tasks: split buffer stride below the tile count + 1");
    hipLaunchKernelGGL(merge_partition_kernel<kT>, dim3(ceil_div(gx + 1u, 256u / kSplitLanes), ntasks), dim3(256), 0,
                       stream, d_tasks, gx + 1u, splits, stride);
  }
  hipLaunchKernelGGL(merge_tasks_kernel<kT>, dim3(gx, ntasks), dim3(kT), 0, stream, d_tasks,
                     static_cast<const uint32_t*>(splits), stride);
  return GPQ_OK;
}
}  // namespace

gpq_status merge_tasks(const MergeTask* d_tasks, uint32_t ntasks, uint64_t max_total, hipStream_t stream,
                       uint32_t tile, unsigned long long* acct, uint32_t* splits, uint32_t stride,
                       uint64_t partition_min) {
  if (ntasks == 0) return GPQ_OK;
  // Below 2^20 records a task's blocks search their splits themselves: the partition launch
  // would cost about what it saves (four heaps of a quarter of the LPs each measured 2.5e8 events/s
  // without it and 1.7e8 with it at 2^20 LPs; one heap and the GAQ gain).
  if (max_total < partition_min) splits = nullptr;
  switch (tile) {
    case 1024: GPQ_TRY(merge_tasks_t<128>(d_tasks, ntasks, max_total, stream, splits, stride)); break;
    case 2048: GPQ_TRY(merge_tasks_t<256>(d_tasks, ntasks, max_total, stream, splits, stride)); break;
    case 4096: GPQ_TRY(merge_tasks_t<512>(d_tasks, ntasks, max_total, stream, splits, stride)); break;
    default:
      set_error("merge tile must be 1024, 2048 or 4096 records");
      return GPQ_ERR_INVALID;
  }
  GPQ_HIP(hipGetLastError());
  if (acct) {
    hipLaunchKernelGGL(acct_tasks_kernel, dim3(1), dim3(256), 0, stream, d_tasks, ntasks, acct);
    GPQ_HIP(hipGetLastError());
  }
  return GPQ_OK;
}

gpq_status select_safe(const Event* const* s, const uint32_t* d_n, uint32_t cap, const uint32_t* d_wend,
                       const Event* d_bound, const uint32_t* d_bound_valid, const SelectBufs& b, Event* deq,
                       Event* rest, hipStream_t stream, unsigned long long* acct) {
  const uint32_t nblk = grid_of(cap, kSelectTile);
  uint32_t* counts = b.scratch;
  uint32_t* offsets = b.scratch + nblk;
  hipLaunchKernelGGL(select_prefix_kernel, dim3(1), dim3(64), 0, stream, s, d_n, d_wend, d_bound, d_bound_valid,
                     b.d_prefix);
  hipLaunchKernelGGL(select_mark_kernel, dim3(grid_of(cap, kThreads) < 4096 ? grid_of(cap, kThreads) : 4096),
                     dim3(kThreads), 0, stream, s, b.d_prefix, b.first);
  hipLaunchKernelGGL(select_count_kernel, dim3(nblk), dim3(kThreads), 0, stream, s, b.d_prefix, b.first, counts);
  hipLaunchKernelGGL(select_scan_kernel, dim3(1), dim3(1024), 0, stream, counts, offsets, nblk, b.d_prefix, b.d_deq_n,
                     b.d_rest_n);
  hipLaunchKernelGGL(select_scatter_kernel, dim3(nblk), dim3(kThreads), 0, stream, s, b.d_prefix, b.first, offsets, deq,
                     rest);
  GPQ_HIP(hipGetLastError());
  if (acct) {
    hipLaunchKernelGGL(acct_count_kernel, dim3(1), dim3(64), 0, stream, b.d_prefix, cap, acct);
    GPQ_HIP(hipGetLastError());
  }
  return GPQ_OK;
}

gpq_status gather_inserts(uint32_t* seg_m, uint32_t seg_cap, uint32_t spill_cap, uint32_t max_total, uint32_t* seg_pre,
                          uint32_t* d_m, WindowState* ws, hipStream_t stream) {
  static_assert(kInsertStripes + 1 <= kWave, "one wave gathers the stripes");
  hipLaunchKernelGGL(gather_inserts_kernel, dim3(1), dim3(kWave), 0, stream, seg_m, seg_cap, spill_cap, max_total,
                     seg_pre, d_m, ws);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status copy_records(const Event* src, Event* dst, const uint32_t* d_n, uint32_t cap, hipStream_t stream) {
  const uint32_t g = grid_of(cap, kThreads);
  hipLaunchKernelGGL(copy_records_kernel, dim3(g < 8192 ? g : 8192), dim3(kThreads), 0, stream, src, dst, d_n, cap);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

}  // namespace gpq

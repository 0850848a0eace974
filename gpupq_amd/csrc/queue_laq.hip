// queue_laq.hip -- local array queue: per-LP sorted rings, wave-per-LP merge of
// the staged inserts, fused with the reduction that publishes the next window.
//
// Reference: src/local_array_queue.cu.  The reference restores order after a step
// with a serial per-LP insertion sort (d_insertion_sort :254-299, sort.cu
// insert_sorted :329-380) or a block-per-LP parallel shift (:187-251), and finds
// the window minimum with a per-LP head scan plus a device-launched cub reduce
// (:116-146, :460-478).  Here one kernel does both: a block owns 256 LPs, its
// waves merge the LPs that received inserts (coalesced reads of the displaced
// tail, one shifted write per moved record), then the block folds the 256 heads
// and the last block publishes gmin / window end / done.
#include <algorithm>
#include <cmath>

#include "laq_calendar.cuh"
#include "laq_merge.cuh"
#include "phold_fused.cuh"
#include "queues.h"

namespace gpq {
namespace {

using namespace laq;

constexpr int kLaqBlock = 256;   // LPs per post block = threads per block (4 waves)

// One block owns kLaqBlock LPs.  Phase 1 reads their counters coalesced into LDS and
// lists the LPs that received inserts; phase 2 merges them, one LP per group of G
// lanes (kLaqBlock / G LPs in flight per block); phase 3 folds the block's heads into
// the window reduction.
template <int G, int SL>
__global__ __launch_bounds__(kLaqBlock, 5) void laq_post_kernel(LaqView q, uint32_t* partial) {
  __shared__ uint32_t s_work[kLaqBlock];
  __shared__ uint32_t s_k[kLaqBlock], s_n[kLaqBlock], s_r[kLaqBlock];
  __shared__ uint32_t s_hist[kLaqBlock];
  __shared__ uint32_t s_count;
  __shared__ uint32_t s_inserted;
  __shared__ uint32_t s_moved;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  if (q.block_live && !q.block_live[blockIdx.x]) {     // a block of unused LPs: nothing staged, no heads
    if (threadIdx.x == 0) partial[blockIdx.x] = kInfBits;
    return;
  }
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
  }
  __syncthreads();
  uint32_t staged = 0;
  if (lp < q.num_lps) {
    staged = q.ins[lp];
    if (staged) {
      const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
      const uint32_t n = q.len[lp];
      if (n + k > (1u << q.cap_log2)) {
        atomicOr(&q.ws->error, kErrQueueOverflow);
        q.ins[lp] = 0;
      } else {
        const uint32_t slot = atomicAdd(&s_count, 1u);
        s_work[slot] = threadIdx.x;
        s_k[threadIdx.x] = k;
        s_n[threadIdx.x] = n;
        s_r[threadIdx.x] = q.rd[lp];
        if (q.stats) atomicAdd(&s_inserted, k);
      }
    }
  }
  __syncthreads();
  const int lane = lane_id();
  const int g = lane % G;
  const int gbase = lane - g;
  const int group = (int)threadIdx.x / G;
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)group; i < s_count; i += kLaqBlock / G) {
    const uint32_t t = s_work[i];
    const uint32_t lpt = blockIdx.x * kLaqBlock + t;
    moved += laq_merge_group<G, SL>(q, lpt, q.stage + (size_t)lpt * q.stage_cap, s_k[t], s_n[t], s_r[t], g, gbase,
                                    s_hist + group * G);
  }
  if (q.stats && g == 0 && moved) atomicAdd(&s_moved, moved);
  __syncthreads();
  if (q.stats && threadIdx.x == 0) {   // work counters: one device atomic per block and counter
    const uint32_t lps_here = min((uint32_t)kLaqBlock, q.num_lps - blockIdx.x * kLaqBlock);
    atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
    atomicAdd(&q.stats[1], (unsigned long long)s_moved);
    atomicAdd(&q.stats[2], (unsigned long long)s_count);
    atomicAdd(&q.stats[3], (unsigned long long)lps_here);
  }
  const uint32_t v = lp < q.num_lps ? q.head[lp] : kInfBits;
  block_min_to_partial(v, partial);
}

// Two-tier post.  N = the near horizon this post uses (ws->near_next when the last window
// reached the old horizon, else ws->near_end).  Per LP with staged events, or with far events
// when the horizon moves: the candidates -- its far events first (when moving), then its
// staged events -- split at N; those below N are listed in LDS and merged into the ring, the
// others are written back to the far store (in place, compacted, then the staged ones after
// them).  Partials per block: ring-head minimum, far minimum, far count.

template <int G, int SL>
__global__ __launch_bounds__(kLaqBlock, 5) void laq_post_far_kernel(LaqView q, uint32_t* partial) {
  __shared__ uint32_t s_work[kLaqBlock];
  __shared__ uint32_t s_k[kLaqBlock], s_n[kLaqBlock], s_r[kLaqBlock], s_f[kLaqBlock];
  __shared__ uint32_t s_hist[kLaqBlock];
  __shared__ uint32_t s_count, s_inserted, s_moved;
  __shared__ unsigned long long s_aux;
  __shared__ Key3 s_near[kLaqBlock / G][kFarList];   // key words only: every record here has dst = its LP
  const uint32_t nb = gridDim.x;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  if (q.block_live && !q.block_live[blockIdx.x]) {     // a block of unused LPs: nothing staged, no heads
    if (threadIdx.x == 0) {
      partial[blockIdx.x] = kInfBits;
      partial[gridDim.x + blockIdx.x] = kInfBits;
      partial[2 * gridDim.x + blockIdx.x] = 0;
    }
    return;
  }
  const WindowState* hw = q.block_win ? &q.block_ws[q.block_win[blockIdx.x]] : q.ws;
  const uint32_t moving = hw->near_next;
  const uint32_t N = moving ? moving : hw->near_end;
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
    s_aux = 0;
  }
  __syncthreads();
  if (lp < q.num_lps) {
    const uint32_t staged = q.ins[lp];
    const uint32_t f = q.nfar[lp];
    if (q.stats && moving && f) atomicAdd(&s_aux, 32ull * f);   // far records read and written once
    if (staged || (moving && f)) {
      const uint32_t slot = atomicAdd(&s_count, 1u);
      s_work[slot] = threadIdx.x;
      s_k[threadIdx.x] = staged < q.stage_cap ? staged : q.stage_cap;
      s_f[threadIdx.x] = f;
      s_n[threadIdx.x] = q.len[lp];
      s_r[threadIdx.x] = q.rd[lp];
      if (q.stats) atomicAdd(&s_inserted, staged);
    }
  }
  __syncthreads();
  const int lane = lane_id();
  const int g = lane % G;
  const int gbase = lane - g;
  const int group = (int)threadIdx.x / G;
  const uint64_t gmask = (G == kWave ? ~0ull : ((1ull << G) - 1ull)) << gbase;
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)group; i < s_count; i += kLaqBlock / G) {
    const uint32_t t = s_work[i];
    const uint32_t lpt = blockIdx.x * kLaqBlock + t;
    const uint32_t k = s_k[t], f = s_f[t];
    const Event* stg = q.stage + (size_t)lpt * q.stage_cap;
    uint32_t n_far, fmin;
    const uint32_t n_near =
        split_candidates<G>(q, lpt, stg, k, f, moving != 0u, N, &s_near[group][0], g, gmask, &n_far, &fmin);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (g == 0) {                       // before the merge, so that n_far and fmin die here
      q.ins[lpt] = 0;
      q.nfar[lpt] = n_far < q.far_cap ? n_far : q.far_cap;
      q.far_min[lpt] = fmin;
    }
    const uint32_t nn = n_near < kFarList ? n_near : kFarList;
    const uint32_t n = s_n[t];
    if (n + nn > (1u << q.cap_log2)) {
      if (g == 0) atomicOr(&q.ws->error, kErrQueueOverflow);
    } else if (nn) {
      moved += laq_merge_group<G, SL>(q, lpt, &s_near[group][0], nn, n, s_r[t], g, gbase, s_hist + group * G);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (q.stats && g == 0 && moved) atomicAdd(&s_moved, moved);
  __syncthreads();
  if (q.stats && threadIdx.x == 0) {
    const uint32_t lps_here = min((uint32_t)kLaqBlock, q.num_lps - blockIdx.x * kLaqBlock);
    atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
    atomicAdd(&q.stats[1], (unsigned long long)s_moved);
    atomicAdd(&q.stats[2], (unsigned long long)s_count);
    atomicAdd(&q.stats[3], (unsigned long long)lps_here);
    if (s_aux) atomicAdd(&q.stats[4], s_aux);
  }
  // partials: ring-head minimum, far minimum, far count
  __shared__ uint32_t s_w[3][kLaqBlock / kWave];
  uint32_t hv = kInfBits, fv = kInfBits, fc = 0;
  if (lp < q.num_lps) {
    hv = q.head[lp];
    fv = q.far_min[lp];
    fc = q.nfar[lp];
  }
  hv = wave_min_u32(hv);
  fv = wave_min_u32(fv);
  fc = wave_sum_u32(fc);
  if (lane == 0) {
    s_w[0][threadIdx.x / kWave] = hv;
    s_w[1][threadIdx.x / kWave] = fv;
    s_w[2][threadIdx.x / kWave] = fc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = s_w[0][0], b = s_w[1][0], cc = s_w[2][0];
    for (int w = 1; w < kLaqBlock / kWave; ++w) {
      a = s_w[0][w] < a ? s_w[0][w] : a;
      b = s_w[1][w] < b ? s_w[1][w] : b;
      cc += s_w[2][w];
    }
    partial[blockIdx.x] = a;
    partial[nb + blockIdx.x] = b;
    partial[2 * nb + blockIdx.x] = cc;
  }
}

// Two-tier post with a horizon per LP (open lists: lookahead = +inf, as the A* model's), after the
// heap-backed LAQ's exchange for A* (src/local_array_queue.cu:150-180,515-681): there, an LP's
// array that holds fewer than A_STAR_MAX_PARALLEL_OL_ENTRIES items takes its heap's best node,
// and the items beyond the array's bound go to the heap.  Here the far store is the LP's
// unsorted backing store, and the horizon -- a full key (ts, src, seq), near_lp[0/1/2][lp] --
// separates the two: every ring key is below it and every far key at or above it, so the ring
// head is the LP's minimum.  Per LP with staged records, or with far records while its ring
// holds fewer than kOpenRefill:
//  1. refill (ring short, far store not empty): the new horizon is chosen so that at most
//     min(C/2, kFarList) - n - k far records come in (C ring slots, n ring records, k staged),
//     and at least one; with no room left and records in the ring, the refill waits;
//  2. split the candidates (the far records when refilling, then the staged ones) at the
//     horizon, as the two-tier post does;
//  3. evict (ring + near candidates > C, ring not empty): the horizon drops to the key of ring
//     record j = C/2 - nn (0 when negative), but stays above the LP's smallest key; ring
//     records and candidates at or above it go to the far store;
//  4. merge the near candidates into the ring.
// Limits (the queue's error flag, GPQ_ERR_CAPACITY): a full far store; more than C records
// with the LP's smallest key (keys are unique in the A* model: (f, h, cell)).
constexpr uint32_t kOpenRefill = 8;     // A_STAR_MAX_PARALLEL_OL_ENTRIES (src/config.h.a_star_heaps:92)
constexpr uint32_t kOpenBuckets = 32;

__device__ __forceinline__ uint32_t key_word(const Event& e, int lvl) {
  return lvl == 0 ? e.ts : (lvl == 1 ? e.src : e.seq);
}

// The key whose words above lvl are (p0, p1), word lvl is v (a carry past 32 bits goes up), the
// words below are 0.  Past the largest key: all ones (above every record, whose ts is a float's
// bits or an A* f, below 2^31).
__device__ __forceinline__ Key3 key_at(uint32_t p0, uint32_t p1, int lvl, uint64_t v) {
  const Key3 top{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  if (lvl == 0) return v >> 32 ? top : Key3{(uint32_t)v, 0u, 0u};
  if (lvl == 1) return v >> 32 ? (p0 == 0xFFFFFFFFu ? top : Key3{p0 + 1u, 0u, 0u}) : Key3{p0, (uint32_t)v, 0u};
  if (!(v >> 32)) return Key3{p0, p1, (uint32_t)v};
  return p1 != 0xFFFFFFFFu ? Key3{p0, p1 + 1u, 0u} : (p0 == 0xFFFFFFFFu ? top : Key3{p0 + 1u, 0u, 0u});
}

// The smallest key above k.
__device__ __forceinline__ Key3 key_next(const Key3& k) {
  if (k.seq != 0xFFFFFFFFu) return Key3{k.ts, k.src, k.seq + 1u};
  return key_at(k.ts, k.src, 1, (uint64_t)k.src + 1u);
}

template <int G>
__device__ __forceinline__ Key3 group_min_key(Key3 k) {
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    const Key3 o{(uint32_t)__shfl_xor((int)k.ts, off, kWave), (uint32_t)__shfl_xor((int)k.src, off, kWave),
                 (uint32_t)__shfl_xor((int)k.seq, off, kWave)};
    if (key_less3(o, k)) k = o;
  }
  return k;
}

// Step 1: the new horizon of LP lp from its f > 0 far records: below it at most `budget`
// records, and at least one.  Word by word (ts, then src, then seq) over the records whose
// higher words equal the prefix already fixed: the range [vmin, vmax] of the word is cut into
// kOpenBuckets buckets of width w; the buckets that fit the budget give the horizon; when the
// first bucket alone is too full, its range is cut again, and a width-1 bucket fixes the word.
template <int G>
__device__ Key3 open_refill_key(const LaqView& q, uint32_t lp, uint32_t f, uint32_t budget, uint32_t* bkt, int g,
                                uint64_t gmask) {
  const Event* far = q.far + (size_t)lp * q.far_cap;
  uint32_t p0 = 0, p1 = 0, lo = 0, hi = 0xFFFFFFFFu;
  int lvl = 0;
  // each round narrows the range 32-fold or fixes a word: at most 7 rounds per word
  for (int round = 0; round < 24; ++round) {
    uint32_t vmin = 0xFFFFFFFFu, vmax = 0;
    for (uint32_t i = (uint32_t)g; i < f; i += G) {
      const Event e = load_event(&far[i]);
      const uint32_t v = key_word(e, lvl);
      if ((lvl < 1 || e.ts == p0) && (lvl < 2 || e.src == p1) && v >= lo && v <= hi) {
        vmin = min(vmin, v);
        vmax = max(vmax, v);
      }
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      vmin = min(vmin, (uint32_t)__shfl_xor((int)vmin, off, kWave));
      vmax = max(vmax, (uint32_t)__shfl_xor((int)vmax, off, kWave));
    }
    const uint64_t w = ((uint64_t)vmax - vmin + kOpenBuckets) / kOpenBuckets;   // >= 1: the range fits
    for (uint32_t b = (uint32_t)g; b < kOpenBuckets; b += G) bkt[b] = 0;
    wave_lds_sync();
    for (uint32_t i = (uint32_t)g; i < f; i += G) {
      const Event e = load_event(&far[i]);
      const uint32_t v = key_word(e, lvl);
      if ((lvl < 1 || e.ts == p0) && (lvl < 2 || e.src == p1) && v >= vmin && v <= vmax)
        atomicAdd(&bkt[(uint32_t)(((uint64_t)v - vmin) / w)], 1u);
    }
    wave_lds_sync();
    uint32_t cum = 0, nb = 0;
    for (uint32_t b = 0; b < kOpenBuckets; ++b) {
      cum += bkt[b];
      if (cum > budget) break;
      nb = b + 1;
    }
    wave_lds_sync();                               // the buckets are read before the next round clears them
    if (nb) return key_at(p0, p1, lvl, (uint64_t)vmin + nb * w);
    if (w > 1) {                                   // the first bucket alone is too full: cut it again
      lo = vmin;
      hi = (uint32_t)((uint64_t)vmin + w - 1u);
      continue;
    }
    if (lvl == 2) break;                           // one full key held more than the budget: take it
    if (lvl == 0) p0 = vmin;
    else p1 = vmin;
    ++lvl;
    lo = 0;
    hi = 0xFFFFFFFFu;
  }
  // copies of one full key fill more than the budget: the smallest far key and its copies
  Key3 m{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  for (uint32_t i = (uint32_t)g; i < f; i += G) {
    const Event e = load_event(&far[i]);
    const Key3 k{e.ts, e.src, e.seq};
    if (key_less3(k, m)) m = k;
  }
  return key_next(group_min_key<G>(m));
}

// Step 3: ring run of *n records (> 0) at slot r, near candidates near[0 .. *nn), with
// *n + *nn > C.  The new horizon x is the key of ring record j = C/2 - nn (0 when negative), but
// at least the key after the LP's smallest, so that the LP's minimum stays in the ring.  Moves
// the ring's records from the first at or above x on, and the candidates at or above x, to the
// far store after its *n_far records; updates *n, *nn, *n_far, *fmin and the horizon *H.
template <int G>
__device__ void open_evict(const LaqView& q, uint32_t lp, uint32_t r, uint32_t* n_io, Key3* near, uint32_t* nn_io,
                           uint32_t* n_far_io, uint32_t* fmin_io, Key3* H_io, int g, uint64_t gmask) {
  const uint64_t below = (1ull << lane_id()) - 1ull;
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  const uint32_t half = (1u << q.cap_log2) / 2u;
  const uint32_t n = *n_io, nn = *nn_io;
  const Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  Event* far = q.far + (size_t)lp * q.far_cap;
  const uint32_t j = half > nn ? half - nn : 0u;   // j < n, since n + nn > C >= half
  const Event h = load_event(&ring[r & mask]);     // the smallest key: the run's first or a candidate's
  Key3 m{h.ts, h.src, h.seq};
  for (uint32_t i = (uint32_t)g; i < nn; i += G)
    if (key_less3(near[i], m)) m = near[i];
  m = group_min_key<G>(m);
  const Event ej = load_event(&ring[(r + j) & mask]);
  Key3 x{ej.ts, ej.src, ej.seq};
  const Key3 m1 = key_next(m);
  if (key_less3(x, m1)) x = m1;
  uint32_t j0 = 0;                                 // ring records below x: a prefix of the sorted run
  for (uint32_t base = 0; base < n; base += G) {
    const uint32_t i = base + (uint32_t)g;
    bool lt = false;
    if (i < n) {
      const Event e = load_event(&ring[(r + i) & mask]);
      lt = key_less3(Key3{e.ts, e.src, e.seq}, x);
    }
    j0 += (uint32_t)__popcll(__ballot(lt) & gmask);
  }
  uint32_t nf = *n_far_io, fmin = *fmin_io;
  for (uint32_t i = j0 + (uint32_t)g; i < n; i += G) {   // ring records [j0, n): x and above
    const Event e = load_event(&ring[(r + i) & mask]);
    const uint32_t dst = nf + (i - j0);
    if (dst < q.far_cap) store_event(&far[dst], e);
    else atomicOr(&q.ws->error, kErrQueueOverflow);
    fmin = min(fmin, e.ts);
  }
  nf += n - j0;
  uint32_t kept = 0;                               // candidates: below x kept (compacted in place), the rest out
  for (uint32_t base = 0; base < nn; base += G) {
    const uint32_t i = base + (uint32_t)g;
    const bool valid = i < nn;
    const Key3 k = valid ? near[i] : Key3{0u, 0u, 0u};
    const bool keep = valid && key_less3(k, x), out = valid && !keep;
    const uint64_t bk = __ballot(keep) & gmask, bo = __ballot(out) & gmask;
    if (keep) near[kept + (uint32_t)__popcll(bk & below)] = k;
    if (out) {
      const uint32_t dst = nf + (uint32_t)__popcll(bo & below);
      if (dst < q.far_cap) store_event(&far[dst], Event{k.ts, lp, k.src, k.seq});
      else atomicOr(&q.ws->error, kErrQueueOverflow);
      fmin = min(fmin, k.ts);
    }
    kept += (uint32_t)__popcll(bk);
    nf += (uint32_t)__popcll(bo);
  }
#pragma unroll
  for (int off = 1; off < G; off <<= 1) fmin = min(fmin, (uint32_t)__shfl_xor((int)fmin, off, kWave));
  *n_io = j0;
  *nn_io = kept;
  *n_far_io = nf;
  *fmin_io = fmin;
  *H_io = x;
}

template <int G, int SL>
__global__ __launch_bounds__(kLaqBlock) void laq_post_open_kernel(LaqView q, uint32_t* partial) {
  __shared__ uint32_t s_work[kLaqBlock];
  __shared__ uint32_t s_k[kLaqBlock], s_n[kLaqBlock], s_r[kLaqBlock], s_f[kLaqBlock];
  __shared__ uint32_t s_hist[kLaqBlock];
  __shared__ uint32_t s_bkt[kLaqBlock / G][kOpenBuckets];
  __shared__ uint32_t s_count, s_inserted, s_moved;
  __shared__ unsigned long long s_aux;
  __shared__ Key3 s_near[kLaqBlock / G][kFarList];
  const uint32_t nb = gridDim.x;
  const uint32_t L = q.num_lps;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  const uint32_t C = 1u << q.cap_log2;
  const uint32_t refill_below = min(kOpenRefill, max(1u, C / 4u));
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
    s_aux = 0;
  }
  __syncthreads();
  if (lp < L) {
    const uint32_t staged = q.ins[lp];
    const uint32_t f = q.nfar[lp];
    const uint32_t n = q.len[lp];
    if (staged || (f && n < refill_below)) {
      const uint32_t slot = atomicAdd(&s_count, 1u);
      s_work[slot] = threadIdx.x;
      s_k[threadIdx.x] = staged < q.stage_cap ? staged : q.stage_cap;
      s_f[threadIdx.x] = f;
      s_n[threadIdx.x] = n;
      s_r[threadIdx.x] = q.rd[lp];
      if (q.stats) atomicAdd(&s_inserted, staged);
    }
  }
  __syncthreads();
  const int lane = lane_id();
  const int g = lane % G;
  const int gbase = lane - g;
  const int group = (int)threadIdx.x / G;
  const uint64_t gmask = (G == kWave ? ~0ull : ((1ull << G) - 1ull)) << gbase;
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)group; i < s_count; i += kLaqBlock / G) {
    const uint32_t t = s_work[i];
    const uint32_t lpt = blockIdx.x * kLaqBlock + t;
    const uint32_t k = s_k[t], f = s_f[t], n0 = s_n[t], r = s_r[t];
    uint32_t n = n0;
    Key3 H{q.near_lp[lpt], q.near_lp[L + lpt], q.near_lp[2 * L + lpt]};
    bool refill = f && n < refill_below;
    if (refill) {
      const uint32_t room = min(C / 2u, kFarList);
      const uint32_t budget = room > n + k ? room - n - k : 0u;
      if (!budget && n) {
        refill = false;                            // no room: the ring's records are popped first
      } else {
        H = open_refill_key<G>(q, lpt, f, budget ? budget : 1u, &s_bkt[group][0], g, gmask);
        if (q.stats && g == 0) atomicAdd(&s_aux, 32ull * f);   // far records read and written once
      }
    }
    const Event* stg = q.stage + (size_t)lpt * q.stage_cap;
    uint32_t n_far, fmin;
    const uint32_t n_near = split_candidates_by<G>(
        q, lpt, stg, k, f, refill, [H](const Event& e) { return key_less3(Key3{e.ts, e.src, e.seq}, H); },
        &s_near[group][0], g, gmask, &n_far, &fmin);
    wave_lds_sync();
    uint32_t nn = n_near < kFarList ? n_near : kFarList;
    if (n + nn > C && n) {
      open_evict<G>(q, lpt, r, &n, &s_near[group][0], &nn, &n_far, &fmin, &H, g, gmask);
      wave_lds_sync();
    }
    const bool fits = n + nn <= C;                 // else: more than C records of the LP's smallest key
    if (g == 0) {
      if (!fits) atomicOr(&q.ws->error, kErrQueueOverflow);
      q.ins[lpt] = 0;
      q.nfar[lpt] = n_far < q.far_cap ? n_far : q.far_cap;
      q.far_min[lpt] = fmin;
      q.near_lp[lpt] = H.ts;
      q.near_lp[L + lpt] = H.src;
      q.near_lp[2 * L + lpt] = H.seq;
      if (n != n0) {
        q.len[lpt] = n;
        if (!n) q.head[lpt] = kInfBits;            // a record merged at index 0 sets it again
      }
    }
    if (nn && fits) moved += laq_merge_group<G, SL>(q, lpt, &s_near[group][0], nn, n, r, g, gbase, s_hist + group * G);
    __builtin_amdgcn_wave_barrier();
  }
  if (q.stats && g == 0 && moved) atomicAdd(&s_moved, moved);
  __syncthreads();
  if (q.stats && threadIdx.x == 0) {
    const uint32_t lps_here = min((uint32_t)kLaqBlock, q.num_lps - blockIdx.x * kLaqBlock);
    atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
    atomicAdd(&q.stats[1], (unsigned long long)s_moved);
    atomicAdd(&q.stats[2], (unsigned long long)s_count);
    atomicAdd(&q.stats[3], (unsigned long long)lps_here);
    if (s_aux) atomicAdd(&q.stats[4], s_aux);
  }
  // partials: ring-head minimum, far minimum, far count
  __shared__ uint32_t s_w[3][kLaqBlock / kWave];
  uint32_t hv = kInfBits, fv = kInfBits, fc = 0;
  if (lp < q.num_lps) {
    hv = q.head[lp];
    fv = q.far_min[lp];
    fc = q.nfar[lp];
  }
  hv = wave_min_u32(hv);
  fv = wave_min_u32(fv);
  fc = wave_sum_u32(fc);
  if (lane == 0) {
    s_w[0][threadIdx.x / kWave] = hv;
    s_w[1][threadIdx.x / kWave] = fv;
    s_w[2][threadIdx.x / kWave] = fc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = s_w[0][0], b = s_w[1][0], cc = s_w[2][0];
    for (int w = 1; w < kLaqBlock / kWave; ++w) {
      a = s_w[0][w] < a ? s_w[0][w] : a;
      b = s_w[1][w] < b ? s_w[1][w] : b;
      cc += s_w[2][w];
    }
    partial[blockIdx.x] = a;
    partial[nb + blockIdx.x] = b;
    partial[2 * nb + blockIdx.x] = cc;
  }
}

// One block: gmin = min(ring heads, far minima); the window ends at
// min(gmin + lookahead, horizon, near horizon); when the next window could reach the near
// horizon, its move to gmin + 2 lookahead + step is scheduled for the next post.
__global__ void laq_far_publish_kernel(const uint32_t* partial, uint32_t nb, WindowState* ws, float lookahead,
                                       float horizon, float step) {
  __shared__ uint32_t s_a[16], s_b[16], s_c[16];
  uint32_t a = kInfBits, b = kInfBits, c = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) {
    a = partial[i] < a ? partial[i] : a;
    b = partial[nb + i] < b ? partial[nb + i] : b;
    c += partial[2 * nb + i];
  }
  a = wave_min_u32(a);
  b = wave_min_u32(b);
  c = wave_sum_u32(c);
  if (lane_id() == 0) {
    s_a[threadIdx.x / kWave] = a;
    s_b[threadIdx.x / kWave] = b;
    s_c[threadIdx.x / kWave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) {
      a = s_a[i] < a ? s_a[i] : a;
      b = s_b[i] < b ? s_b[i] : b;
      c += s_c[i];
    }
    a = s_a[0] < a ? s_a[0] : a;
    b = s_b[0] < b ? s_b[0] : b;
    if (ws->near_next) {
      ws->near_end = ws->near_next;
      ws->near_next = 0;
    }
    const uint32_t g = a < b ? a : b;
    const float gf = bits_to_f(g);
    float wend = gf + lookahead;
    if (horizon < wend) wend = horizon;
    const float nearf = bits_to_f(ws->near_end);
    if (nearf < wend) wend = nearf;
    // scheduled one window ahead: the next window, which can reach gmin + 2 lookahead, then
    // starts below the moved horizon instead of being cut short at the old one
    if (g != kInfBits && gf + 2.0f * lookahead > nearf) {
      float next = gf + 2.0f * lookahead + step;
      if (!(next > nearf)) next = nearf + step;
      ws->near_next = f_to_bits(next);
    }
    ws->gmin = g;
    ws->wend = f_to_bits(wend);
    ws->done = (g == kInfBits || !(gf < horizon)) ? 1u : 0u;
    ws->windows += 1;
    (void)c;
  }
}

// queue_post after a model step that merged and committed in one kernel (the PHOLD step of
// drain mode 2, which owns the second staging buffer): advance_window.  settle: only the
// buffer choice is reset (both buffers are empty after a settle pass).
__global__ void laq_step_publish_kernel(WindowState* ws, float lookahead, float end_time, float step,
                                        uint32_t settle) {
  if (settle) {
    ws->sel = 0;
    return;
  }
  advance_window(ws, lookahead, end_time, step);
}

__global__ void laq_init_kernel(LaqView q) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp < q.num_lps) {
    q.rd[lp] = 0;
    q.len[lp] = 0;
    q.ins[lp] = 0;
    q.head[lp] = kInfBits;
    if (q.far_cap) {
      q.nfar[lp] = 0;
      q.far_min[lp] = kInfBits;
    }
    if (q.near_lp) {                           // open lists: everything to the ring until it fills
      q.near_lp[lp] = 0xFFFFFFFFu;
      q.near_lp[q.num_lps + lp] = 0xFFFFFFFFu;
      q.near_lp[2 * q.num_lps + lp] = 0xFFFFFFFFu;
    }
  }
  if (lp == 0) {
    q.ws->gmin = kInfBits;
    q.ws->wend = 0;
    q.ws->done = 0;
    q.ws->near_end = q.near_lp ? kInfBits : 0u;   // open lists: no global horizon
    q.ws->near_next = 0;
    q.ws->windows = 0;
    q.ws->error = 0;
    q.ws->sel = 0;
  }
}

__global__ void laq_clear_kernel(LaqView q, uint32_t lp) { q.clear((int)lp); }

// Empties every LP and unmarks every id (decrease-key mode): the queue as after init.
__global__ void laq_empty_kernel(LaqView q, uint32_t words) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < q.num_lps || i < words; i += stride) {
    if (i < q.num_lps) {
      q.len[i] = 0;
      q.ins[i] = 0;
      q.head[i] = kInfBits;
      q.upd_n[i] = 0;
    }
    if (i < words) q.in_queue[i] = 0;
  }
}

// Decrease-key pass of queue_post (src/local_array_queue.cu:350-416, execute_update), after the
// staged inserts are merged: each LP's buffered updates in buffer order.  An update whose id
// has an entry replaces it when its key is smaller and is dropped otherwise; an id whose entry
// was removed after the update was buffered is inserted anew.  One thread per LP moves the
// record down its sorted run; the block's partial minimum is lowered to the new head.
__global__ void laq_update_kernel(LaqView q, uint32_t* partial) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= q.num_lps) return;
  const uint32_t nu = q.upd_n[lp];
  if (!nu) return;
  const uint32_t C = 1u << q.cap_log2, mask = C - 1u;
  Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  const uint32_t r = q.rd[lp];
  uint32_t n = q.len[lp];
  const uint32_t cnt = nu < q.upd_cap ? nu : q.upd_cap;
  for (uint32_t u = 0; u < cnt; ++u) {
    const Event e = load_event(&q.upd[(size_t)lp * q.upd_cap + u]);
    uint32_t p = n;
    for (uint32_t i = 0; i < n; ++i)
      if (ring[(r + i) & mask].seq == e.seq) {
        p = i;
        break;
      }
    if (p < n) {
      if (!key_less(e, load_event(&ring[(r + p) & mask]))) continue;   // not cheaper: dropped
    } else {
      if (n == C) {
        atomicOr(&q.ws->error, kErrQueueOverflow);
        continue;
      }
      atomicOr(&q.in_queue[e.seq >> 5], 1u << (e.seq & 31u));   // back in the queue
      ++n;                                                     // slot p = old n is free
    }
    uint32_t j = p;   // larger records from j to p - 1 move up one slot
    while (j > 0) {
      const Event prev = load_event(&ring[(r + j - 1u) & mask]);
      if (!key_less(e, prev)) break;
      store_event(&ring[(r + j) & mask], prev);
      --j;
    }
    store_event(&ring[(r + j) & mask], e);
  }
  q.len[lp] = n;
  q.upd_n[lp] = 0;
  const uint32_t h = n ? ring[r].ts : kInfBits;
  q.head[lp] = h;
  atomicMin(&partial[lp / kLaqBlock], h);
}

// ---- calendar mode (laq_calendar.cuh) -------------------------------------------------

// queue_post_init: the staged initial events go to the list over[0]
// (the window ends, and so the logs, are known only once every initial event is in); the block
// minima of their timestamps feed cal_init_publish_kernel.
// The block reserves one run for all its LPs' staged events (one atomic per block: an atomic per
// wave on the list counter took 192 us per launch at 2^20 LPs), and each thread writes its LP's
// events at its prefix in the run.
__global__ __launch_bounds__(kLaqBlock) void cal_post_init_kernel(LaqView q, uint32_t* partial) {
  __shared__ uint32_t s_wave[kLaqBlock / kWave];
  __shared__ uint32_t s_base;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  const int lane = lane_id(), wave = (int)threadIdx.x / kWave;
  uint32_t k0 = 0, k = 0;
  if (lp < q.num_lps) {
    k0 = q.ins[lp];
    k = k0 < q.stage_cap ? k0 : q.stage_cap;
  }
  uint32_t incl = k;                                   // inclusive prefix of k over the wave
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < kLaqBlock / kWave; ++w) {    // wave totals -> exclusive wave offsets
      const uint32_t t = s_wave[w];
      s_wave[w] = tot;
      tot += t;
    }
    s_base = tot ? atomicAdd(&q.over_n[0], tot) : 0u;
  }
  __syncthreads();
  const uint32_t start = s_base + s_wave[wave] + (incl - k);
  uint32_t m = kInfBits;
  for (uint32_t i = 0; i < k; ++i) {
    const Event e = load_event(&q.stage[(size_t)lp * q.stage_cap + i]);
    m = e.ts < m ? e.ts : m;
    if (start + i < q.over_cap) store_event(&q.over[start + i], e);
    else atomicOr(&q.ws->error, kErrQueueOverflow);
  }
  if (k0) q.ins[lp] = 0;
  block_min_to_partial(m, partial);
}

// Window publication after a post_init chunk: gmin = the smaller of the chunk's minimum and
// the minimum of the chunks before, the window end one lookahead above it (capped).
__global__ void cal_init_publish_kernel(const uint32_t* partial, uint32_t count, WindowState* ws, float lookahead,
                                        float horizon) {
  __shared__ uint32_t s_w[16];
  uint32_t m = kInfBits;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) m = partial[i] < m ? partial[i] : m;
  m = wave_min_u32(m);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t g = ws->gmin;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) g = s_w[i] < g ? s_w[i] : g;
    float wend = bits_to_f(g) + lookahead;
    if (horizon < wend) wend = horizon;
    ws->gmin = g;
    ws->wend = f_to_bits(wend);
    ws->done = (g == kInfBits || !(bits_to_f(g) < horizon)) ? 1u : 0u;
    ws->windows += 1;
  }
}

// The ends of windows 0 .. B from the first one (published from the initial events), each one
// lookahead after the previous in float arithmetic, capped at the end time.
__global__ void cal_start_kernel(LaqView q, float lookahead, float end_time) {
  CalState* st = q.cal_st;
  const uint32_t B = q.cal_windows;
  st->cur = 0;
  st->fin = 0;
  uint32_t w = q.ws->wend;
  st->w[0] = w;
  for (uint32_t i = 1; i <= B; ++i) {
    float nw = bits_to_f(w) + lookahead;
    if (end_time < nw) nw = end_time;
    w = f_to_bits(nw);
    st->w[i] = w;
  }
  q.ws->sel = 0;
}

// Files every initial event (the list over[0]) in the log of its window (windows 0 .. B - 1)
// and LP block, in over[1] (read by the first step), or with the parked ones.
__global__ __launch_bounds__(cal::kBlock) void cal_fill_kernel(LaqView q, float end_time) {
  __shared__ uint32_t sw[kCalMaxWindows];
  __shared__ cal::ListAppend la;
  cal::load_windows(q, 0u, sw);
  __syncthreads();
  const uint32_t B = q.cal_windows, end_bits = f_to_bits(end_time);
  const uint32_t n0 = q.over_n[0], n = n0 < q.over_cap ? n0 : q.over_cap;
  const Event* src = q.over;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {   // block-uniform trip count
    const uint32_t i = base + threadIdx.x;
    const bool has = i < n;
    const Event e = has ? load_event(&src[i]) : Event{};
    const uint32_t r = has ? cal::route_of(sw, B, e.ts, end_bits, 0u) : cal::Route::kNone;
    cal::block_file(q, 0u, 1u, r, e, la);
  }
}

__global__ void cal_fill_done_kernel(LaqView q) { q.over_n[0] = 0; }

// Pending records: every log (grid-stride over the B x blocks counters), both beyond-the-calendar
// lists and the parked ones (by the first thread), and the staging area (one LP per thread:
// events inserted since the last queue_post).
__global__ void cal_count_kernel(LaqView q, unsigned long long* sum) {
  __shared__ unsigned long long s[16];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (t < q.num_lps) n = min(q.ins[t], q.stage_cap);
  const uint32_t logs = q.cal_windows * q.cal_blocks;
  for (uint32_t i = t; i < logs; i += gridDim.x * blockDim.x) n += min(q.wlog_n[(size_t)i << cal::kLogCountShift], q.cal_sub_cap);
  if (t == 0) {
    for (uint32_t p = 0; p < 2; ++p) n += min(q.over_n[p], q.over_cap);
    n += min(*q.parked_n, q.parked_cap);
  }
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, kWave);
  if (lane_id() == 0) s[threadIdx.x / kWave] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) tot += s[w];
    if (tot) atomicAdd(sum, tot);
  }
}

// Every pending record owned by LP lp (unsorted) into out[0 .. *n) (at most cap), grid-stride
// over the logs of its block, both lists, the parked ones and its staging slots.
__device__ __forceinline__ void gather_one(const Event* e, uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  if (e->dst == lp) {
    const uint32_t k = atomicAdd(n, 1u);
    if (k < cap) out[k] = *e;
  }
}

__global__ void cal_gather_kernel(LaqView q, uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t b = lp >> q.cal_block_log2;
  for (uint32_t x = 0; x < q.cal_windows; ++x) {
    const uint32_t c = min(*cal::log_count(q, x, b), q.cal_sub_cap);
    const Event* lg = cal::log_base(q, x, b);
    for (uint32_t i = t; i < c; i += stride) gather_one(&lg[i], lp, out, cap, n);
  }
  for (uint32_t p = 0; p < 2; ++p) {
    const uint32_t c = min(q.over_n[p], q.over_cap);
    for (uint32_t i = t; i < c; i += stride) gather_one(&q.over[(size_t)p * q.over_cap + i], lp, out, cap, n);
  }
  const uint32_t c = min(*q.parked_n, q.parked_cap);
  for (uint32_t i = t; i < c; i += stride) gather_one(&q.parked[i], lp, out, cap, n);
  if (t == 0) {
    const uint32_t k0 = min(q.ins[lp], q.stage_cap);
    for (uint32_t i = 0; i < k0; ++i) gather_one(&q.stage[(size_t)lp * q.stage_cap + i], lp, out, cap, n);
  }
}

}  // namespace

gpq_status LaqQueue::init() {
  if (cfg.calendar_windows) return init_calendar();
  GPQ_CHECK_ARG(is_pow2(cfg.capacity_per_lp) && cfg.capacity_per_lp >= 2,
                "LAQ: capacity_per_lp must be a power of two >= 2");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LAQ: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG(cfg.capacity_per_lp <= 8u * 64u, "LAQ: capacity_per_lp must be <= 512");
  GPQ_CHECK_ARG((uint64_t)cfg.num_lps * cfg.capacity_per_lp < (1ull << 31),
                "LAQ: num_lps * capacity_per_lp must be < 2^31 (int item indices, src/queue.cuh:51)");
  const uint32_t L = cfg.num_lps;
  v_.num_lps = L;
  v_.cap_log2 = log2u(cfg.capacity_per_lp);
  v_.stage_cap = cfg.stage_per_lp;
  // lanes per LP merge: enough to hold the run (kSlots records per lane); staged events are
  // merged G - 1 at a time
  group_ = 8;
  while (group_ < 64 && group_ * (uint32_t)kSlots < cfg.capacity_per_lp) group_ *= 2;
  GPQ_TRY(arena.alloc(&v_.ring, (size_t)L << v_.cap_log2));
  GPQ_TRY(arena.alloc(&v_.stage, (size_t)L * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.rd, L));
  GPQ_TRY(arena.alloc(&v_.len, L));
  GPQ_TRY(arena.alloc(&v_.ins, L));
  GPQ_TRY(arena.alloc(&v_.head, L));
  GPQ_TRY(arena.alloc(&sum_, 1));
  v_.ws = ws;
  v_.stats = stats;
  partial_slots = ceil_div(L, kLaqBlock);
  if (cfg.far_per_lp) {
    // lookahead = +inf: open lists (a horizon per LP); else a global near horizon
    const bool open = std::isinf(cfg.lookahead) && cfg.lookahead > 0.0f;
    GPQ_CHECK_ARG(open || (std::isfinite(cfg.lookahead) && cfg.lookahead > 0.0f),
                  "LAQ two-tier: needs a finite lookahead > 0, or lookahead = +inf (open lists)");
    GPQ_CHECK_ARG(cfg.far_per_lp <= 4096, "LAQ two-tier: far_per_lp must be <= 4096");
    GPQ_CHECK_ARG(group_ <= 16, "LAQ two-tier: capacity_per_lp must be <= 128");
    GPQ_CHECK_ARG(cfg.stage_per_lp <= kFarList, "LAQ two-tier: stage_per_lp must be <= 64");
    GPQ_CHECK_ARG((uint64_t)L * cfg.far_per_lp < (1ull << 32), "LAQ two-tier: num_lps * far_per_lp too large");
    GPQ_CHECK_ARG(!open || cfg.capacity_per_lp >= 2u * cfg.stage_per_lp,
                  "LAQ two-tier open lists: capacity_per_lp must be >= 2 * stage_per_lp");
    v_.far_cap = cfg.far_per_lp;
    far_step_ = open ? 0.0f : (cfg.far_step > 0.0f ? cfg.far_step : 4.0f * cfg.lookahead);
    GPQ_TRY(arena.alloc(&v_.far, (size_t)L * v_.far_cap));
    GPQ_TRY(arena.alloc(&v_.nfar, L));
    GPQ_TRY(arena.alloc(&v_.far_min, L));
    if (open) GPQ_TRY(arena.alloc(&v_.near_lp, 3 * (size_t)L));   // the horizon's three key words per LP
  }
  if (cfg.update_ids) {
    GPQ_CHECK_ARG(!cfg.far_per_lp, "LAQ decrease-key mode needs the one-tier LAQ (far_per_lp = 0)");
    v_.update_ids = cfg.update_ids;
    v_.upd_cap = cfg.update_per_lp ? cfg.update_per_lp : cfg.stage_per_lp;
    GPQ_CHECK_ARG(v_.upd_cap <= 1024, "LAQ: update_per_lp must be <= 1024");
    const size_t words = ((size_t)cfg.update_ids + 31) / 32;
    GPQ_TRY(arena.alloc(&v_.in_queue, words));
    GPQ_TRY(arena.alloc(&v_.upd, (size_t)L * v_.upd_cap));
    GPQ_TRY(arena.alloc(&v_.upd_n, L));
    GPQ_HIP(hipMemsetAsync(v_.in_queue, 0, sizeof(uint32_t) * words, stream));
    GPQ_HIP(hipMemsetAsync(v_.upd_n, 0, sizeof(uint32_t) * L, stream));
  }
  GPQ_TRY(arena.alloc(&partial, (v_.far_cap ? 3 : 1) * (size_t)partial_slots));
  hipLaunchKernelGGL(laq_init_kernel, dim3(ceil_div(L, 256)), dim3(256), 0, stream, v_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::init_calendar() {
  const uint32_t L = cfg.num_lps, B = cfg.calendar_windows;
  GPQ_CHECK_ARG(is_pow2(B) && B >= 8 && B <= kCalMaxWindows, "LAQ calendar: calendar_windows must be a power of "
                "two in [8, 256]");
  GPQ_CHECK_ARG(cfg.capacity_per_lp >= 1 && cfg.capacity_per_lp <= 64,
                "LAQ calendar: capacity_per_lp (a window's records per LP, on average over the LPs) must be in [1, 64]");
  GPQ_CHECK_ARG(cfg.far_per_lp >= 1 && cfg.far_per_lp <= 4096, "LAQ calendar: far_per_lp must be in [1, 4096]");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LAQ calendar: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG(std::isfinite(cfg.lookahead) && cfg.lookahead > 0.0f && std::isfinite(cfg.horizon),
                "LAQ calendar: needs a finite lookahead > 0 and a finite horizon");
  GPQ_CHECK_ARG(!cfg.update_ids, "LAQ calendar: no decrease-key mode");
  // one log per window and block of kCalBlockLps LPs (the commit block), each holding
  // capacity_per_lp records per LP of the block
  const uint32_t blocks = (uint32_t)ceil_div(L, kCalBlockLps);
  const uint64_t sub_cap = (uint64_t)kCalBlockLps * cfg.capacity_per_lp;
  GPQ_CHECK_ARG((uint64_t)blocks * sub_cap < (1ull << 32) && (uint64_t)L * cfg.far_per_lp < (1ull << 32),
                "LAQ calendar: num_lps too large for the window logs or the lists");
  static_assert(kCalBlockLps * 64u <= 65536u, "a block's log indexes fit 16 bits");
  v_.num_lps = L;
  v_.cap_log2 = 1;                    // no sorted run in this mode: a two-slot ring keeps the words valid
  v_.stage_cap = cfg.stage_per_lp;
  v_.cal_windows = B;
  v_.cal_blocks = blocks;
  v_.cal_sub_cap = (uint32_t)sub_cap;
  v_.cal_block_log2 = log2u(kCalBlockLps);
  v_.wlog_cap = blocks * v_.cal_sub_cap;
  v_.over_cap = L * cfg.far_per_lp;
  v_.parked_cap = L * cfg.far_per_lp;
  far_step_ = 0.0f;
  GPQ_TRY(arena.alloc(&v_.ring, (size_t)L << v_.cap_log2));
  GPQ_TRY(arena.alloc(&v_.stage, (size_t)L * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.rd, L));
  GPQ_TRY(arena.alloc(&v_.len, L));
  GPQ_TRY(arena.alloc(&v_.ins, L));
  GPQ_TRY(arena.alloc(&v_.head, L));
  GPQ_TRY(arena.alloc(&sum_, 1));
  // the logs' counters (16 words apart), then over_n[2] and parked_n, each on a 128-byte line of
  // its own: every block of a step may reserve a run on them, and a line that such atomics share
  // with words that blocks read (or with each other) slows every access to it
  const size_t lists_at = (((size_t)B * blocks << cal::kLogCountShift) + 31u) & ~(size_t)31u;
  const size_t counters = lists_at + 64u;
  GPQ_TRY(arena.alloc(&v_.wlog, (size_t)B * v_.wlog_cap));
  GPQ_TRY(arena.alloc(&v_.wlog_n, counters));
  v_.over_n = v_.wlog_n + lists_at;
  GPQ_TRY(arena.alloc(&v_.over, 2 * (size_t)v_.over_cap));
  GPQ_TRY(arena.alloc(&v_.parked, (size_t)v_.parked_cap));
  GPQ_TRY(arena.alloc(&v_.cal_st, 1));
  GPQ_TRY(arena.alloc(&gather_, kGatherCap + 1));
  v_.parked_n = v_.wlog_n + lists_at + 32u;
  GPQ_HIP(hipMemsetAsync(v_.wlog_n, 0, sizeof(uint32_t) * counters, stream));
  GPQ_HIP(hipMemsetAsync(v_.cal_st, 0, sizeof(CalState), stream));
  v_.ws = ws;
  v_.stats = stats;
  partial_slots = ceil_div(L, kLaqBlock);
  GPQ_TRY(arena.alloc(&partial, (size_t)partial_slots));
  hipLaunchKernelGGL(laq_init_kernel, dim3(ceil_div(L, 256)), dim3(256), 0, stream, v_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::calendar_start() {
  GPQ_CHECK_ARG(v_.wlog, "LAQ: calendar_start needs calendar mode");
  hipLaunchKernelGGL(cal_start_kernel, dim3(1), dim3(1), 0, stream, v_, cfg.lookahead, cfg.horizon);
  hipLaunchKernelGGL(cal_fill_kernel, dim3(cal::kGroupBlocks), dim3(cal::kBlock), 0, stream, v_, cfg.horizon);
  hipLaunchKernelGGL(cal_fill_done_kernel, dim3(1), dim3(1), 0, stream, v_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::post() {
  if (v_.wlog) {             // the calendar step's last block has moved the window already
    if (profile_mark) GPQ_HIP(hipEventRecord(profile_mark, stream));
    return GPQ_OK;
  }
  if (v_.stage_alt) {        // the model step merged and committed: only the window moves
    if (profile_mark) GPQ_HIP(hipEventRecord(profile_mark, stream));
    hipLaunchKernelGGL(laq_step_publish_kernel, dim3(1), dim3(1), 0, stream, ws, cfg.lookahead, cfg.horizon,
                       far_step_, 0u);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }
  return merge_publish();
}

gpq_status LaqQueue::reset_stage_select() {
  hipLaunchKernelGGL(laq_step_publish_kernel, dim3(1), dim3(1), 0, stream, ws, cfg.lookahead, cfg.horizon, far_step_,
                     1u);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::empty_all() {
  GPQ_CHECK_ARG(v_.in_queue, "LAQ: empty_all is for the decrease-key mode");
  const uint32_t words = (v_.update_ids + 31) / 32;
  const uint32_t n = std::max(cfg.num_lps, words);
  hipLaunchKernelGGL(laq_empty_kernel, dim3(std::min<uint32_t>(ceil_div(n, 256), 4096)), dim3(256), 0, stream, v_,
                     words);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::post_init() {
  if (!v_.wlog) return merge_publish();
  hipLaunchKernelGGL(cal_post_init_kernel, dim3(partial_slots), dim3(kLaqBlock), 0, stream, v_, partial);
  hipLaunchKernelGGL(cal_init_publish_kernel, dim3(1), dim3(1024), 0, stream, partial, partial_slots, ws,
                     cfg.lookahead, cfg.horizon);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::merge_publish() {
  GPQ_TRY(merge_only());
  if (v_.in_queue) {
    hipLaunchKernelGGL(laq_update_kernel, dim3(ceil_div(cfg.num_lps, 256)), dim3(256), 0, stream, v_, partial);
    GPQ_HIP(hipGetLastError());
  }
  if (!v_.far_cap) return publish_window();
  if (profile_mark) GPQ_HIP(hipEventRecord(profile_mark, stream));
  hipLaunchKernelGGL(laq_far_publish_kernel, dim3(1), dim3(1024), 0, stream, partial, partial_slots, ws, cfg.lookahead,
                     cfg.horizon, far_step_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::merge_only() {
  static_assert(kLaqBlock == (int)LaqQueue::kBlockLps, "LAQ post block size");
  const dim3 grid(partial_slots), block(kLaqBlock);
  if (v_.near_lp) {
    if (group_ == 8)
      hipLaunchKernelGGL((laq_post_open_kernel<8, kSlots>), grid, block, 0, stream, v_, partial);
    else
      hipLaunchKernelGGL((laq_post_open_kernel<16, kSlots>), grid, block, 0, stream, v_, partial);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }
  if (v_.far_cap) {
    if (group_ == 8)
      hipLaunchKernelGGL((laq_post_far_kernel<8, kSlots>), grid, block, 0, stream, v_, partial);
    else
      hipLaunchKernelGGL((laq_post_far_kernel<16, kSlots>), grid, block, 0, stream, v_, partial);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }
  if (group_ == 8)
    hipLaunchKernelGGL((laq_post_kernel<8, kSlots>), grid, block, 0, stream, v_, partial);
  else if (group_ == 16)
    hipLaunchKernelGGL((laq_post_kernel<16, kSlots>), grid, block, 0, stream, v_, partial);
  else if (group_ == 32)
    hipLaunchKernelGGL((laq_post_kernel<32, kSlots>), grid, block, 0, stream, v_, partial);
  else
    hipLaunchKernelGGL((laq_post_kernel<64, kSlots>), grid, block, 0, stream, v_, partial);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::enable_alt_stage() {
  if (v_.stage_alt) return GPQ_OK;
  GPQ_TRY(arena.alloc(&v_.stage_alt, (size_t)cfg.num_lps * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.ins_alt, cfg.num_lps));
  GPQ_HIP(hipMemsetAsync(v_.ins_alt, 0, sizeof(uint32_t) * cfg.num_lps, stream));
  return GPQ_OK;
}

gpq_status LaqQueue::pending(uint64_t* n) {
  if (v_.wlog) {
    GPQ_HIP(hipMemsetAsync(sum_, 0, sizeof(uint64_t), stream));
    hipLaunchKernelGGL(cal_count_kernel, dim3(ceil_div(cfg.num_lps, 256)), dim3(256), 0, stream, v_,
                       reinterpret_cast<unsigned long long*>(sum_));
    GPQ_HIP(hipGetLastError());
    GPQ_HIP(hipMemcpyAsync(n, sum_, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    GPQ_HIP(hipStreamSynchronize(stream));
    return GPQ_OK;
  }
  GPQ_TRY(sum_u32(v_.len, cfg.num_lps, sum_, n));
  uint64_t x = 0;
  if (v_.far_cap) {
    GPQ_TRY(sum_u32(v_.nfar, cfg.num_lps, sum_, &x));
    *n += x;
  }
  if (v_.ins_alt) {          // a merge-and-commit step leaves the events it scheduled staged
    GPQ_TRY(sum_u32(v_.ins, cfg.num_lps, sum_, &x));
    *n += x;
    GPQ_TRY(sum_u32(v_.ins_alt, cfg.num_lps, sum_, &x));
    *n += x;
  }
  return GPQ_OK;
}

gpq_status LaqQueue::calendar_gather(uint32_t lp, std::vector<Event>* out) {
  uint32_t* d_n = reinterpret_cast<uint32_t*>(gather_ + kGatherCap);
  GPQ_HIP(hipMemsetAsync(d_n, 0, sizeof(uint32_t), stream));
  hipLaunchKernelGGL(cal_gather_kernel, dim3(1024), dim3(256), 0, stream, v_, lp, gather_, kGatherCap, d_n);
  GPQ_HIP(hipGetLastError());
  uint32_t n = 0;
  GPQ_HIP(hipMemcpyAsync(&n, d_n, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  if (n > kGatherCap) {
    set_error("LAQ calendar: LP " + std::to_string(lp) + " holds more than " + std::to_string(kGatherCap) +
              " records");
    return GPQ_ERR_CAPACITY;
  }
  out->resize(n);
  if (n) {
    GPQ_HIP(hipMemcpyAsync(out->data(), gather_, sizeof(Event) * n, hipMemcpyDeviceToHost, stream));
    GPQ_HIP(hipStreamSynchronize(stream));
  }
  return GPQ_OK;
}

gpq_status LaqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  if (v_.wlog) {
    std::vector<Event> all;
    GPQ_TRY(calendar_gather(lp, &all));
    *n = (uint32_t)all.size();
    return GPQ_OK;
  }
  uint32_t f = 0;
  GPQ_HIP(hipMemcpyAsync(n, v_.len + lp, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  if (v_.far_cap) GPQ_HIP(hipMemcpyAsync(&f, v_.nfar + lp, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *n += f;
  return GPQ_OK;
}

gpq_status LaqQueue::clear_lp(uint32_t lp) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  if (v_.wlog) {
    set_error("LAQ calendar mode: per-LP clear is not provided");
    return GPQ_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(laq_clear_kernel, dim3(1), dim3(1), 0, stream, v_, lp);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  if (v_.wlog) {
    std::vector<Event> all;
    GPQ_TRY(calendar_gather(lp, &all));
    std::sort(all.begin(), all.end(), [](const Event& a, const Event& b) {
      if (a.ts != b.ts) return a.ts < b.ts;
      if (a.src != b.src) return a.src < b.src;
      return a.seq < b.seq;
    });
    *n = (uint32_t)all.size();
    for (uint32_t i = 0; i < *n && i < cap; ++i) out[i] = all[i];
    return GPQ_OK;
  }
  uint32_t rd = 0, len = 0, nf = 0;
  const uint32_t C = 1u << v_.cap_log2;
  std::vector<Event> ring(C), far(v_.far_cap);
  GPQ_HIP(hipMemcpyAsync(&rd, v_.rd + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(&len, v_.len + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(ring.data(), v_.ring + ((size_t)lp << v_.cap_log2), sizeof(Event) * C,
                         hipMemcpyDeviceToHost, stream));
  if (v_.far_cap) {
    GPQ_HIP(hipMemcpyAsync(&nf, v_.nfar + lp, 4, hipMemcpyDeviceToHost, stream));
    GPQ_HIP(hipMemcpyAsync(far.data(), v_.far + (size_t)lp * v_.far_cap, sizeof(Event) * v_.far_cap,
                           hipMemcpyDeviceToHost, stream));
  }
  uint32_t ni[2] = {0, 0};
  std::vector<Event> stg[2];
  if (v_.ins_alt) {          // staged events of a merge-and-commit step, in either buffer
    const Event* bufs[2] = {v_.stage, v_.stage_alt};
    const uint32_t* cnts[2] = {v_.ins, v_.ins_alt};
    for (int b = 0; b < 2; ++b) {
      stg[b].resize(v_.stage_cap);
      GPQ_HIP(hipMemcpyAsync(&ni[b], cnts[b] + lp, 4, hipMemcpyDeviceToHost, stream));
      GPQ_HIP(hipMemcpyAsync(stg[b].data(), bufs[b] + (size_t)lp * v_.stage_cap, sizeof(Event) * v_.stage_cap,
                             hipMemcpyDeviceToHost, stream));
    }
  }
  GPQ_HIP(hipStreamSynchronize(stream));
  std::vector<Event> all;
  for (uint32_t i = 0; i < len; ++i) all.push_back(ring[(rd + i) & (C - 1)]);
  std::vector<Event> rest(far.begin(), far.begin() + std::min(nf, v_.far_cap));   // far store: unsorted
  for (int b = 0; b < 2; ++b)
    rest.insert(rest.end(), stg[b].begin(), stg[b].begin() + std::min(ni[b], v_.stage_cap));
  all.insert(all.end(), rest.begin(), rest.end());
  std::sort(all.begin(), all.end(), [](const Event& a, const Event& b) {   // the ring part is sorted already
    if (a.ts != b.ts) return a.ts < b.ts;
    if (a.src != b.src) return a.src < b.src;
    return a.seq < b.seq;
  });
  *n = (uint32_t)all.size();
  for (uint32_t i = 0; i < *n && i < cap; ++i) out[i] = all[i];
  return GPQ_OK;
}

}  // namespace gpq

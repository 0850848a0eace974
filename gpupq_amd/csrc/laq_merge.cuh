// laq_merge.cuh -- device building blocks of the local array queue's merge, shared by the
// queue's post kernels (queue_laq.hip) and the PHOLD step that merges and commits in one pass
// (phold.hip).  Reference: src/local_array_queue.cu (the per-LP insertion sort :254-299,
// src/sort.cu:329-380) and the heap-backed variant's near/far split (:515-681).
#pragma once

#include "../../include/gpupq_device/queue_laq.cuh"
#include "gpq_device.h"

namespace gpq {
namespace laq {

constexpr int kSlots = 8;                  // default ring records held per lane: a group of G lanes covers 8G records
constexpr uint32_t kFarList = 64;          // two-tier: near events one LP may take in one pass (merge input)

struct Key3 {
  uint32_t ts, src, seq;
};

__device__ __forceinline__ Key3 shfl_key(const Event& e, int src_lane) {
  return Key3{(uint32_t)__shfl((int)e.ts, src_lane, kWave), (uint32_t)__shfl((int)e.src, src_lane, kWave),
              (uint32_t)__shfl((int)e.seq, src_lane, kWave)};
}

__device__ __forceinline__ bool key_less3(const Key3& a, const Event& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

__device__ __forceinline__ bool key_less3(const Key3& a, const Key3& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

// Staged event i of LP lp: from the LP's staging slots in HBM, or from a key-only list in LDS
// (records staged for LP lp all carry dst = lp).
__device__ __forceinline__ Event staged_at(const Event* stg, uint32_t i, uint32_t) { return load_event(&stg[i]); }
__device__ __forceinline__ Event staged_at(const Key3* stg, uint32_t i, uint32_t lp) {
  const Key3 k = stg[i];
  return Event{k.ts, lp, k.src, k.seq};
}

// Orders the LDS histogram updates of the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads the run of n records of LP lp at ring slot r for laq_merge_pass<..., kPre = true>:
// record j into lane j % G of the group, register slot j / G (key words only).
template <int G, int SL>
__device__ __forceinline__ void load_run(const LaqView& q, uint32_t lp, uint32_t n, uint32_t r, int g,
                                         Key3* old) {
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  const Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  const uint32_t ns = (n + G - 1) / G;
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    const uint32_t j = (uint32_t)(s * G + g);
    old[s] = Key3{0u, 0u, 0u};
    if ((uint32_t)s < ns && j < n) {
      const Event o = load_event(&ring[(r + j) & mask]);
      old[s] = Key3{o.ts, o.src, o.seq};
    }
  }
}

// One merge pass: the k <= G - 1 staged events stg[0 .. k) of LP lp go into its sorted run of
// n records at ring slot r, by a group of G lanes (lane g of the group, group starting at wave
// lane gbase).  The run is loaded at once (record j in lane j % G, register slot j / G), one
// staged event per lane.  Old record j ends at final index j + c_j, c_j = number of new
// events with a smaller key, and c is nondecreasing in j; the new event of rank a lands at
// a + #{j : c_j <= a}.  A histogram of c over [0, k] (LDS, one bin per lane) gives both: its
// inclusive prefix places the new events, and bin a counts the old records that stay put
// when the run grows by a slots at the front and k - a at the back.  The a with the largest
// bin is taken, so the fewest records move: a = 0 is the reference's backward insertion
// (src/sort.cu:329-356), a = k its forward variant (src/sort.cu:358-379), chosen here per LP
// and pass.  Updates n and r and returns the records moved.
//
// kDrain: the records below wend (final indices [0, d), d returned in *d_out) are taken out
// of the run: it then starts at index d (n and r count from there), the head word is set to
// the record at index d, and the taken records stay in the ring slots just before the new
// start for the caller to commit.  Without kDrain the head changes only when a new event
// lands at index 0.
//
// The pass exists twice with one body (laq_merge_pass.inc): laq_merge_pass, an ordinary function
// (the post kernels, whose registers stay within the five-blocks-per-CU bound this way), and
// laq_merge_pass_inl, force-inlined (the merge-and-commit kernel, measured faster this way).
//
// kPre: the run is not loaded here but taken from *preloaded, filled by load_run (with the same n
// and r) before the caller's other loads, so that their latencies overlap.
template <int SL>
struct Run {
  Key3 k[SL];
};

template <int G, int SL, bool kDrain, typename Stg, bool kPre = false>
__device__ uint32_t laq_merge_pass(const LaqView& q, uint32_t lp, const Stg* stg, uint32_t k,
                                    uint32_t& n, uint32_t& r, int g, int gbase, uint32_t* hist,
                                    uint32_t wend, uint32_t* d_out, const Run<SL>* preloaded = nullptr) {
#include "laq_merge_pass.inc"
}

template <int G, int SL, bool kDrain, typename Stg, bool kPre = false>
__device__ __forceinline__ uint32_t laq_merge_pass_inl(const LaqView& q, uint32_t lp, const Stg* stg, uint32_t k,
                                                       uint32_t& n, uint32_t& r, int g, int gbase, uint32_t* hist,
                                                       uint32_t wend, uint32_t* d_out,
                                                       const Run<SL>* preloaded = nullptr) {
#include "laq_merge_pass.inc"
}

// Merges the k_total staged events of one LP into its sorted run of n records at ring slot r,
// G - 1 per pass, and writes the LP's ring indices back.  Returns the records moved.
template <int G, int SL, typename Stg>
__device__ uint32_t laq_merge_group(const LaqView& q, uint32_t lp, const Stg* stg, uint32_t k_total, uint32_t n,
                                    uint32_t r, int g, int gbase, uint32_t* hist) {
  uint32_t moved = 0;
  for (uint32_t done = 0; done < k_total;) {
    if (done) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");   // previous pass's ring writes
    const uint32_t k = min(k_total - done, (uint32_t)G - 1u);
    moved += laq_merge_pass<G, SL, false>(q, lp, stg + done, k, n, r, g, gbase, hist, 0u, nullptr);
    done += k;
  }
  if (g == 0) {
    q.rd[lp] = r;
    q.len[lp] = n;
    q.ins[lp] = 0;
  }
  return moved;
}

// Two-tier split of the candidates of LP lp -- its far records first when the near horizon
// moves, then its k staged records -- at the horizon N, by a group of G lanes.  Keys below N
// are listed in `near` (LDS, every record of LP lp has dst = lp); the others are compacted
// into the far store, in place after the far records kept untouched (all of them when the
// horizon does not move, none when it does).  Returns the number of near keys; *n_far and
// *fmin get the far count and the smallest far ts bits (+inf when empty).
// split_candidates_by takes the test for "below the horizon" as a predicate on the record.
template <int G, class BelowFn>
__device__ __forceinline__ uint32_t split_candidates_by(const LaqView& q, uint32_t lp, const Event* stg, uint32_t k,
                                                        uint32_t f, bool moving, BelowFn below_horizon, Key3* near,
                                                        int g, uint64_t gmask, uint32_t* n_far_out,
                                                        uint32_t* fmin_out) {
  const int lane = lane_id();
  const uint32_t cands = (moving ? f : 0u) + k;
  Event* far = q.far + (size_t)lp * q.far_cap;
  uint32_t n_near = 0, n_far = moving ? 0u : f;
  uint32_t fmin = moving ? kInfBits : q.far_min[lp];
  // kChunk candidates per lane are loaded before any is placed, so their latencies overlap.
  // In-place compaction stays safe: a far record is written to an index no larger than the
  // candidate index being placed, and every candidate of the chunk is already loaded.
  constexpr uint32_t kChunk = 4;
  const uint32_t nf = moving ? f : 0u;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t base = 0; base < cands; base += kChunk * G) {
    Event ev[kChunk];
#pragma unroll
    for (uint32_t u = 0; u < kChunk; ++u) {
      const uint32_t c = base + u * G + (uint32_t)g;
      ev[u] = Event{kInfBits, 0u, 0u, 0u};
      if (c < cands) ev[u] = load_event(c < nf ? &far[c] : &stg[c - nf]);
    }
#pragma unroll
    for (uint32_t u = 0; u < kChunk; ++u) {
      if (base + u * G >= cands) break;       // uniform across the group
      const Event& e = ev[u];
      const bool valid = base + u * G + (uint32_t)g < cands;
      const bool is_near = valid && below_horizon(e);
      const bool is_far = valid && !is_near;
      const uint64_t bn = __ballot(is_near) & gmask, bf = __ballot(is_far) & gmask;
      const uint32_t pn = n_near + (uint32_t)__popcll(bn & below);
      const uint32_t pf = n_far + (uint32_t)__popcll(bf & below);
      if (is_near) {
        if (pn < kFarList) near[pn] = Key3{e.ts, e.src, e.seq};
        else atomicOr(&q.ws->error, kErrQueueOverflow);
      }
      if (is_far) {
        if (pf < q.far_cap) store_event(&far[pf], e);
        else atomicOr(&q.ws->error, kErrQueueOverflow);
        fmin = e.ts < fmin ? e.ts : fmin;
      }
      n_near += (uint32_t)__popcll(bn);
      n_far += (uint32_t)__popcll(bf);
    }
  }
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)fmin, off, kWave);
    fmin = o < fmin ? o : fmin;
  }
  *n_far_out = n_far;
  *fmin_out = fmin;
  return n_near;
}

template <int G>
__device__ __forceinline__ uint32_t split_candidates(const LaqView& q, uint32_t lp, const Event* stg, uint32_t k,
                                                     uint32_t f, bool moving, uint32_t N, Key3* near, int g,
                                                     uint64_t gmask, uint32_t* n_far_out, uint32_t* fmin_out) {
  return split_candidates_by<G>(q, lp, stg, k, f, moving, [N](const Event& e) { return e.ts < N; }, near, g, gmask,
                                n_far_out, fmin_out);
}

}  // namespace laq
}  // namespace gpq

// laq_calendar.cuh -- calendar mode of the local array queue and the PHOLD step on it.
//
// The merge-and-commit PHOLD step (drain mode 2) ends window k at w_k = min(w_{k-1} + lookahead,
// end time) in float arithmetic, so every window end is known as soon as the first one is.  In
// calendar mode the queue uses that: it keeps one append log per upcoming window (a calendar
// queue, R. Brown, CACM 31(10), 1988, with one day per window, shared by all LPs), and an event
// is appended to the log of the window it falls in.  Step c is one launch with two kinds of
// blocks that run side by side:
//   * commit blocks: per LP, the staged events of window c in key order (sorted in registers),
//     their successors appended to the logs of their windows, aggregated per block and window
//     so that every log grows by coalesced runs; a successor for window c + 1 goes straight to
//     that window's staging area;
//   * group blocks: the log of window c + 1, grouped by LP into the other staging area (the
//     staging area and its counters are small enough to stay in the device's last-level cache
//     while they are written in random order).
// No sorted run is kept, merged or shifted.
//
// Layout (B = cal_windows, a power of two; L = num_lps):
//   wlog   [B][wlog_cap] Event   window x's log is wlog[x mod B], count wlog_n[x mod B]
//   over   [2][over_cap] Event   events beyond the calendar: appended to over[(x + 1) & 1] by the
//                                steps that stage window x and read when window x + 1 is staged
//   parked [parked_cap]  Event   events at or past the end time: never committed, only kept
//   stage / stage_alt [S][L]     staging areas of windows with even / odd index, slot-major
//   ins / ins_alt [L]            their counts
//   cal_st                       `cur` (the window the next step commits) and the ends of windows
//                                x in [cur, cur + B] at w[x mod 2B]
//
// Reference: the per-LP queue of src/local_array_queue.cu and its heap-backed far tier
// (:515-681); the step is handle_next_event (src/phold.cu:81-117) for every event of a window.
#pragma once

#include "phold_fused.cuh"

namespace gpq {
namespace cal {

// LPs per commit block.  Every commit block adds its run to the counter of each window it
// files successors for (about 40 of them per step), so fewer, larger blocks mean less
// contention on those counters: 256 LPs 1.12e10, 512 LPs 1.25-1.28e10, 1024 LPs 1.25-1.28e10
// events/s at 2^20 LPs (same box, the staging threads kept at 98 K).
constexpr int kBlock = 512;
constexpr int kRegs = 8;          // staged records one thread sorts in registers (fast path)
constexpr int kWaveCand = 64;     // staged records of one LP on the slow path (one per lane)
constexpr int kGroupBlocks = 2048;        // blocks of the standalone filing / staging launches
constexpr uint32_t kStepGroupBlocks = 192;   // staging blocks of a step (beside the commit blocks; with 256-thread
                                             // blocks 384 measured best of 128, 256, 384, 512, 1024, 2048)

struct Key {
  uint32_t ts, src, seq;
};

__device__ __forceinline__ bool key_lt(const Key& a, const Key& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

// Compare-exchange of keys a and b (ascending) held as three word arrays indexed by constants,
// so that the sort network stays in registers (a network over an array of Key structs is
// placed in scratch memory by the compiler).
template <int A, int B>
__device__ __forceinline__ void cswap3(uint32_t* t, uint32_t* s, uint32_t* q) {
  const bool lt = t[B] != t[A] ? t[B] < t[A] : (s[B] != s[A] ? s[B] < s[A] : q[B] < q[A]);
  const uint32_t t0 = lt ? t[B] : t[A], t1 = lt ? t[A] : t[B];
  const uint32_t s0 = lt ? s[B] : s[A], s1 = lt ? s[A] : s[B];
  const uint32_t q0 = lt ? q[B] : q[A], q1 = lt ? q[A] : q[B];
  t[A] = t0; t[B] = t1;
  s[A] = s0; s[B] = s1;
  q[A] = q0; q[B] = q1;
}

// Batcher's odd-even merge sort of eight keys (19 comparators).
__device__ __forceinline__ void sort8(uint32_t* t, uint32_t* s, uint32_t* q) {
  cswap3<0, 1>(t, s, q); cswap3<2, 3>(t, s, q); cswap3<4, 5>(t, s, q); cswap3<6, 7>(t, s, q);
  cswap3<0, 2>(t, s, q); cswap3<1, 3>(t, s, q); cswap3<4, 6>(t, s, q); cswap3<5, 7>(t, s, q);
  cswap3<1, 2>(t, s, q); cswap3<5, 6>(t, s, q);
  cswap3<0, 4>(t, s, q); cswap3<1, 5>(t, s, q); cswap3<2, 6>(t, s, q); cswap3<3, 7>(t, s, q);
  cswap3<2, 4>(t, s, q); cswap3<3, 5>(t, s, q);
  cswap3<1, 2>(t, s, q); cswap3<3, 4>(t, s, q); cswap3<5, 6>(t, s, q);
}

// Window offset of timestamp bits ts: the smallest o in [o_min, B) with ts < sw[o], where
// sw[o] is the end of window cur + o; B when ts >= sw[B - 1] (beyond the calendar).
__device__ __forceinline__ uint32_t find_window(const uint32_t* sw, uint32_t B, uint32_t ts, uint32_t o_min) {
  if (!(ts < sw[B - 1])) return B;
  uint32_t lo = o_min, hi = B - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ts < sw[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Where an event goes, seen from window x (the window being committed, or staged): a window
// offset o < B (the log of window x + o), kOver (beyond the calendar) or kParked (at or past the
// end time).
struct Route {
  static constexpr uint32_t kOver = kCalMaxWindows;
  static constexpr uint32_t kParked = kCalMaxWindows + 1;
  static constexpr uint32_t kNone = kCalMaxWindows + 2;   // no record
};

__device__ __forceinline__ uint32_t route_of(const uint32_t* sw, uint32_t B, uint32_t ts, uint32_t end_bits,
                                             uint32_t o_min) {
  if (!(ts < end_bits)) return Route::kParked;
  const uint32_t o = find_window(sw, B, ts, o_min);
  return o < B ? o : Route::kOver;
}

// The counter and record array of route r seen from window x; the beyond-the-calendar list is
// over[op].
__device__ __forceinline__ uint32_t* route_count(const LaqView& q, uint32_t x, uint32_t op, uint32_t r) {
  if (r == Route::kParked) return q.parked_n;
  if (r == Route::kOver) return &q.over_n[op];
  return &q.wlog_n[(x + r) & (q.cal_windows - 1u)];
}

__device__ __forceinline__ Event* route_slot(const LaqView& q, uint32_t x, uint32_t op, uint32_t r, uint32_t pos,
                                             bool* ok) {
  if (r == Route::kParked) {
    *ok = pos < q.parked_cap;
    return q.parked + pos;
  }
  if (r == Route::kOver) {
    *ok = pos < q.over_cap;
    return q.over + (size_t)op * q.over_cap + pos;
  }
  *ok = pos < q.wlog_cap;
  return q.wlog + (size_t)((x + r) & (q.cal_windows - 1u)) * q.wlog_cap + pos;
}

// One event appended by its own atomic (the rare paths).
__device__ __forceinline__ void append_one(const LaqView& q, uint32_t x, uint32_t op, uint32_t r, const Event& e) {
  const uint32_t pos = atomicAdd(route_count(q, x, op, r), 1u);
  bool ok;
  Event* p = route_slot(q, x, op, r, pos, &ok);
  if (ok) store_event(p, e);
  else atomicOr(&q.ws->error, kErrQueueOverflow);
}

__device__ __forceinline__ uint32_t route_index(uint32_t r, uint32_t B) { return r < B ? r : B + (r - kCalMaxWindows); }
__device__ __forceinline__ uint32_t index_route(uint32_t i, uint32_t B) { return i < B ? i : kCalMaxWindows + (i - B); }

// Block-aggregated append: every thread of the block calls with at most one record (route
// kNone: none); each takes a rank in the block's LDS count of its route, one thread per used
// route reserves the block's run with a single atomic, and the records are written into the
// runs.  x and op are block-uniform.
struct BlockAppend {
  uint32_t cnt[kCalMaxWindows + 2];
  uint32_t base[kCalMaxWindows + 2];
};

__device__ __forceinline__ void block_append(const LaqView& q, uint32_t x, uint32_t op, uint32_t r, const Event& e,
                                             BlockAppend& ba) {
  const uint32_t B = q.cal_windows;
  for (uint32_t i = threadIdx.x; i < B + 2u; i += blockDim.x) ba.cnt[i] = 0u;
  __syncthreads();
  uint32_t rank = 0;
  if (r != Route::kNone) rank = atomicAdd(&ba.cnt[route_index(r, B)], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < B + 2u; i += blockDim.x) {
    const uint32_t n = ba.cnt[i];
    ba.base[i] = n ? atomicAdd(route_count(q, x, op, index_route(i, B)), n) : 0u;
  }
  __syncthreads();
  if (r != Route::kNone) {
    bool ok;
    Event* p = route_slot(q, x, op, r, ba.base[route_index(r, B)] + rank, &ok);
    if (ok) store_event(p, e);
    else atomicOr(&q.ws->error, kErrQueueOverflow);
  }
  __syncthreads();
}

// The staging area of window x (slot-major: slot s of LP lp at s * num_lps + lp, so that the
// grouping pass writes into few, densely filled lines -- slots 0 .. 3 of all LPs -- and the
// LPs of a wave read slot s at consecutive addresses; while the initial events are populated
// the even area is LP-major, as for every other LAQ mode).
struct Stage {
  Event* rec;
  uint32_t* n;
};

__device__ __forceinline__ Stage stage_of(const LaqView& q, uint32_t x) {
  return (x & 1u) ? Stage{q.stage_alt, q.ins_alt} : Stage{q.stage, q.ins};
}

__device__ __forceinline__ size_t gslot(const LaqView& q, uint32_t slot, uint32_t lp) {
  return (size_t)slot * q.num_lps + lp;
}

// queue_insert into a staging area: e goes to its LP's next slot.
__device__ __forceinline__ void stage_one(const LaqView& q, const Stage& st, const Event& e) {
  const uint32_t slot = atomicAdd(&st.n[e.dst], 1u);
  if (slot < q.stage_cap) store_event(&st.rec[gslot(q, slot, e.dst)], e);
  else atomicOr(&q.ws->error, kErrStageOverflow);
}

// Block-wide: the ends of windows x .. x + B - 1 into sw (every thread of the block calls).
__device__ __forceinline__ void load_windows(const LaqView& q, uint32_t x, uint32_t* sw) {
  const uint32_t B = q.cal_windows, M = 2u * B - 1u;
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) sw[i] = q.cal_st->w[(x + i) & M];
}

// The successor of a committed event (src/phold.cu:93-116): draw {n, local, 0, 0}, destination
// self with probability p_self else uniform over the instance, ts' = (ts + lookahead) + Exp(1).
__device__ __forceinline__ Event successor(const FusedModel& m, uint32_t lp, uint32_t n, uint32_t ts) {
  const uint32_t local = lp - m.base;
  const uint4 r = philox4x32_10(make_uint4(n, local, 0u, 0u), make_uint2(m.seed, 0u));
  const uint32_t dst = r.y < m.p_self ? lp : m.base + (uint32_t)(((uint64_t)r.z * m.lps) >> 32);
  const float t_new = (bits_to_f(ts) + m.lookahead) + exp_from_u32(r.x);
  return Event{f_to_bits(t_new), dst, local, m.init_per_lp + n};
}

// Slow path, one wave for one LP with more than kRegs staged events: lane i holds event i and
// its rank among them; the successors are filed by their own atomics.
__device__ __forceinline__ void slow_lp(const LaqView& q, const FusedModel& m, const uint32_t* sw, uint32_t c,
                                        uint32_t end_bits, uint32_t lp, Key* sorted) {
  const int lane = lane_id();
  const Stage st = stage_of(q, c), next = stage_of(q, c + 1u);
  uint32_t k = st.n[lp];
  if (k > q.stage_cap) k = q.stage_cap;
  if (k > (uint32_t)kWaveCand) {
    if (lane == 0) atomicOr(&q.ws->error, kErrStageOverflow);
    k = kWaveCand;
  }
  const bool has = (uint32_t)lane < k;
  const Event e = has ? load_event(&st.rec[gslot(q, (uint32_t)lane, lp)]) : Event{kInfBits, 0u, 0u, 0u};
  const Key me{e.ts, e.src, e.seq};
  uint32_t rank = 0;
  for (uint32_t j = 0; j < k; ++j) {
    const Key o{(uint32_t)__shfl((int)e.ts, (int)j, kWave), (uint32_t)__shfl((int)e.src, (int)j, kWave),
                (uint32_t)__shfl((int)e.seq, (int)j, kWave)};
    rank += key_lt(o, me) ? 1u : 0u;
  }
  const uint32_t cnt = m.count[lp];
  if (has) {
    sorted[rank] = me;
    const Event ne = successor(m, lp, cnt + rank, e.ts);
    const uint32_t r = route_of(sw, q.cal_windows, ne.ts, end_bits, 1u);
    if (r == 1u) stage_one(q, next, ne);
    else append_one(q, c, c & 1u, r, ne);
  }
  laq::wave_lds_sync();
  if (lane == 0) {
    uint64_t h = m.hash[lp];
    uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
    for (uint32_t r = 0; r < k; ++r) {
      const Key x = sorted[r];
      h = trace_step(h, x.ts, x.src, x.seq);
      if (m.last_ts) {
        if (x.ts < last) atomicOr(&q.ws->error, kErrBadState);
        last = x.ts;
      }
    }
    m.count[lp] = cnt + k;
    m.hash[lp] = h;
    if (m.last_ts) m.last_ts[lp] = last;
    st.n[lp] = 0;
  }
  laq::wave_lds_sync();
}

// Commit blocks of step c (window c), block `blk` of the LPs.  Fast path, one thread per LP with
// at most kRegs staged events: sorted in registers and committed in order; a successor for
// window c + 1 goes straight to that window's staging area, every other one takes a rank in the
// block's LDS count of its route, one thread per route reserves the block's run with a single
// atomic, and the successors are written into the runs.  LPs with more staged events are handled
// afterwards, one wave each (slow_lp).
__device__ __forceinline__ void commit_block(const LaqView& q, const FusedModel& m, float end_time, uint32_t c,
                                             uint32_t blk) {
  __shared__ uint32_t sw[kCalMaxWindows];
  __shared__ uint32_t s_cnt[kCalMaxWindows + 2], s_base[kCalMaxWindows + 2];
  __shared__ uint32_t s_slow[kBlock];
  __shared__ uint32_t s_nslow;
  __shared__ Key s_sorted[kBlock / kWave][kWaveCand];
  load_windows(q, c, sw);
  const uint32_t B = q.cal_windows;
  for (uint32_t i = threadIdx.x; i < B + 2u; i += blockDim.x) s_cnt[i] = 0u;
  if (threadIdx.x == 0) s_nslow = 0;
  __syncthreads();
  const Stage st = stage_of(q, c), next = stage_of(q, c + 1u);
  const uint32_t end_bits = f_to_bits(end_time);
  const uint32_t lp = blk * kBlock + threadIdx.x;
  uint4 ne[kRegs];      // successors as uint4: an array of the 16-byte-aligned Event goes to scratch memory
  uint32_t route[kRegs], rank[kRegs];
  uint32_t cn = 0;
  if (lp < q.num_lps && lp - m.base < m.lps) {
    const uint32_t k = st.n[lp];
    if (k > (uint32_t)kRegs) {
      s_slow[atomicAdd(&s_nslow, 1u)] = lp;
    } else if (k) {
      cn = k;
      uint32_t kt[kRegs], ks[kRegs], kq[kRegs];
#pragma unroll
      for (int s = 0; s < kRegs; ++s) {
        if ((uint32_t)s < cn) {
          const Event e = load_event(&st.rec[gslot(q, (uint32_t)s, lp)]);
          kt[s] = e.ts;
          ks[s] = e.src;
          kq[s] = e.seq;
        } else {
          kt[s] = ks[s] = kq[s] = 0xFFFFFFFFu;
        }
      }
      const uint32_t cnt = m.count[lp];
      uint64_t h = m.hash[lp];
      uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
      sort8(kt, ks, kq);
#pragma unroll
      for (int j = 0; j < kRegs; ++j) {
        if ((uint32_t)j < cn) {
          h = trace_step(h, kt[j], ks[j], kq[j]);
          if (m.last_ts) {
            if (kt[j] < last) atomicOr(&q.ws->error, kErrBadState);
            last = kt[j];
          }
          const Event sc = successor(m, lp, cnt + (uint32_t)j, kt[j]);
          ne[j] = make_uint4(sc.ts, sc.dst, sc.src, sc.seq);
          route[j] = route_of(sw, B, sc.ts, end_bits, 1u);
          if (route[j] == 1u) {                 // window c + 1: straight to its staging area
            stage_one(q, next, sc);
            route[j] = Route::kNone;
          } else {
            rank[j] = atomicAdd(&s_cnt[route_index(route[j], B)], 1u);
          }
        }
      }
      m.count[lp] = cnt + cn;
      m.hash[lp] = h;
      if (m.last_ts) m.last_ts[lp] = last;
      st.n[lp] = 0;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < B + 2u; i += blockDim.x) {
    const uint32_t n = s_cnt[i];
    s_base[i] = n ? atomicAdd(route_count(q, c, c & 1u, index_route(i, B)), n) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRegs; ++j) {
    if ((uint32_t)j < cn && route[j] != Route::kNone) {
      bool ok;
      Event* p = route_slot(q, c, c & 1u, route[j], s_base[route_index(route[j], B)] + rank[j], &ok);
      if (ok) *reinterpret_cast<uint4*>(p) = ne[j];
      else atomicOr(&q.ws->error, kErrQueueOverflow);
    }
  }
  const uint32_t w = threadIdx.x / kWave;
  for (uint32_t i = w; i < s_nslow; i += kBlock / kWave) slow_lp(q, m, sw, c, end_bits, s_slow[i], s_sorted[w]);
}

// Group blocks staging window x (block blk of nblk): every record of its log goes to the
// staging area of its LP; every record of the list over[x & 1] goes to that staging area
// (window x), to the log of its window, on to over[(x + 1) & 1], or with the parked ones.
__device__ __forceinline__ void group_block(const LaqView& q, float end_time, uint32_t x, uint32_t blk,
                                            uint32_t nblk) {
  __shared__ uint32_t sw[kCalMaxWindows];
  __shared__ BlockAppend ba;
  load_windows(q, x, sw);
  __syncthreads();
  const uint32_t B = q.cal_windows;
  const Stage st = stage_of(q, x);
  const uint32_t n = q.wlog_n[x & (B - 1u)];
  const uint32_t nl = n < q.wlog_cap ? n : q.wlog_cap;
  const Event* lg = q.wlog + (size_t)(x & (B - 1u)) * q.wlog_cap;
  const uint32_t stride = nblk * blockDim.x;
  // One record per thread and round.  Two or more in flight per thread (loads, then slot
  // atomics, then stores) measured slower: 1.02e10 (four) and 1.05e10 (eight) against 1.08e10
  // events/s; the pass is bound by the rate of the random atomics and stores, not their latency.
  for (uint32_t i = blk * blockDim.x + threadIdx.x; i < nl; i += stride) stage_one(q, st, load_event(&lg[i]));
  const uint32_t rp = x & 1u, wp = rp ^ 1u;
  const uint32_t no0 = q.over_n[rp];
  const uint32_t no = no0 < q.over_cap ? no0 : q.over_cap;
  const Event* ov = q.over + (size_t)rp * q.over_cap;
  const uint32_t end_bits = f_to_bits(end_time);
  for (uint32_t base = blk * blockDim.x; base < no; base += stride) {   // block-uniform trip count
    const uint32_t i = base + threadIdx.x;
    const bool has = i < no;
    const Event e = has ? load_event(&ov[i]) : Event{};
    uint32_t r = has ? route_of(sw, B, e.ts, end_bits, 0u) : Route::kNone;
    if (r == 0u) {
      stage_one(q, st, e);
      r = Route::kNone;
    }
    block_append(q, x, wp, r, e, ba);
  }
}

}  // namespace cal
}  // namespace gpq

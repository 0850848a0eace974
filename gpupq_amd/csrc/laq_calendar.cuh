// laq_calendar.cuh -- calendar mode of the local array queue and the PHOLD step on it.
//
// The merge-and-commit PHOLD step (drain mode 2) ends window k at w_k = min(w_{k-1} + lookahead,
// end time) in float arithmetic, so every window end is known as soon as the first one is.  In
// calendar mode the queue uses that: it keeps one append log per upcoming window (a calendar
// queue, R. Brown, CACM 31(10), 1988, with one day per window), split by the block of LPs the
// events are for, and an event is appended to the log of its window and LP block.  Step c is one
// launch with two kinds of blocks that run side by side:
//   * commit blocks, one per block of kBlock LPs: the block's log of window c is grouped by LP in
//     LDS, each LP's events are sorted in registers and committed in key order, and their
//     successors are appended to the logs of their windows and LP blocks (one atomic each);
//   * a few list blocks: the events beyond the calendar (over[(c + 1) & 1]) whose window the
//     calendar now reaches move into the logs.
// No sorted run is kept, merged or shifted, and no separate pass groups the events by LP.
//
// Layout (B = cal_windows, a power of two; L = num_lps; K = kBlock LPs per block):
//   wlog   [B][blocks][sub_cap]  Event   window x, LP block b: count wlog_n[((x mod B) * blocks + b) * 16]
//   over   [2][over_cap]         Event   events beyond the calendar: appended to over[c & 1] by step
//                                        c, read (and moved on) by step c + 1
//   parked [parked_cap]          Event   events at or past the end time: never committed, only kept
//   cal_st                               `cur` (the window the next step commits) and the ends of
//                                        windows x in [cur, cur + B] at w[x mod 2B]
// A step appends only to windows c + 1 .. c + B - 1, so the log slot of window c (the one its
// commit blocks read and then empty) takes no appends during the step.
//
// Reference: the per-LP queue of src/local_array_queue.cu and its heap-backed far tier
// (:515-681); the step is handle_next_event (src/phold.cu:81-117) for every event of a window.
#pragma once

#include "phold_fused.cuh"

namespace gpq {
namespace cal {

// LPs per commit block, and so per log.  Smaller blocks mean more logs and shorter runs of the
// self-addressed successors, larger ones fewer blocks in flight: at 2^20 LPs 64 LPs 1.37e10,
// 128 LPs 1.59-1.66e10, 256 LPs 1.74-1.83e10, 512 LPs 1.60e10, 1024 LPs 1.42e10 events/s.
constexpr int kBlock = 256;
constexpr int kRegs = 8;          // events of one LP sorted in registers by one thread (fast path)
constexpr int kWaveCand = 64;     // events of one LP on the slow path (one per lane)
constexpr int kLogStage = 896;    // log records whose keys a commit block may keep in LDS while grouping
constexpr int kGroupBlocks = 2048;        // blocks of the standalone filing launch
constexpr uint32_t kStepListBlocks = 32;  // blocks of a step that move the beyond-the-calendar events

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

// Window offset of timestamp bits ts: the smallest o in [o_min, n) with ts < sw[o], where
// sw[o] is the end of window x + o; n when ts >= sw[n - 1] (beyond the windows looked at).
__device__ __forceinline__ uint32_t find_window(const uint32_t* sw, uint32_t n, uint32_t ts, uint32_t o_min) {
  if (!(ts < sw[n - 1])) return n;
  uint32_t lo = o_min, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ts < sw[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Where an event goes, seen from window x with n windows looked at: a window offset o < n
// (the log of window x + o), kOver (beyond them) or kParked (at or past the end time).
struct Route {
  static constexpr uint32_t kOver = kCalMaxWindows;
  static constexpr uint32_t kParked = kCalMaxWindows + 1;
  static constexpr uint32_t kNone = kCalMaxWindows + 2;   // no record
};
constexpr uint32_t kRemote = 0x10000u;   // flag on a window route: the event is for another block's LPs

__device__ __forceinline__ uint32_t route_of(const uint32_t* sw, uint32_t n, uint32_t ts, uint32_t end_bits,
                                             uint32_t o_min) {
  if (!(ts < end_bits)) return Route::kParked;
  const uint32_t o = find_window(sw, n, ts, o_min);
  return o < n ? o : Route::kOver;
}

// The log of window x for LP block b: its counter and its first record.  Counters sit 16 words
// (64 bytes) apart: packed, the counters of one window for 32 neighbouring LP blocks shared a
// 128-byte line, and every block's successors for those blocks' logs met on it (about 2,000
// atomics per line and step at 2^20 LPs).  Spread out, the step took 154 against 159 us (nine
// A/B pairs at 2^20 LPs) and the instances of one calendar ran 9% faster; 32 words apart gave no
// more.
constexpr uint32_t kLogCountShift = 4;
__device__ __forceinline__ uint32_t* log_count(const LaqView& q, uint32_t x, uint32_t b) {
  return &q.wlog_n[((size_t)(x & (q.cal_windows - 1u)) * q.cal_blocks + b) << kLogCountShift];
}
__device__ __forceinline__ Event* log_base(const LaqView& q, uint32_t x, uint32_t b) {
  return q.wlog + (size_t)(x & (q.cal_windows - 1u)) * q.wlog_cap + (size_t)b * q.cal_sub_cap;
}

// e appended to the log of window x for its LP's block, by its own atomic.
__device__ __forceinline__ void log_append(const LaqView& q, uint32_t x, const Event& e) {
  const uint32_t b = e.dst >> q.cal_block_log2;
  const uint32_t pos = atomicAdd(log_count(q, x, b), 1u);
  if (pos < q.cal_sub_cap) store_event(log_base(q, x, b) + pos, e);
  else atomicOr(&q.ws->error, kErrQueueOverflow);
}

// The lists: beyond the calendar (over[op]) and parked.
__device__ __forceinline__ uint32_t* list_count(const LaqView& q, uint32_t op, uint32_t r) {
  return r == Route::kParked ? q.parked_n : &q.over_n[op];
}
__device__ __forceinline__ void list_store(const LaqView& q, uint32_t op, uint32_t r, uint32_t pos, const uint4& e) {
  const bool parked = r == Route::kParked;
  if (pos < (parked ? q.parked_cap : q.over_cap))
    *reinterpret_cast<uint4*>(parked ? q.parked + pos : q.over + (size_t)op * q.over_cap + pos) = e;
  else
    atomicOr(&q.ws->error, kErrQueueOverflow);
}

// One event filed by its own atomic: route r seen from window x (a window offset, kOver: the
// list over[op], or kParked).
__device__ __forceinline__ void append_one(const LaqView& q, uint32_t x, uint32_t op, uint32_t r, const Event& e) {
  if (r < kCalMaxWindows) {
    log_append(q, x + r, e);
    return;
  }
  list_store(q, op, r, atomicAdd(list_count(q, op, r), 1u), make_uint4(e.ts, e.dst, e.src, e.seq));
}

// Block-wide filing of at most one event per thread: a window route goes to its log by the
// event's own atomic; the list routes take a rank in the block's LDS count of their list, one
// thread per list reserves the block's run with a single atomic, and the events are written
// into the runs.  op is block-uniform; every thread of the block calls.
struct ListAppend {
  uint32_t cnt[2];
  uint32_t base[2];
};

__device__ __forceinline__ void block_file(const LaqView& q, uint32_t x, uint32_t op, uint32_t r, const Event& e,
                                           ListAppend& la) {
  if (threadIdx.x < 2) la.cnt[threadIdx.x] = 0u;
  __syncthreads();
  const bool list = r == Route::kOver || r == Route::kParked;
  const uint32_t li = r == Route::kParked ? 1u : 0u;
  uint32_t rank = 0;
  if (r < kCalMaxWindows) log_append(q, x + r, e);
  else if (list) rank = atomicAdd(&la.cnt[li], 1u);
  __syncthreads();
  if (threadIdx.x < 2) {
    const uint32_t n = la.cnt[threadIdx.x];
    la.base[threadIdx.x] = n ? atomicAdd(list_count(q, op, threadIdx.x ? Route::kParked : Route::kOver), n) : 0u;
  }
  __syncthreads();
  if (list) list_store(q, op, r, la.base[li] + rank, make_uint4(e.ts, e.dst, e.src, e.seq));
  __syncthreads();
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

// Slow path, one wave for one LP with more than kRegs events in window c: the block's log of the
// window is scanned for them (lane i gets the LP's i-th; with kStage, the keys the grouping kept
// in LDS -- st, ss, sq, and the LP's index in the block sl -- stand for the log's first
// kLogStage records), each lane ranks its event among the others, and the successors are filed
// by their own atomics.
template <bool kStage>
__device__ __forceinline__ void slow_lp(const LaqView& q, const FusedModel& m, const uint32_t* sw, uint32_t c,
                                        uint32_t end_bits, uint32_t lp, const Event* lg, uint32_t n, Key* keys,
                                        const uint32_t* st, const uint32_t* ss, const uint32_t* sq,
                                        const uint8_t* sl) {
  const int lane = lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t l = lp & (kBlock - 1u);
  uint32_t k = 0;
  for (uint32_t base = 0; base < n; base += kWave) {     // wave-uniform trip count
    const uint32_t i = base + (uint32_t)lane;
    Event x{kInfBits, 0xFFFFFFFFu, 0u, 0u};
    if (kStage && i < n && i < (uint32_t)kLogStage) {
      if (sl[i] == l) x = Event{st[i], lp, ss[i], sq[i]};
    } else if (i < n) {
      x = load_event(&lg[i]);
    }
    const bool mine = x.dst == lp;
    const uint64_t mm = __ballot(mine);
    const uint32_t pos = k + (uint32_t)__popcll(mm & below);
    if (mine && pos < (uint32_t)kWaveCand) keys[pos] = Key{x.ts, x.src, x.seq};
    k += (uint32_t)__popcll(mm);
  }
  if (k > (uint32_t)kWaveCand) {
    if (lane == 0) atomicOr(&q.ws->error, kErrStageOverflow);
    k = kWaveCand;
  }
  laq::wave_lds_sync();
  const bool has = (uint32_t)lane < k;
  const Key me = has ? keys[lane] : Key{kInfBits, 0u, 0u};
  uint32_t rank = 0;
  for (uint32_t j = 0; j < k; ++j) {
    const Key o{(uint32_t)__shfl((int)me.ts, (int)j, kWave), (uint32_t)__shfl((int)me.src, (int)j, kWave),
                (uint32_t)__shfl((int)me.seq, (int)j, kWave)};
    rank += key_lt(o, me) ? 1u : 0u;
  }
  laq::wave_lds_sync();                                  // every lane has read its key
  const uint32_t cnt = m.count[lp];
  Event ne{};
  uint32_t r = Route::kNone;
  if (has) {
    keys[rank] = me;
    ne = successor(m, lp, cnt + rank, me.ts);
    r = route_of(sw, q.cal_windows, ne.ts, end_bits, 1u);
    if (r < kCalMaxWindows) log_append(q, c + r, ne);
  }
  // the list successors: one reservation per list for the wave.  Near the end time almost every
  // successor is parked, and with an atomic per lane on the one list counter the last steps took
  // up to 410 us instead of 110-160.
#pragma unroll
  for (uint32_t li = 0; li < 2u; ++li) {
    const uint32_t route = li ? Route::kParked : Route::kOver;
    const uint64_t mm = __ballot(r == route);
    if (!mm) continue;                                   // uniform across the wave
    const int leader = __builtin_ctzll(mm);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(list_count(q, c & 1u, route), (uint32_t)__popcll(mm));
    base = (uint32_t)__shfl((int)base, leader, kWave);
    if (r == route) list_store(q, c & 1u, route, base + (uint32_t)__popcll(mm & below), make_uint4(ne.ts, ne.dst, ne.src, ne.seq));
  }
  laq::wave_lds_sync();
  if (lane == 0) {
    uint64_t h = m.hash[lp];
    uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
    for (uint32_t r = 0; r < k; ++r) {
      const Key x = keys[r];
      h = trace_step(h, x.ts, x.src, x.seq);
      if (m.last_ts) {
        if (x.ts < last) atomicOr(&q.ws->error, kErrBadState);
        last = x.ts;
      }
    }
    m.count[lp] = cnt + k;
    m.hash[lp] = h;
    if (m.last_ts) m.last_ts[lp] = last;
  }
  laq::wave_lds_sync();
}

// Commit block `blk` of step c (window c): the block's log of the window is grouped by LP in LDS
// (counts, and the log index of each LP's first kRegs events).  Fast path, one thread per LP
// with at most kRegs events: sorted in registers and committed in order.  A successor for one of
// this block's LPs (half of them in PHOLD: the self-addressed ones) and one for a list take a
// rank in the block's LDS count of their window or list, one thread per window and list
// reserves the block's run, and they are written into the runs; a successor for another block
// goes to that block's log by its own atomic, all of a thread's issued before any is waited for.
// LPs with more events are handled afterwards, one wave each (slow_lp).  The log is emptied at
// the end.
// kStage: the keys of the log's first kLogStage records are kept in LDS while grouping, and the
// commit and the slow path read them there (the instances of one calendar: 1.81 -> 1.93e10
// events/s; one 2^20-LP instance: 2-4% slower, so it keeps kStage off).
template <bool kStage>
__device__ __forceinline__ void commit_block(const LaqView& q, const FusedModel& m, float end_time, uint32_t c,
                                             uint32_t blk) {
  __shared__ uint32_t sw[kCalMaxWindows];
  __shared__ uint32_t s_cnt[kBlock];
  __shared__ uint16_t s_idx[kBlock][kRegs];
  __shared__ uint32_t s_lcnt[2], s_lbase[2];
  __shared__ uint32_t s_wcnt[kCalMaxWindows], s_wbase[kCalMaxWindows];
  __shared__ uint32_t s_slow[kBlock];
  __shared__ uint32_t s_nslow;
  __shared__ Key s_keys[kBlock / kWave][kWaveCand];
  constexpr int kSt = kStage ? kLogStage : 1;
  __shared__ uint32_t s_lt[kSt], s_ls[kSt], s_lq[kSt];   // with kStage: keys of the log's first records
  __shared__ uint8_t s_ll[kSt];                          // and their LP's index in the block
  static_assert(kBlock <= 256, "an LP's index in its block fits a byte");
  const uint32_t lp0 = blk << q.cal_block_log2;
  const uint32_t lp = lp0 + threadIdx.x;
  const bool in_range = lp < q.num_lps && lp - m.base < m.lps;
  const Event* lg = log_base(q, c, blk);
  // issued together with the window ends' load: the log's length, the LP of the log record this
  // thread groups first (slot threadIdx.x exists: a block's log has at least kBlock slots; used
  // only below the length) and the thread's LP state, so that their latencies overlap
  const uint32_t n0 = *log_count(q, c, blk);
  Event e0{};
  if (kStage) e0 = load_event(&lg[threadIdx.x]);
  const uint32_t d0 = kStage ? e0.dst : lg[threadIdx.x].dst;
  const uint32_t cnt_pre = in_range ? m.count[lp] : 0u;
  const uint64_t h_pre = in_range ? m.hash[lp] : 0ull;
  load_windows(q, c, sw);
  const uint32_t B = q.cal_windows;
  s_cnt[threadIdx.x] = 0u;
  for (uint32_t i = threadIdx.x; i < B; i += kBlock) s_wcnt[i] = 0u;
  if (threadIdx.x < 2) s_lcnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) s_nslow = 0;
  __syncthreads();
  const uint32_t n = n0 < q.cal_sub_cap ? n0 : q.cal_sub_cap;
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
    Event e{};
    if (kStage) e = i == threadIdx.x ? e0 : load_event(&lg[i]);
    const uint32_t l = (kStage ? e.dst : (i == threadIdx.x ? d0 : lg[i].dst)) - lp0;   // < kBlock: this block's LPs only
    if (l >= (uint32_t)kBlock) {
      atomicOr(&q.ws->error, kErrBadState);
      continue;
    }
    if (kStage && i < (uint32_t)kLogStage) {            // kept for the commit and the slow path
      s_lt[i] = e.ts;
      s_ls[i] = e.src;
      s_lq[i] = e.seq;
      s_ll[i] = (uint8_t)l;
    }
    const uint32_t k = atomicAdd(&s_cnt[l], 1u);
    if (k < (uint32_t)kRegs) s_idx[l][k] = (uint16_t)i;
  }
  __syncthreads();
  const uint32_t end_bits = f_to_bits(end_time);
  uint4 ne[kRegs];      // list successors as uint4: an array of the 16-byte-aligned Event goes to scratch memory
  uint32_t route[kRegs], rank[kRegs];
  uint32_t cn = 0;
  if (in_range) {
    const uint32_t k = s_cnt[threadIdx.x];
    if (k > (uint32_t)kRegs) {
      s_slow[atomicAdd(&s_nslow, 1u)] = lp;
    } else if (k) {
      cn = k;
      uint32_t kt[kRegs], ks[kRegs], kq[kRegs];
#pragma unroll
      for (int s = 0; s < kRegs; ++s) {
        if ((uint32_t)s < cn) {
          const uint32_t ix = s_idx[threadIdx.x][s];
          if (kStage && ix < (uint32_t)kLogStage) {
            kt[s] = s_lt[ix];
            ks[s] = s_ls[ix];
            kq[s] = s_lq[ix];
          } else {
            const Event e = load_event(&lg[ix]);
            kt[s] = e.ts;
            ks[s] = e.src;
            kq[s] = e.seq;
          }
        } else {
          kt[s] = ks[s] = kq[s] = 0xFFFFFFFFu;
        }
      }
      const uint32_t cnt = cnt_pre;
      uint64_t h = h_pre;
      uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
      sort8(kt, ks, kq);
#pragma unroll
      for (int j = 0; j < kRegs; ++j) {
        route[j] = Route::kNone;
        if ((uint32_t)j < cn) {
          h = trace_step(h, kt[j], ks[j], kq[j]);
          if (m.last_ts) {
            if (kt[j] < last) atomicOr(&q.ws->error, kErrBadState);
            last = kt[j];
          }
          const Event sc = successor(m, lp, cnt + (uint32_t)j, kt[j]);
          const uint32_t r = route_of(sw, B, sc.ts, end_bits, 1u);
          route[j] = r;
          ne[j] = make_uint4(sc.ts, sc.dst, sc.src, sc.seq);
          if (r < kCalMaxWindows) {
            if ((sc.dst >> q.cal_block_log2) == blk) rank[j] = atomicAdd(&s_wcnt[r], 1u);   // this block's log
            else route[j] = r | kRemote;
          } else {
            rank[j] = atomicAdd(&s_lcnt[r == Route::kParked ? 1u : 0u], 1u);
          }
        }
      }
      m.count[lp] = cnt + cn;
      m.hash[lp] = h;
      if (m.last_ts) m.last_ts[lp] = last;
    }
  }
  // successors for other blocks' logs: every slot atomic of the thread in flight at once
  uint32_t pos[kRegs];
#pragma unroll
  for (int j = 0; j < kRegs; ++j)
    pos[j] = (uint32_t)j < cn && (route[j] & kRemote)
                 ? atomicAdd(log_count(q, c + (route[j] & ~kRemote), ne[j].y >> q.cal_block_log2), 1u)
                 : 0u;
  __syncthreads();
  // one run per window (this block's own log) and per list for the block
  for (uint32_t i = threadIdx.x; i < B; i += kBlock) {
    const uint32_t t = s_wcnt[i];
    s_wbase[i] = t ? atomicAdd(log_count(q, c + i, blk), t) : 0u;
  }
  if (threadIdx.x < 2) {
    const uint32_t t = s_lcnt[threadIdx.x];
    s_lbase[threadIdx.x] = t ? atomicAdd(list_count(q, c & 1u, threadIdx.x ? Route::kParked : Route::kOver), t) : 0u;
  }
#pragma unroll
  for (int j = 0; j < kRegs; ++j) {
    if ((uint32_t)j < cn && (route[j] & kRemote)) {
      const uint32_t r = route[j] & ~kRemote;
      if (pos[j] < q.cal_sub_cap) *reinterpret_cast<uint4*>(log_base(q, c + r, ne[j].y >> q.cal_block_log2) + pos[j]) = ne[j];
      else atomicOr(&q.ws->error, kErrQueueOverflow);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRegs; ++j) {
    if ((uint32_t)j >= cn || (route[j] & kRemote)) continue;
    const uint32_t r = route[j];
    if (r < kCalMaxWindows) {
      const uint32_t p = s_wbase[r] + rank[j];
      if (p < q.cal_sub_cap) *reinterpret_cast<uint4*>(log_base(q, c + r, blk) + p) = ne[j];
      else atomicOr(&q.ws->error, kErrQueueOverflow);
    } else {
      list_store(q, c & 1u, r, s_lbase[r == Route::kParked ? 1u : 0u] + rank[j], ne[j]);
    }
  }
  const uint32_t w = threadIdx.x / kWave;
  for (uint32_t i = w; i < s_nslow; i += kBlock / kWave) slow_lp<kStage>(q, m, sw, c, end_bits, s_slow[i], lg, n, s_keys[w], s_lt, s_ls, s_lq, s_ll);
  __syncthreads();
  if (threadIdx.x == 0) *log_count(q, c, blk) = 0u;     // the block's log of window c is consumed
}

// The end of the step that committed window c (its logs emptied by the commit blocks): the list
// the step's list blocks read is emptied; the window advances by the lookahead (capped at the end
// time) and the end of window c + 1 + B joins the ring; the window that reached the end time ends
// the run.  Only the fields it owns are written: other blocks of the step may be setting the
// error flag.
__device__ __forceinline__ void publish(const LaqView& q, float lookahead, float end_time) {
  WindowState* ws = q.ws;
  CalState* st = q.cal_st;
  const uint32_t B = q.cal_windows, M = 2u * B - 1u;
  const uint32_t c = st->cur;
  const uint32_t wend = ws->wend;
  q.over_n[(c + 1u) & 1u] = 0;
  ws->windows += 1;
  ws->gmin = wend;                                 // a lower bound of the pending minimum
  if (!(bits_to_f(wend) < end_time)) {
    ws->done = 1;
  } else {
    const uint32_t nc = c + 1u;
    st->cur = nc;
    float nw = bits_to_f(st->w[(nc + B - 1u) & M]) + lookahead;
    if (end_time < nw) nw = end_time;
    st->w[(nc + B) & M] = f_to_bits(nw);
    ws->wend = st->w[nc & M];
  }
}

// Called by thread 0 of every block of a step once the block is done: the last block publishes.
// No fence is needed: the words publish() reads are written by no block of the step, the list it
// empties was read (and the values used) by the list blocks before they counted themselves, the
// error flag it leaves alone, and its writes reach the next step across the kernel boundary.
__device__ __forceinline__ void block_done(const LaqView& q, float lookahead, float end_time) {
  if (atomicAdd(&q.cal_st->fin, 1u) == gridDim.x - 1u) {
    publish(q, lookahead, end_time);
    q.cal_st->fin = 0u;
  }
}

// List block `blk` of `nblk` in the step that commits window x - 1: the events of over[x & 1]
// (beyond the windows the step before could reach) are seen from window x over its first B - 1
// windows -- x .. x + B - 2, all of them beyond the window being committed and short of the
// one whose log slot it is reading -- and go to the log of their window, on to over[(x + 1) & 1],
// or with the parked ones.
__device__ __forceinline__ void list_block(const LaqView& q, float end_time, uint32_t x, uint32_t blk,
                                           uint32_t nblk) {
  __shared__ uint32_t sw[kCalMaxWindows];
  __shared__ ListAppend la;
  load_windows(q, x, sw);
  __syncthreads();
  const uint32_t B = q.cal_windows;
  const uint32_t rp = x & 1u, wp = rp ^ 1u;
  const uint32_t no0 = q.over_n[rp];
  const uint32_t no = no0 < q.over_cap ? no0 : q.over_cap;
  const Event* ov = q.over + (size_t)rp * q.over_cap;
  const uint32_t end_bits = f_to_bits(end_time);
  const uint32_t stride = nblk * blockDim.x;
  for (uint32_t base = blk * blockDim.x; base < no; base += stride) {   // block-uniform trip count
    const uint32_t i = base + threadIdx.x;
    const bool has = i < no;
    const Event e = has ? load_event(&ov[i]) : Event{};
    const uint32_t r = has ? route_of(sw, B - 1u, e.ts, end_bits, 0u) : Route::kNone;
    block_file(q, x, wp, r, e, la);
  }
}

}  // namespace cal
}  // namespace gpq

// queue_laq.cuh -- device API of the local array queue (LAQ): one sorted ring per
// LP.  Reference: src/local_array_queue.cu (insert :486-513, peek :760-779,
// set_done :710-728, is_empty/length/clear :795-822, insert_or_update :830-846).
//
// HBM layout (LP-major so that a wave reading one LP's run is contiguous):
//   ring  [L][C]  Event   sorted run of LP lp at slots rd[lp] .. rd[lp]+len[lp]-1 (mod C)
//   stage [L][S]  Event   this step's inserts for LP lp (unsorted), count ins[lp]
//   head  [L]     u32     ts bits of the ring head, +inf when empty (read coalesced by peek)
//   rd, len, ins [L] u32
// Two-tier mode (far_cap > 0): the ring keeps only events below the near horizon
// ws->near_end; the others wait unsorted in
//   far   [L][F]  Event   count nfar[lp], smallest ts bits far_min[lp] (+inf when empty)
// and move into the ring when the horizon advances.  With lookahead = +inf (open lists, the
// A* model's) each LP has a horizon of its own, a full key (ts, src, seq) in near_lp[0/1/2][L]:
// the ring keeps the LP's records below it, so the ring head is the LP's minimum.  A post moves
// the smallest far records into a ring that holds fewer than 8, and the ring's largest into the
// far store when it would overflow (the heap-backed LAQ's exchange,
// src/local_array_queue.cu:150-180,515-681).
// Inserts never touch the ring: they land in the staging slots and are merged by
// queue_post.  The reference appends after fel_write and insertion-sorts in place
// (src/local_array_queue.cu:486-513, src/sort.cu:329-356).
#pragma once

#include "core.cuh"

namespace gpq {

constexpr uint32_t kCalMaxWindows = 256;

// Calendar mode: the window index `cur` of the step that commits next, and the ends of the
// windows x in [cur, cur + B] at w[x mod 2B] (float bits; every window after the one that
// reaches the end time ends at the end time).  `fin` counts the blocks of the running step that
// have finished (the last one moves the window and sets it back to 0); it has a cache line of its
// own, away from the words every block reads when it starts.
struct CalState {
  uint32_t cur;
  uint32_t pad[3];
  uint32_t w[2 * kCalMaxWindows];
  alignas(128) uint32_t fin;
};

struct LaqView {
  Event* ring;
  Event* stage;
  uint32_t* rd;
  uint32_t* len;
  uint32_t* ins;
  uint32_t* head;
  WindowState* ws;
  unsigned long long* stats;   // work counters, nullptr when not collected
  const uint8_t* block_live;    // per 256-LP post block, 0 = holds no LP in use (skipped); nullptr = all live
  // Two-tier mode with several windows (a pool of instances): the near horizon of post block b
  // is that of block_ws[block_win[b]]; nullptr = ws (one window).
  const uint32_t* block_win;
  const WindowState* block_ws;
  // Second staging buffer of the PHOLD merge-and-commit step (nullptr otherwise): a step reads
  // one buffer while the events it schedules land in the other.
  Event* stage_alt;
  uint32_t* ins_alt;
  // Decrease-key mode (one-tier only; in_queue == nullptr = off): one bit per item id
  // (Event.seq) marks an id with an entry in the queue, staged or in a ring; an
  // insert_or_update of a marked id goes to the LP's update buffer, applied by queue_post.
  uint32_t* in_queue;
  Event* upd;                   // [L][upd_cap]
  uint32_t* upd_n;
  uint32_t upd_cap;
  uint32_t update_ids;
  // Calendar mode (cal_windows > 0; the PHOLD merge-and-commit step only): one append log per
  // upcoming window and block of 2^cal_block_log2 LPs: the events of window cur + o for the LPs of
  // block b are at wlog + ((cur + o) mod cal_windows) * wlog_cap + b * cal_sub_cap, count
  // wlog_n[(((cur + o) mod cal_windows) * cal_blocks + b) * 16]; events beyond the calendar in over[p]
  // (p = parity of the step that appended them), count over_n[p]; events at or past the end
  // time in parked, count *parked_n.  A step's commit blocks read the log of their own LPs for
  // its window and group it by LP in LDS.
  Event* wlog;
  uint32_t* wlog_n;
  Event* over;
  uint32_t* over_n;
  Event* parked;
  uint32_t* parked_n;
  CalState* cal_st;
  uint32_t cal_windows;
  uint32_t wlog_cap;            // records per window: cal_blocks * cal_sub_cap
  uint32_t over_cap;
  uint32_t parked_cap;
  uint32_t cal_sub_cap;         // records per window and block
  uint32_t cal_blocks;          // blocks of 2^cal_block_log2 LPs (the commit block size)
  uint32_t cal_block_log2;
  uint32_t cal_pad;
  Event* far;                   // two-tier mode only (far_cap > 0)
  uint32_t* nfar;
  uint32_t* far_min;
  uint32_t* near_lp;            // two-tier open-list mode (lookahead = +inf) only: per-LP horizon, [3][L]
  uint32_t far_cap;
  uint32_t num_lps;
  uint32_t cap_log2;
  uint32_t stage_cap;

  __device__ __forceinline__ uint32_t slots() const { return num_lps; }

  // queue_insert: staged for the owning LP; false (and a device flag) when the
  // LP's staging area is full.
  __device__ __forceinline__ bool insert(const Event& e) const { return stage_insert(stage, ins, stage_cap, ws, e); }

  // queue_peek: the head of LP lp if its ts is inside the current safe window.
  __device__ __forceinline__ int peek(Event** item, int lp) const {
    const uint32_t h = head[lp];
    if (!(h < ws->wend)) return -1;
    const uint32_t idx = ((uint32_t)lp << cap_log2) | rd[lp];
    *item = ring + idx;
    return (int)idx;
  }

  // queue_set_done: removes the head; the new head's ts is refreshed for the window scan.
  __device__ __forceinline__ void set_done(int index) const {
    if (in_queue) {                       // the id leaves the queue
      const uint32_t id = ring[index].seq;
      if (id < update_ids) atomicAnd(&in_queue[id >> 5], ~(1u << (id & 31u)));
    }
    const uint32_t lp = (uint32_t)index >> cap_log2;
    const uint32_t mask = (1u << cap_log2) - 1u;
    const uint32_t r = (rd[lp] + 1u) & mask;
    const uint32_t n = len[lp] - 1u;
    rd[lp] = r;
    len[lp] = n;
    head[lp] = n ? ring[(lp << cap_log2) | r].ts : kInfBits;
  }

  __device__ __forceinline__ bool is_empty(int lp) const {
    return len[lp] == 0 && ins[lp] == 0 && (far_cap == 0 || nfar[lp] == 0);
  }
  // Pending events of LP lp: the sorted ring plus, in two-tier mode, the far store (staged
  // inserts of the current step are not counted until queue_post folds them in).
  __device__ __forceinline__ int length(int lp) const { return (int)(len[lp] + (far_cap ? nfar[lp] : 0u)); }
  __device__ __forceinline__ void clear(int lp) const {
    const uint32_t mask = (1u << cap_log2) - 1u;
    if (in_queue) {                       // every id of the LP leaves the queue
      for (uint32_t i = 0; i < len[lp]; ++i) {
        const uint32_t id = ring[((uint32_t)lp << cap_log2) | ((rd[lp] + i) & mask)].seq;
        if (id < update_ids) atomicAnd(&in_queue[id >> 5], ~(1u << (id & 31u)));
      }
      upd_n[lp] = 0;
    }
    rd[lp] = (rd[lp] + len[lp]) & mask;
    len[lp] = 0;
    head[lp] = kInfBits;
    if (far_cap) {
      nfar[lp] = 0;
      far_min[lp] = kInfBits;
    }
    if (near_lp) {
      near_lp[lp] = 0xFFFFFFFFu;
      near_lp[num_lps + lp] = 0xFFFFFFFFu;
      near_lp[2 * num_lps + lp] = 0xFFFFFFFFu;
    }
  }
  // queue_insert_or_update (src/local_array_queue.cu:830-846).  Without decrease-key mode a
  // plain insert (PHOLD events carry no identity).  With it: the first insert of an id marks
  // it and is staged; a later one, while the id has an entry, goes to the update buffer of
  // LP lp (= e.dst), and queue_post replaces the entry when the new key is smaller.
  __device__ __forceinline__ void insert_or_update(const Event& e, int lp) const {
    if (!in_queue) {
      insert(e);
      return;
    }
    if (e.seq >= update_ids) {
      atomicOr(&ws->error, kErrBadState);
      return;
    }
    const uint32_t bit = 1u << (e.seq & 31u);
    if (atomicOr(&in_queue[e.seq >> 5], bit) & bit) {
      const uint32_t slot = atomicAdd(&upd_n[lp], 1u);
      if (slot < upd_cap) store_event(&upd[(size_t)lp * upd_cap + slot], e);
      else atomicOr(&ws->error, kErrStageOverflow);
    } else {
      insert(e);
    }
  }
};

}  // namespace gpq

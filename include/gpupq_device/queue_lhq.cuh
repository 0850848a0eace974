// queue_lhq.cuh -- device API of the local heap queue (LHQ): one binary min-heap per
// LP, owned by one thread.  Reference: src/local_heap_queue.cu (insert :398-419,
// copy_last :128-145, heapify_down :249-267, heapify_up :216-233, peek :430-447,
// set_done :451-467).
//
// HBM layout: slot-major, heap[i * L + lp] is slot i of LP lp's heap, so the
// threads of a wave (consecutive LPs) read their roots and upper levels from
// consecutive addresses; the reference keeps each heap contiguous (fel[lp * FEL_SIZE + i]).
// head[lp] caches the root's ts bits for the coalesced window scan.
#pragma once

#include "core.cuh"

namespace gpq {

struct LhqView {
  Event* heap;
  Event* stage;
  uint32_t* len;
  uint32_t* ins;
  uint32_t* head;
  WindowState* ws;
  unsigned long long* stats;
  uint32_t num_lps;
  uint32_t cap;
  uint32_t stage_cap;

  __device__ __forceinline__ Event* slot(uint32_t lp, uint32_t i) const { return heap + (size_t)i * num_lps + lp; }
  __device__ __forceinline__ uint32_t slots() const { return num_lps; }
  __device__ __forceinline__ bool insert(const Event& e) const { return stage_insert(stage, ins, stage_cap, ws, e); }

  // queue_peek: the root of LP lp if inside the window; the index is the LP.
  __device__ __forceinline__ int peek(Event** item, int lp) const {
    if (!(head[lp] < ws->wend)) return -1;
    *item = slot((uint32_t)lp, 0);
    return lp;
  }

  // queue_set_done: removes the root -- the last slot moves to the root and sifts down
  // (the reference defers this to copy_last + heapify_down in queue_post).  Returns moves.
  __device__ __forceinline__ uint32_t pop_root(uint32_t lp) const {
    const uint32_t n = len[lp] - 1u;
    uint32_t moves = 0;
    if (n > 0) {
      const Event x = load_event(slot(lp, n));
      uint32_t i = 0;
      for (;;) {
        uint32_t c = 2u * i + 1u;
        if (c >= n) break;
        Event ec = load_event(slot(lp, c));
        if (c + 1u < n) {
          const Event er = load_event(slot(lp, c + 1u));
          if (key_less(er, ec)) { ec = er; c = c + 1u; }
        }
        if (!key_less(ec, x)) break;
        store_event(slot(lp, i), ec);
        i = c;
        ++moves;
      }
      store_event(slot(lp, i), x);
      head[lp] = i == 0 ? x.ts : slot(lp, 0)->ts;
    } else {
      head[lp] = kInfBits;
    }
    len[lp] = n;
    return moves;
  }
  __device__ __forceinline__ void set_done(int index) const { (void)pop_root((uint32_t)index); }


  __device__ __forceinline__ bool is_empty(int lp) const { return len[lp] == 0 && ins[lp] == 0; }
  __device__ __forceinline__ int length(int lp) const { return (int)len[lp]; }
  __device__ __forceinline__ void clear(int lp) const {
    len[lp] = 0;
    head[lp] = kInfBits;
  }
  __device__ __forceinline__ void insert_or_update(const Event& e, int) const { insert(e); }
};

}  // namespace gpq

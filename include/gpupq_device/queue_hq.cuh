// queue_hq.cuh -- device API of the global parallel heap (HQ): a binary heap whose nodes
// hold up to r sorted records, every record of a node <= every record of its children.
// Reference: src/heap_queue.cu (insert :1809-1826, peek :1860-1896, set_done :1898-1904,
// post :1609-1719, partition :478-547, delete-update :631-784, insert-update :817-960,
// copy_last :975-1094).
//
// HBM layout: nodes[v * r + i] (node v, position i), counts[v]; the step's inserts are
// appended to enq; queue_post leaves the next step's safe events in deq (peek slot = position).
#pragma once

#include "core.cuh"

namespace gpq {

constexpr int kHqMaxLevels = 20;
constexpr int kHqMaxChunks = 8;

struct HqState {
  uint32_t last;           // highest node holding records
  uint32_t m;              // inserts of this step (totalled over the stripes by queue_post)
  uint32_t mcount;         // records in the merged buffer
  uint32_t prefix, deq_n, rest_n;
  uint32_t bound_valid;
  uint32_t carry_n;        // records that travel down as insert chunks
  uint32_t select;         // 1: queue_post selects the next step's events
  uint32_t pad[3];
  Event bound;             // the root's largest record while nodes exist below the root
  const Event* mptr;       // = merged buffer (pointer slot for the selection)
  // insert-update chunks: target node, size, current data pointer, next data pointer
  uint32_t chunk_target[kHqMaxChunks];
  uint32_t chunk_n[kHqMaxChunks];
  const Event* chunk_ptr[kHqMaxChunks];
  Event* chunk_alt[kHqMaxChunks];
};

// Several heaps (HQ_NUM_QUEUES): heap h owns the LPs [h * lps_per_heap, (h + 1) * lps_per_heap),
// its state st[h], its insert area enq + h * enq_cap and its dequeue slots deq + h * lps_per_heap.
struct HqView {
  Event* enq;              // heap h's striped insert area (core.cuh) at enq + h * area
  Event* deq;
  HqState* st;
  WindowState* ws;
  uint32_t* seg_m;         // heap h's kInsertStripes + 1 append counters at seg_m + h * kStripeCounterWords
  uint32_t enq_cap;        // per heap: inserts one step may hold (and the spill segment's size)
  uint32_t seg_cap;
  uint32_t num_lps;
  uint32_t heaps;
  uint32_t lps_per_heap;
  uint32_t pad;
  uint64_t area;           // records of one heap's insert area: kInsertStripes * seg_cap + enq_cap

  __device__ __forceinline__ uint32_t slots() const { return num_lps; }
  // heap_queue_insert (src/heap_queue.cu:1809-1826): into the insert area of the LP's heap,
  // one counter atomic per wave, stripe and heap (every active lane of the wave takes part).
  __device__ __forceinline__ bool insert(const Event& e) const {
    bool ok = true;
    if (heaps == 1) {
      ok = striped_append(enq, seg_m, seg_cap, enq_cap, e);
    } else {
      const uint32_t h = e.dst / lps_per_heap;
      uint64_t pending = __ballot(1);
      while (pending) {                                  // one round per heap present in the wave
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lh = (uint32_t)__shfl((int)h, leader, kWave);
        const uint64_t same = __ballot(h == lh) & pending;
        if (h == lh) ok = striped_append(enq + lh * area, seg_m + lh * kStripeCounterWords, seg_cap, enq_cap, e);
        pending &= ~same;
      }
    }
    if (!ok) atomicOr(&ws->error, kErrStageOverflow);
    return ok;
  }
  // heap_queue_peek(item, pos): dequeue slot pos, i.e. slot pos % lps_per_heap of heap
  // pos / lps_per_heap (src/heap_queue.cu:1860-1896).
  __device__ __forceinline__ int peek(Event** item, int pos) const {
    const uint32_t h = (uint32_t)pos / lps_per_heap, i = (uint32_t)pos - h * lps_per_heap;
    if (i >= st[h].deq_n) return -1;
    *item = deq + pos;
    return pos;
  }
  __device__ __forceinline__ void set_done(int) const {}   // src/heap_queue.cu:1898-1904 (PHOLD)
  __device__ __forceinline__ void insert_or_update(const Event& e, int) const { insert(e); }
};

}  // namespace gpq

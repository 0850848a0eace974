// queue_gaq.cuh -- device API of the global array queue (GAQ): every pending event in
// one globally sorted array.  Reference: src/global_array_queue.cu (insert :347-365,
// mark_safe_events :147-179, peek :327-341, set_done :343-345).
//
// HBM layout: two sorted arrays fel[0], fel[1] (ping-pong, state->cur selects the live
// one), an append buffer for the step's inserts, and the dequeue buffer filled by
// queue_pre with the first record of each LP inside the safe window.  The reference
// keeps a `done` flag in every record and re-sorts the whole array each step; here the
// selected records are compacted out and the sorted inserts are merged in (one pass).
#pragma once

#include "core.cuh"

namespace gpq {

struct GaqState {
  Event* fel[2];
  const Event* front;   // fel[cur]
  uint32_t cur;
  uint32_t n;           // records in fel[cur]
  uint32_t m;           // inserts of this step (totalled over the stripes by queue_post)
  uint32_t prefix;      // records scanned by the last selection
  uint32_t deq_n;       // records in the dequeue buffer
  uint32_t rest_n;      // unselected records of the scanned prefix
  uint32_t selected;    // 1 between queue_pre and queue_post
  uint32_t n_next;
};

struct GaqView {
  Event* enq;           // striped insert area (core.cuh): kInsertStripes * seg_cap + enq_cap records
  Event* deq;
  GaqState* st;
  WindowState* ws;
  uint32_t* seg_m;      // kInsertStripes + 1 append counters, kStripeWords apart (core.cuh)
  uint32_t enq_cap;     // inserts one step may hold (and the spill segment's size)
  uint32_t seg_cap;
  uint32_t num_lps;
  uint32_t pad;

  __device__ __forceinline__ uint32_t slots() const { return num_lps; }

  // queue_insert: appended to the step's striped insert area (one counter atomic per wave).
  __device__ __forceinline__ bool insert(const Event& e) const {
    if (striped_append(enq, seg_m, seg_cap, enq_cap, e)) return true;
    atomicOr(&ws->error, kErrQueueOverflow);
    return false;
  }
  // queue_peek(item, pos): dequeue-buffer slot pos (src/global_array_queue.cu:327-341).
  __device__ __forceinline__ int peek(Event** item, int pos) const {
    if ((uint32_t)pos >= st->deq_n) return -1;
    *item = deq + pos;
    return pos;
  }
  // Selected records already left the array in queue_pre (src/global_array_queue.cu:343-345).
  __device__ __forceinline__ void set_done(int) const {}
  __device__ __forceinline__ void insert_or_update(const Event& e, int) const { insert(e); }
};

}  // namespace gpq

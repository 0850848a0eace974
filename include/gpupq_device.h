/*
 * gpupq_device.h -- device operator API of the priority queues, for model kernels written
 * outside the library (header-only HIP, gfx950).
 *
 * The reference's model kernels call free device functions bound to one queue at compile time
 * (src/queue.cuh:46-70):
 *   bool queue_insert(queue_item item);
 *   int  queue_peek(queue_item **item, int lp);          (HQ: heap_queue_peek(item, pos))
 *   void queue_set_done(int index);
 *   bool queue_is_empty(int lp);  int queue_length(int lp);  void queue_clear(int lp);
 *   void queue_insert_or_update(queue_item item, int lp);
 * Here the queue is a runtime object: gpq_queue_device_view (include/gpupq.h) hands the caller
 * the view struct of its queue's kind, the kernel takes it by value, and the same verbs are its
 * methods -- or the free functions below, which keep the reference's names with the view as
 * first argument.  Semantics:
 *   - insert: callable from any thread for any LP; the item is staged and folded in by
 *     gpq_queue_post; false (and a device error flag, GPQ_ERR_CAPACITY at the next host call)
 *     when the LP's staging area or the insert area is full.
 *   - peek(&item, slot): for the per-LP queues slot = LP, for HQ and GAQ slot = dequeue
 *     position (0 .. num_lps - 1); returns an index >= 0 and points *item into queue storage
 *     when that slot holds an event inside the current safe window, else -1.  The pointer stays
 *     valid until set_done or the next gpq_queue_post.
 *   - set_done(index): removes the peeked event; called by the slot's single owner thread.
 *   - is_empty / length / clear (per-LP queues only, by the LP's owner thread).
 *   - insert_or_update(item, lp): a plain insert, except on a LAQ in decrease-key mode, where an
 *     item whose id (seq) already has an entry replaces it at the next post if its key is smaller.
 * Order of keys: (ts, src, seq), ts as float32 bits (>= 0).  A kernel that uses the view runs
 * on the queue's stream (gpq_queue_stream) between gpq_queue_pre and gpq_queue_post.
 */
#ifndef GPUPQ_DEVICE_H
#define GPUPQ_DEVICE_H

#include "gpupq_device/core.cuh"
#include "gpupq_device/queue_gaq.cuh"
#include "gpupq_device/queue_hq.cuh"
#include "gpupq_device/queue_laq.cuh"
#include "gpupq_device/queue_lhq.cuh"
#include "gpupq_device/queue_lsq.cuh"
#include "gpupq.h"

namespace gpq {

template <class Q>
__device__ __forceinline__ uint32_t queue_slots(const Q& q) { return q.slots(); }
template <class Q>
__device__ __forceinline__ bool queue_insert(const Q& q, const Event& item) { return q.insert(item); }
template <class Q>
__device__ __forceinline__ int queue_peek(const Q& q, Event** item, int slot) { return q.peek(item, slot); }
template <class Q>
__device__ __forceinline__ void queue_set_done(const Q& q, int index) { q.set_done(index); }
template <class Q>
__device__ __forceinline__ bool queue_is_empty(const Q& q, int lp) { return q.is_empty(lp); }
template <class Q>
__device__ __forceinline__ int queue_length(const Q& q, int lp) { return q.length(lp); }
template <class Q>
__device__ __forceinline__ void queue_clear(const Q& q, int lp) { q.clear(lp); }
template <class Q>
__device__ __forceinline__ void queue_insert_or_update(const Q& q, const Event& item, int lp) {
  q.insert_or_update(item, lp);
}

// The queue kind whose view struct is View.
template <class View> struct view_kind;
template <> struct view_kind<LaqView> { static constexpr gpq_queue_kind value = GPQ_LAQ; };
template <> struct view_kind<LsqView> { static constexpr gpq_queue_kind value = GPQ_LSQ; };
template <> struct view_kind<LhqView> { static constexpr gpq_queue_kind value = GPQ_LHQ; };
template <> struct view_kind<HqView> { static constexpr gpq_queue_kind value = GPQ_HQ; };
template <> struct view_kind<GaqView> { static constexpr gpq_queue_kind value = GPQ_GAQ; };

// Host side: q's view as View; GPQ_ERR_INVALID when q is of another kind.
template <class View>
inline gpq_status get_view(gpq_queue* q, View* out) {
  gpq_queue_kind kind = GPQ_LAQ;
  View v;
  const gpq_status s = gpq_queue_device_view(q, &kind, &v, sizeof(View));
  if (s != GPQ_OK) return s;
  if (kind != view_kind<View>::value) return GPQ_ERR_INVALID;
  *out = v;
  return GPQ_OK;
}

}  // namespace gpq

#endif

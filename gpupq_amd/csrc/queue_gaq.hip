// queue_gaq.hip -- global array queue.
//   queue_pre : one safe event per LP selected from the sorted array (parallel prefix
//               search + per-LP atomicMin + stable compaction), where the reference runs
//               a single-thread scan (mark_safe_events, src/global_array_queue.cu:147-179).
//   queue_post: the step's inserts are sorted (tile sort + merge passes) and merged with
//               the array minus the selected records into the other buffer; the reference
//               appends them and mergesorts the whole array (src/global_array_queue.cu:62-74,
//               229-262).  Both are device-resident: no host reads between kernels.
#include <algorithm>

#include "queues.h"

namespace gpq {
namespace {

__global__ void gaq_init_kernel(GaqState* st, Event* fel0, Event* fel1, uint32_t* first, uint32_t num_lps,
                                WindowState* ws) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < num_lps) first[i] = 0xFFFFFFFFu;
  if (i == 0) {
    st->fel[0] = fel0;
    st->fel[1] = fel1;
    st->front = fel0;
    st->cur = 0;
    st->n = 0;
    st->m = 0;
    st->prefix = 0;
    st->deq_n = 0;
    st->rest_n = 0;
    st->selected = 0;
    st->n_next = 0;
    *ws = WindowState{kInfBits, 0u, 0u, 0u, 0ull, 0u, 0u};
  }
}

__global__ void gaq_mark_selected_kernel(GaqState* st) { st->selected = 1; }

// Builds the merge task: (unselected prefix ++ rest of the array) with the sorted inserts.
__global__ void gaq_plan_kernel(GaqState* st, MergeTask* task, const Event* sorted_ins, Event* rest, uint32_t cap,
                                uint32_t enq_cap, WindowState* ws) {
  const uint32_t m = st->m < enq_cap ? st->m : enq_cap;
  Event* live = st->fel[st->cur];
  MergeTask t{};
  if (st->selected) {
    t.a1 = rest;
    t.n_a1 = st->rest_n;
    t.a2 = live + st->prefix;
    t.n_a2 = st->n - st->prefix;
  } else {
    t.a1 = live;
    t.n_a1 = st->n;
    t.a2 = live;
    t.n_a2 = 0;
  }
  t.b = sorted_ins;
  t.n_b = m;
  uint64_t total = (uint64_t)t.n_a1 + t.n_a2 + t.n_b;
  if (total > cap) {
    atomicOr(&ws->error, kErrQueueOverflow);
    t.n_b = 0;                                    // keep the array, drop the inserts
    total = (uint64_t)t.n_a1 + t.n_a2;
  }
  t.out[0] = t.out[1] = t.out[2] = st->fel[st->cur ^ 1u];
  t.end[0] = t.end[1] = t.end[2] = (uint32_t)total;
  *task = t;
  st->n_next = (uint32_t)total;
  st->m = 0;
  st->deq_n = 0;
  st->selected = 0;
}

// Swaps the live buffer and publishes gmin = first record, wend, done.
__global__ void gaq_finish_kernel(GaqState* st, WindowState* ws, float lookahead, float horizon) {
  st->cur ^= 1u;
  st->n = st->n_next;
  const Event* live = st->fel[st->cur];
  st->front = live;
  const uint32_t g = st->n ? live[0].ts : kInfBits;
  float wend = bits_to_f(g) + lookahead;
  if (horizon < wend) wend = horizon;
  ws->gmin = g;
  ws->wend = f_to_bits(wend);
  ws->done = (g == kInfBits || !(bits_to_f(g) < horizon)) ? 1u : 0u;
  ws->windows += 1;
}

}  // namespace

gpq_status GaqQueue::init() {
  const uint64_t cap = cfg.total_capacity ? cfg.total_capacity : (uint64_t)cfg.num_lps * cfg.capacity_per_lp;
  GPQ_CHECK_ARG(cap >= 1 && cap < (1ull << 31), "GAQ: total capacity must be in [1, 2^31)");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1, "GAQ: stage_per_lp must be >= 1");
  const uint64_t enq_cap = (uint64_t)cfg.num_lps * cfg.stage_per_lp;
  GPQ_CHECK_ARG(enq_cap < (1ull << 31), "GAQ: num_lps * stage_per_lp must be < 2^31");
  cap_ = (uint32_t)cap;
  tile_ = cfg.sort_tile ? cfg.sort_tile : kSortTile;
  GPQ_CHECK_ARG(valid_sort_tile(tile_), "GAQ: sort_tile must be 0, 1024, 2048 or 4096");
  v_.num_lps = cfg.num_lps;
  v_.enq_cap = (uint32_t)enq_cap;
  v_.seg_cap = insert_stripe_cap(enq_cap);
  GPQ_TRY(arena.alloc(&fel_[0], cap));
  GPQ_TRY(arena.alloc(&fel_[1], cap));
  GPQ_TRY(arena.alloc(&rest_, cap));
  GPQ_TRY(arena.alloc(&v_.enq, (size_t)kInsertStripes * v_.seg_cap + enq_cap));
  GPQ_TRY(arena.alloc(&v_.seg_m, kStripeCounterWords));
  GPQ_TRY(arena.alloc(&seg_pre_, kInsertStripes + 2));
  GPQ_HIP(hipMemsetAsync(v_.seg_m, 0, sizeof(uint32_t) * kStripeCounterWords, stream));
  GPQ_TRY(arena.alloc(&sort_[0], enq_cap));
  GPQ_TRY(arena.alloc(&sort_[1], enq_cap));
  GPQ_TRY(arena.alloc(&v_.deq, cfg.num_lps));
  GPQ_TRY(arena.alloc(&v_.st, 1));
  GPQ_TRY(arena.alloc(&task_, 1));
  split_stride_ = merge_split_words(cap, tile_);
  GPQ_TRY(arena.alloc(&splits_, split_stride_));
  GPQ_TRY(arena.alloc(&first_, cfg.num_lps));
  GPQ_TRY(arena.alloc(&scratch_, 2 * (size_t)ceil_div(cap, 1024) + 2));
  v_.ws = ws;
  partial_slots = 1;
  GPQ_TRY(arena.alloc(&partial, 1));
  hipLaunchKernelGGL(gaq_init_kernel, dim3(ceil_div(cfg.num_lps, 256)), dim3(256), 0, stream, v_.st, fel_[0], fel_[1],
                     first_, cfg.num_lps, ws);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status GaqQueue::pre() {
  GaqState* st = v_.st;
  const SelectBufs b{first_, scratch_, &st->prefix, &st->deq_n, &st->rest_n};
  GPQ_TRY(select_safe(&st->front, &st->n, cap_, &ws->wend, nullptr, nullptr, b, v_.deq, rest_, stream,
                      acct(kStatSelected)));
  hipLaunchKernelGGL(gaq_mark_selected_kernel, dim3(1), dim3(1), 0, stream, st);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status GaqQueue::post() {
  Event* sorted = nullptr;
  GPQ_TRY(gather_inserts(v_.seg_m, v_.seg_cap, v_.enq_cap, v_.enq_cap, seg_pre_, &v_.st->m, ws, stream));
  GPQ_TRY(sort_records(v_.enq, sort_[0], sort_[1], &v_.st->m, v_.enq_cap, stream, &sorted, tile_,
                       acct(kStatSorted), seg_pre_, v_.seg_cap));
  hipLaunchKernelGGL(gaq_plan_kernel, dim3(1), dim3(1), 0, stream, v_.st, task_, sorted, rest_, cap_, v_.enq_cap, ws);
  GPQ_HIP(hipGetLastError());
  GPQ_TRY(merge_tasks(task_, 1, (uint64_t)cap_, stream, tile_, acct(kStatMerged), splits_, split_stride_));
  hipLaunchKernelGGL(gaq_finish_kernel, dim3(1), dim3(1), 0, stream, v_.st, ws, cfg.lookahead, cfg.horizon);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status GaqQueue::snapshot(std::vector<Event>* out) {
  GaqState st;
  GPQ_HIP(hipMemcpyAsync(&st, v_.st, sizeof(st), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  out->resize(st.n);
  if (st.n)
    GPQ_HIP(hipMemcpyAsync(out->data(), st.fel[st.cur], sizeof(Event) * st.n, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status GaqQueue::pending(uint64_t* n) {
  GaqState st;
  GPQ_HIP(hipMemcpyAsync(&st, v_.st, sizeof(st), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *n = st.n - (st.selected ? st.deq_n : 0u);   // records handed out by queue_pre are gone
  return GPQ_OK;
}

gpq_status GaqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  std::vector<Event> all;
  GPQ_TRY(snapshot(&all));
  *n = (uint32_t)std::count_if(all.begin(), all.end(), [lp](const Event& e) { return e.dst == lp; });
  return GPQ_OK;
}

gpq_status GaqQueue::clear_lp(uint32_t) {
  set_error("GAQ: per-LP clear is not provided (the reference's GAQ has no queue_clear)");
  return GPQ_ERR_UNSUPPORTED;
}

gpq_status GaqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  std::vector<Event> all;
  GPQ_TRY(snapshot(&all));
  uint32_t k = 0;
  for (const Event& e : all)
    if (e.dst == lp) {
      if (k < cap) out[k] = e;
      ++k;
    }
  *n = k;
  return GPQ_OK;
}

}  // namespace gpq

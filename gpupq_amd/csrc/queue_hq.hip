// queue_hq.hip -- global parallel heap with r-record nodes (Deo and Prasad, "Parallel
// heap: an optimal parallel priority queue", J. Supercomputing 6, 1992), restated for a
// device-resident schedule.  The reference (src/heap_queue.cu) drives each node merge
// from host pthreads with device-side launches, cudaStreamSynchronize and D2H copies per
// level; here every step is a fixed launch sequence whose kernels read the sizes and
// node choices from device memory, written by one-thread planner kernels.
//
// queue_post:
//  1. sort the step's inserts; M = merge(root, inserts)            (src/heap_queue.cu:1541-1553, 1448-1470)
//  2. publish the window from M[0]; select one safe record per LP of M with ts < window
//     end and key <= max(root) while nodes exist below the root -> dequeue buffer
//                                                                   (:478-547, 1560-1604)
//  3. R = M minus the selection: the first r records become the root, the rest are
//     carried down as insert chunks                                 (:418-466, 1097-1249)
//  4. a short root is refilled from the last node (copy_last, :975-1094)
//  5. delete-update, level by level: every dirty node whose largest record exceeds a
//     child's smallest merges with both children; it keeps the smallest records, the
//     child with the larger maximum (or the non-full right child) takes the largest,
//     the other the middle, and both children are checked at the next level
//                                                                   (:631-784)
//  6. insert-update: each chunk travels root -> target node, every node on the path
//     keeping its smallest records                                  (:817-960)
#include <algorithm>

#include "queues.h"

namespace gpq {
namespace {

__device__ __forceinline__ uint32_t depth_of(uint32_t v) { return 31u - __clz(v + 1u); }

__global__ void hq_init_kernel(HqState* st, uint32_t* counts, uint8_t* dirty, uint32_t nodes, uint32_t* first,
                               uint32_t num_lps, const Event* merged, WindowState* ws) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nodes) {
    counts[i] = 0;
    dirty[i] = 0;
  }
  if (i < num_lps) first[i] = 0xFFFFFFFFu;
  if (i == 0) {
    HqState s{};
    s.mptr = merged;
    *st = s;
    *ws = WindowState{kInfBits, 0u, 0u, 0u, 0ull, 0u, 0u};
  }
}

// Step 1: M = merge(root, sorted inserts); the selection bound is the root's maximum.
__global__ void hq_plan_merge_root(HqState* st, const uint32_t* counts, const Event* nodes, const Event* ins,
                                   Event* merged, uint32_t enq_cap, MergeTask* task) {
  const uint32_t m = st->m < enq_cap ? st->m : enq_cap;
  const uint32_t c0 = counts[0];
  MergeTask t{};
  t.a1 = nodes;
  t.n_a1 = c0;
  t.a2 = nodes;
  t.n_a2 = 0;
  t.b = ins;
  t.n_b = m;
  t.out[0] = t.out[1] = t.out[2] = merged;
  t.end[0] = t.end[1] = t.end[2] = c0 + m;
  *task = t;
  st->mcount = c0 + m;
  st->m = 0;
  st->bound_valid = st->last > 0 ? 1u : 0u;
  st->bound = c0 ? nodes[c0 - 1] : Event{0u, 0u, 0u, 0u};
}

// Step 2, window: the smallest record over every heap's merged buffer (heap h's at
// merged + h * mstride) starts it; the reference takes the minimum over its heaps' dequeue
// buffers (src/heap_queue.cu:1783-1806).
__global__ void hq_window_kernel(HqState* st, uint32_t heaps, const Event* merged, size_t mstride, WindowState* ws,
                                 float lookahead, float horizon, uint32_t select) {
  uint32_t g = kInfBits;
  for (uint32_t h = 0; h < heaps; ++h) {
    const uint32_t v = st[h].mcount ? merged[h * mstride].ts : kInfBits;
    g = v < g ? v : g;
    st[h].select = select;
    if (!select) {
      st[h].prefix = 0;
      st[h].deq_n = 0;
      st[h].rest_n = 0;
    }
  }
  float wend = bits_to_f(g) + lookahead;
  if (horizon < wend) wend = horizon;
  ws->gmin = g;
  ws->wend = f_to_bits(wend);
  ws->done = (g == kInfBits || !(bits_to_f(g) < horizon)) ? 1u : 0u;
  ws->windows += 1;
}

// Step 3: R = (unselected part of the scanned prefix) ++ M[prefix:]; root <- R[0:r), carry <- rest.
__global__ void hq_plan_root(HqState* st, uint32_t* counts, Event* nodes, const Event* merged, const Event* rest,
                             Event* carry, uint32_t r, MergeTask* task) {
  const uint32_t tail = st->mcount - st->prefix;
  const uint32_t rn = st->rest_n + tail;
  const uint32_t root_n = rn < r ? rn : r;
  MergeTask t{};
  t.a1 = rest;
  t.n_a1 = st->rest_n;
  t.a2 = merged + st->prefix;
  t.n_a2 = tail;
  t.b = merged;
  t.n_b = 0;
  t.out[0] = nodes;
  t.out[1] = carry;
  t.out[2] = carry;
  t.end[0] = root_n;
  t.end[1] = rn;
  t.end[2] = rn;
  *task = t;
  counts[0] = root_n;
  st->carry_n = rn - root_n;
}

// Step 4: copy_last -- records from the tail of the last node fill a short root.
__global__ void hq_plan_refill(HqState* st, uint32_t* counts, Event* nodes, uint8_t* dirty, uint32_t r, Event* tmp,
                               MergeTask* tasks) {
  MergeTask t{}, c{};
  const uint32_t c0 = counts[0];
  if (c0 < r && st->last > 0) {
    const uint32_t last = st->last;
    const uint32_t take = min(r - c0, counts[last]);
    t.a1 = nodes;
    t.n_a1 = c0;
    t.a2 = nodes;
    t.b = nodes + (size_t)last * r + (counts[last] - take);
    t.n_b = take;
    t.out[0] = t.out[1] = t.out[2] = tmp;
    t.end[0] = t.end[1] = t.end[2] = c0 + take;
    c.a1 = tmp;
    c.n_a1 = c0 + take;
    c.a2 = tmp;
    c.b = tmp;
    c.out[0] = c.out[1] = c.out[2] = nodes;
    c.end[0] = c.end[1] = c.end[2] = c0 + take;
    counts[0] = c0 + take;
    counts[last] -= take;
    uint32_t l = last;
    while (l > 0 && counts[l] == 0) --l;
    st->last = l;
  }
  dirty[0] = 1;
  tasks[0] = t;
  tasks[1] = c;
}

// Step 5, one level: a task slot per node of the level.
__global__ void hq_plan_level(HqState* st, const uint32_t* counts, Event* nodes, uint8_t* dirty, uint32_t r,
                              uint32_t num_nodes, uint32_t level, Event* tmp2, Event* tmp1, MergeTask* t1,
                              MergeTask* t2, MergeTask* t3) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t first = (1u << level) - 1u;
  const uint32_t v = first + s;
  if (s >= (1u << level)) return;
  MergeTask a{}, b{}, c{};   // slots past the last node launch as empty tasks
  if (v < num_nodes && dirty[v]) {
    dirty[v] = 0;
    const uint32_t l = 2u * v + 1u, rr = 2u * v + 2u;
    const uint32_t cl = l < num_nodes ? counts[l] : 0u;
    const uint32_t cr = rr < num_nodes ? counts[rr] : 0u;
    const uint32_t cv = counts[v];
    if (cv > 0 && cl + cr > 0) {
      const Event maxv = nodes[(size_t)v * r + cv - 1];
      const bool viol = (cl && key_less(nodes[(size_t)l * r], maxv)) || (cr && key_less(nodes[(size_t)rr * r], maxv));
      if (viol) {
        Event* T = tmp2 + (size_t)s * 2u * r;
        Event* V = tmp1 + (size_t)s * r;
        a.a1 = nodes + (size_t)l * r;
        a.n_a1 = cl;
        a.a2 = a.a1;
        a.b = nodes + (size_t)rr * r;
        a.n_b = cr;
        a.out[0] = a.out[1] = a.out[2] = T;
        a.end[0] = a.end[1] = a.end[2] = cl + cr;
        uint32_t big, mid;
        if (cr == 0) { big = l; mid = rr; }
        else if (cl == 0) { big = rr; mid = l; }
        else if (cr < r || key_less(nodes[(size_t)l * r + cl - 1], nodes[(size_t)rr * r + cr - 1])) { big = rr; mid = l; }
        else { big = l; mid = rr; }
        const uint32_t cmid = mid == l ? cl : cr;
        b.a1 = nodes + (size_t)v * r;
        b.n_a1 = cv;
        b.a2 = b.a1;
        b.b = T;
        b.n_b = cl + cr;
        b.out[0] = V;
        b.out[1] = nodes + (size_t)mid * r;
        b.out[2] = nodes + (size_t)big * r;
        b.end[0] = cv;
        b.end[1] = cv + cmid;
        b.end[2] = cv + cl + cr;
        c.a1 = V;
        c.n_a1 = cv;
        c.a2 = V;
        c.b = V;
        c.out[0] = c.out[1] = c.out[2] = nodes + (size_t)v * r;
        c.end[0] = c.end[1] = c.end[2] = cv;
        if (cl) dirty[l] = 1;
        if (cr) dirty[rr] = 1;
      }
    }
  }
  t1[s] = a;
  t2[s] = b;
  t3[s] = c;
}

// Step 6 set-up: carried records are cut into chunks, each targeting the last node's free
// space or a new node at the end of the heap.
__global__ void hq_plan_chunks(HqState* st, const uint32_t* counts, const Event* carry, Event* chunk_buf, uint32_t r,
                               uint32_t num_nodes, WindowState* ws) {
  uint32_t last = st->last;
  uint32_t cnt_last = counts[last];
  uint32_t off = 0;
  for (int c = 0; c < kHqMaxChunks; ++c) {
    st->chunk_n[c] = 0;
    st->chunk_target[c] = 0;
    if (off >= st->carry_n) continue;
    uint32_t t, room;
    if (last > 0 && cnt_last < r) {
      t = last;
      room = r - cnt_last;
    } else {
      t = last + 1u;
      room = r;
      if (t >= num_nodes) {
        atomicOr(&ws->error, kErrHeapOverflow);
        off = st->carry_n;
        continue;
      }
      last = t;
      cnt_last = 0;
    }
    const uint32_t n = min(room, st->carry_n - off);
    st->chunk_target[c] = t;
    st->chunk_n[c] = n;
    st->chunk_ptr[c] = carry + off;
    st->chunk_alt[c] = chunk_buf + (size_t)c * r;
    cnt_last += n;
    off += n;
  }
  if (off < st->carry_n) atomicOr(&ws->error, kErrHeapOverflow);
}

// Step 6, one level of one chunk's root-to-target path (the root is skipped: carried
// records are never smaller than the root's).
__global__ void hq_plan_insert_level(HqState* st, uint32_t* counts, Event* nodes, uint32_t r, uint32_t chunk,
                                     uint32_t level, Event* tmp, MergeTask* tasks) {
  MergeTask t{}, c{};
  const uint32_t n = st->chunk_n[chunk];
  const uint32_t target = st->chunk_target[chunk];
  if (n > 0 && level <= depth_of(target)) {
    const uint32_t dt = depth_of(target);
    const uint32_t p = ((target + 1u) >> (dt - level)) - 1u;
    const uint32_t cp = counts[p];
    Event* node = nodes + (size_t)p * r;
    t.a1 = node;
    t.n_a1 = cp;
    t.a2 = node;
    t.b = st->chunk_ptr[chunk];
    t.n_b = n;
    c.a1 = tmp;
    c.a2 = tmp;
    c.b = tmp;
    c.out[0] = c.out[1] = c.out[2] = node;
    if (level < dt) {                       // node keeps its smallest cp records, the rest moves on
      Event* alt = st->chunk_alt[chunk];
      t.out[0] = tmp;
      t.out[1] = t.out[2] = alt;
      t.end[0] = cp;
      t.end[1] = t.end[2] = cp + n;
      c.n_a1 = cp;
      c.end[0] = c.end[1] = c.end[2] = cp;
      st->chunk_alt[chunk] = const_cast<Event*>(st->chunk_ptr[chunk]);
      st->chunk_ptr[chunk] = alt;
    } else {                                // target: the chunk joins the node
      t.out[0] = t.out[1] = t.out[2] = tmp;
      t.end[0] = t.end[1] = t.end[2] = cp + n;
      c.n_a1 = cp + n;
      c.end[0] = c.end[1] = c.end[2] = cp + n;
      counts[target] = cp + n;
      if (target > st->last) st->last = target;
    }
  }
  tasks[0] = t;
  tasks[1] = c;
}

__global__ void hq_pending_kernel(const uint32_t* counts, uint32_t num_nodes, unsigned long long* out) {
  if (threadIdx.x || blockIdx.x) return;
  unsigned long long s = 0;
  for (uint32_t v = 0; v < num_nodes; ++v) s += counts[v];
  *out = s;
}

uint32_t levels_with_children(uint32_t nodes) {
  uint32_t lv = 0;
  while (((1u << lv) - 1u) < nodes && 2u * ((1u << lv) - 1u) + 1u < nodes) ++lv;
  return lv;   // levels 0 .. lv-1 contain a node with a child
}

}  // namespace

gpq_status HqQueue::init() {
  GPQ_CHECK_ARG(cfg.node_size >= 1, "HQ: node_size (r) must be >= 1");
  heaps_ = cfg.heaps ? cfg.heaps : 1u;
  GPQ_CHECK_ARG(heaps_ <= cfg.num_lps && heaps_ <= 1024, "HQ: heaps must be in [1, min(num_lps, 1024)]");
  lps_per_heap_ = (uint32_t)ceil_div(cfg.num_lps, heaps_);
  const uint64_t total = (uint64_t)lps_per_heap_ * (cfg.capacity_per_lp ? cfg.capacity_per_lp : 1);
  const uint64_t nodes = cfg.heap_nodes ? cfg.heap_nodes : (total + cfg.node_size - 1) / cfg.node_size + 1;
  GPQ_CHECK_ARG(nodes >= 1 && nodes < (1u << 20), "HQ: bad heap_nodes");
  r_ = cfg.node_size;
  tile_ = cfg.sort_tile ? cfg.sort_tile : kSortTile;
  GPQ_CHECK_ARG(valid_sort_tile(tile_), "HQ: sort_tile must be 0, 1024, 2048 or 4096");
  nodes_ = (uint32_t)nodes;
  // every heap takes a whole step's inserts (the events of one step may all go to one heap)
  const uint64_t enq_cap = (uint64_t)cfg.num_lps * (cfg.stage_per_lp ? cfg.stage_per_lp : 1);
  GPQ_CHECK_ARG(enq_cap + r_ < (1ull << 31), "HQ: num_lps * stage_per_lp + node_size must be < 2^31");
  GPQ_CHECK_ARG(enq_cap <= (uint64_t)(kHqMaxChunks - 1) * r_, "HQ: num_lps * stage_per_lp must be <= 7 * node_size");
  GPQ_CHECK_ARG((uint64_t)heaps_ * nodes_ * r_ < (1ull << 34), "HQ: heap too large");
  v_.num_lps = cfg.num_lps;
  v_.enq_cap = (uint32_t)enq_cap;
  v_.heaps = heaps_;
  v_.lps_per_heap = lps_per_heap_;
  levels_ = levels_with_children(nodes_);
  slots_ = levels_ ? (1u << (levels_ - 1)) : 1u;
  depth_ = 0;
  while ((1u << (depth_ + 1)) - 1u < nodes_) ++depth_;   // deepest level index of any node
  mcap_ = r_ + enq_cap;
  GPQ_TRY(arena.alloc(&nodes_buf_, (size_t)heaps_ * nodes_ * r_));
  GPQ_TRY(arena.alloc(&counts_, (size_t)heaps_ * nodes_));
  GPQ_TRY(arena.alloc(&dirty_, (size_t)heaps_ * nodes_));
  v_.seg_cap = insert_stripe_cap(enq_cap);
  v_.area = (uint64_t)kInsertStripes * v_.seg_cap + enq_cap;
  GPQ_TRY(arena.alloc(&v_.enq, (size_t)heaps_ * v_.area));
  GPQ_TRY(arena.alloc(&v_.seg_m, (size_t)heaps_ * kStripeCounterWords));
  GPQ_TRY(arena.alloc(&seg_pre_, (size_t)heaps_ * (kInsertStripes + 2)));
  GPQ_HIP(hipMemsetAsync(v_.seg_m, 0, sizeof(uint32_t) * heaps_ * kStripeCounterWords, stream));
  // working buffers, one set per heap so that the heaps' steps can run concurrently
  scratch_n_ = 2 * (size_t)ceil_div(mcap_, 1024) + 2;
  tasks_n_ = 3 * (size_t)slots_ + 8;
  GPQ_TRY(arena.alloc(&sort_[0], (size_t)heaps_ * enq_cap));
  GPQ_TRY(arena.alloc(&sort_[1], (size_t)heaps_ * enq_cap));
  GPQ_TRY(arena.alloc(&merged_, (size_t)heaps_ * mcap_));
  GPQ_TRY(arena.alloc(&rest_, (size_t)heaps_ * mcap_));
  GPQ_TRY(arena.alloc(&carry_, (size_t)heaps_ * mcap_));
  GPQ_TRY(arena.alloc(&chunk_buf_, (size_t)heaps_ * kHqMaxChunks * r_));
  GPQ_TRY(arena.alloc(&tmp2_, (size_t)heaps_ * slots_ * 2u * r_));
  GPQ_TRY(arena.alloc(&tmp1_, (size_t)heaps_ * (slots_ * (size_t)r_ + r_)));
  GPQ_TRY(arena.alloc(&v_.deq, (size_t)heaps_ * lps_per_heap_));
  GPQ_TRY(arena.alloc(&v_.st, heaps_));
  GPQ_TRY(arena.alloc(&first_, cfg.num_lps));   // by LP: the heaps' LP ranges are disjoint
  GPQ_TRY(arena.alloc(&scratch_, (size_t)heaps_ * scratch_n_));
  GPQ_TRY(arena.alloc(&tasks_, (size_t)heaps_ * tasks_n_));
  split_stride_ = merge_split_words(std::max<uint64_t>(mcap_, 3ull * r_), tile_);
  GPQ_TRY(arena.alloc(&splits_, (size_t)heaps_ * tasks_n_ * split_stride_));
  if (heaps_ > 1) {
    // one stream per heap (default priority), forked from and joined back into the queue's
    // stream twice per step; a hipGraph capture of the step records the branches
    hs_.assign(heaps_, nullptr);
    ev_.assign(2 * heaps_ + 2, nullptr);
    for (auto& x : hs_) GPQ_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    for (auto& e : ev_) GPQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  GPQ_TRY(arena.alloc(&sum_, 1));
  v_.ws = ws;
  partial_slots = 1;
  GPQ_TRY(arena.alloc(&partial, 1));
  const uint32_t n = std::max(nodes_, cfg.num_lps);
  for (uint32_t h = 0; h < heaps_; ++h) {
    hipLaunchKernelGGL(hq_init_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, v_.st + h,
                       counts_ + (size_t)h * nodes_, dirty_ + (size_t)h * nodes_, nodes_, first_, cfg.num_lps,
                       merged_ + (size_t)h * mcap_, ws);
    GPQ_HIP(hipGetLastError());
  }
  return GPQ_OK;
}

HqQueue::~HqQueue() {
  if (stream) (void)hipStreamSynchronize(stream);   // the heap streams join it every step
  for (auto e : ev_) if (e) (void)hipEventDestroy(e);
  for (auto x : hs_) if (x) (void)hipStreamDestroy(x);
}

// One post over every heap: the merges of inserts into the roots, the global window, then per
// heap the selection and the heap maintenance.  With several heaps each heap's part runs on
// its own stream (the reference runs one host pthread per heap, src/heap_queue.cu:1721-1757).
gpq_status HqQueue::step(bool select) {
  if (heaps_ == 1) {
    GPQ_TRY(merge_root(0, stream));
    hipLaunchKernelGGL(hq_window_kernel, dim3(1), dim3(1), 0, stream, v_.st, 1u, merged_, (size_t)mcap_, ws,
                       cfg.lookahead, cfg.horizon, select ? 1u : 0u);
    return step_heap(0, select, stream);
  }
  for (int phase = 0; phase < 2; ++phase) {
    hipEvent_t fork = ev_[2 * heaps_ + phase];
    GPQ_HIP(hipEventRecord(fork, stream));
    for (uint32_t h = 0; h < heaps_; ++h) {
      GPQ_HIP(hipStreamWaitEvent(hs_[h], fork, 0));
      GPQ_TRY(phase ? step_heap(h, select, hs_[h]) : merge_root(h, hs_[h]));
      hipEvent_t join = ev_[phase * heaps_ + h];
      GPQ_HIP(hipEventRecord(join, hs_[h]));
      GPQ_HIP(hipStreamWaitEvent(stream, join, 0));
    }
    if (!phase)
      hipLaunchKernelGGL(hq_window_kernel, dim3(1), dim3(1), 0, stream, v_.st, heaps_, merged_, (size_t)mcap_, ws,
                         cfg.lookahead, cfg.horizon, select ? 1u : 0u);
  }
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

// Step 1 for heap h: M = merge(root, sorted inserts).
gpq_status HqQueue::merge_root(uint32_t h, hipStream_t s) {
  HqState* st = v_.st + h;
  MergeTask* t = tasks_ + (size_t)h * tasks_n_;
  Event* sorted = nullptr;
  uint32_t* pre = seg_pre_ + (size_t)h * (kInsertStripes + 2);
  GPQ_TRY(gather_inserts(v_.seg_m + (size_t)h * kStripeCounterWords, v_.seg_cap, v_.enq_cap, v_.enq_cap, pre, &st->m,
                         ws, s));
  GPQ_TRY(sort_records(v_.enq + (size_t)h * v_.area, sort_[0] + (size_t)h * v_.enq_cap,
                       sort_[1] + (size_t)h * v_.enq_cap, &st->m, v_.enq_cap, s, &sorted, tile_,
                       acct(kStatSorted), pre, v_.seg_cap));
  hipLaunchKernelGGL(hq_plan_merge_root, dim3(1), dim3(1), 0, s, st, counts_ + (size_t)h * nodes_,
                     nodes_buf_ + (size_t)h * nodes_ * r_, sorted, merged_ + (size_t)h * mcap_, v_.enq_cap, t);
  GPQ_TRY(merge_tasks(t, 1, mcap_, s, tile_, acct(kStatMerged), splits_of(t), split_stride_));
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

// Steps 2 (selection) to 6 for heap h, after the window is published.
gpq_status HqQueue::step_heap(uint32_t h, bool select, hipStream_t s) {
  HqState* st = v_.st + h;
  uint32_t* counts = counts_ + (size_t)h * nodes_;
  uint8_t* dirty = dirty_ + (size_t)h * nodes_;
  Event* nodes = nodes_buf_ + (size_t)h * nodes_ * r_;
  Event* merged = merged_ + (size_t)h * mcap_;
  Event* rest = rest_ + (size_t)h * mcap_;
  Event* carry = carry_ + (size_t)h * mcap_;
  Event* chunk_buf = chunk_buf_ + (size_t)h * kHqMaxChunks * r_;
  Event* tmp2 = tmp2_ + (size_t)h * slots_ * 2u * r_;
  Event* tmp1 = tmp1_ + (size_t)h * (slots_ * (size_t)r_ + r_);
  MergeTask* t = tasks_ + (size_t)h * tasks_n_;
  if (select) {
    const SelectBufs b{first_, scratch_ + (size_t)h * scratch_n_, &st->prefix, &st->deq_n, &st->rest_n};
    GPQ_TRY(select_safe(&st->mptr, &st->mcount, (uint32_t)mcap_, &ws->wend, &st->bound, &st->bound_valid, b,
                        v_.deq + (size_t)h * lps_per_heap_, rest, s, acct(kStatSelected)));
  }
  hipLaunchKernelGGL(hq_plan_root, dim3(1), dim3(1), 0, s, st, counts, nodes, merged, rest, carry, r_, t);
  GPQ_TRY(merge_tasks(t, 1, mcap_, s, tile_, acct(kStatMerged), splits_of(t), split_stride_));
  hipLaunchKernelGGL(hq_plan_refill, dim3(1), dim3(1), 0, s, st, counts, nodes, dirty, r_, tmp1 + (size_t)slots_ * r_, t);
  GPQ_TRY(merge_tasks(t, 1, r_, s, tile_, acct(kStatMerged), splits_of(t), split_stride_));
  GPQ_TRY(merge_tasks(t + 1, 1, r_, s, tile_, acct(kStatMerged), splits_of(t + 1), split_stride_));
  MergeTask* t1 = t;
  MergeTask* t2 = t + slots_;
  MergeTask* t3 = t + 2 * (size_t)slots_;
  for (uint32_t lv = 0; lv < levels_; ++lv) {
    const uint32_t ns = 1u << lv;
    hipLaunchKernelGGL(hq_plan_level, dim3(ceil_div(ns, 256)), dim3(256), 0, s, st, counts, nodes, dirty, r_, nodes_,
                       lv, tmp2, tmp1, t1, t2, t3);
    GPQ_TRY(merge_tasks(t1, ns, 2ull * r_, s, tile_, acct(kStatMerged), splits_of(t1), split_stride_));
    GPQ_TRY(merge_tasks(t2, ns, 3ull * r_, s, tile_, acct(kStatMerged), splits_of(t2), split_stride_));
    GPQ_TRY(merge_tasks(t3, ns, r_, s, tile_, acct(kStatMerged), splits_of(t3), split_stride_));
  }
  hipLaunchKernelGGL(hq_plan_chunks, dim3(1), dim3(1), 0, s, st, counts, carry, chunk_buf, r_, nodes_, ws);
  const uint32_t chunks = std::min<uint32_t>(kHqMaxChunks, ceil_div(v_.enq_cap, r_) + 1);
  for (uint32_t c = 0; c < chunks; ++c) {
    for (uint32_t lv = 1; lv <= depth_; ++lv) {
      hipLaunchKernelGGL(hq_plan_insert_level, dim3(1), dim3(1), 0, s, st, counts, nodes, r_, c, lv, tmp1, t);
      GPQ_TRY(merge_tasks(t, 1, 2ull * r_, s, tile_, acct(kStatMerged), splits_of(t), split_stride_));
      GPQ_TRY(merge_tasks(t + 1, 1, 2ull * r_, s, tile_, acct(kStatMerged), splits_of(t + 1), split_stride_));
    }
  }
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status HqQueue::post() { return step(true); }
gpq_status HqQueue::post_init() { return step(false); }

gpq_status HqQueue::pending(uint64_t* n) {
  hipLaunchKernelGGL(hq_pending_kernel, dim3(1), dim3(1), 0, stream, counts_, heaps_ * nodes_,
                     reinterpret_cast<unsigned long long*>(sum_));
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(n, sum_, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status HqQueue::snapshot(std::vector<Event>* out, std::vector<uint32_t>* counts) {
  const uint32_t all = heaps_ * nodes_;   // node v of heap h at h * nodes_ + v, as in nodes_buf_
  counts->resize(all);
  GPQ_HIP(hipMemcpyAsync(counts->data(), counts_, sizeof(uint32_t) * all, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  out->clear();
  for (uint32_t v = 0; v < all; ++v) {
    const uint32_t c = (*counts)[v];
    if (!c) continue;
    const size_t at = out->size();
    out->resize(at + c);
    GPQ_HIP(hipMemcpyAsync(out->data() + at, nodes_buf_ + (size_t)v * r_, sizeof(Event) * c, hipMemcpyDeviceToHost,
                           stream));
  }
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status HqQueue::node_contents(uint32_t v, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(v < heaps_ * nodes_, "node out of range");
  uint32_t c = 0;
  GPQ_HIP(hipMemcpyAsync(&c, counts_ + v, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *n = c;
  const uint32_t k = std::min(c, cap);
  if (k) {
    GPQ_HIP(hipMemcpyAsync(out, nodes_buf_ + (size_t)v * r_, sizeof(Event) * k, hipMemcpyDeviceToHost, stream));
    GPQ_HIP(hipStreamSynchronize(stream));
  }
  return GPQ_OK;
}

gpq_status HqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  std::vector<Event> all;
  std::vector<uint32_t> counts;
  GPQ_TRY(snapshot(&all, &counts));
  *n = (uint32_t)std::count_if(all.begin(), all.end(), [lp](const Event& e) { return e.dst == lp; });
  return GPQ_OK;
}

gpq_status HqQueue::clear_lp(uint32_t) {
  set_error("HQ: per-LP clear is not provided for the shared heap");
  return GPQ_ERR_UNSUPPORTED;
}

gpq_status HqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  std::vector<Event> all;
  std::vector<uint32_t> counts;
  GPQ_TRY(snapshot(&all, &counts));
  std::vector<Event> mine;
  for (const Event& e : all)
    if (e.dst == lp) mine.push_back(e);
  std::sort(mine.begin(), mine.end(), [](const Event& a, const Event& b) {
    if (a.ts != b.ts) return a.ts < b.ts;
    if (a.src != b.src) return a.src < b.src;
    return a.seq < b.seq;
  });
  *n = (uint32_t)mine.size();
  for (uint32_t i = 0; i < mine.size() && i < cap; ++i) out[i] = mine[i];
  return GPQ_OK;
}

}  // namespace gpq

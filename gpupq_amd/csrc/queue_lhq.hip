// queue_lhq.hip -- local heap queue: queue_post sifts each LP's staged events up into
// its binary heap (one thread per LP, as heapify_up in src/local_heap_queue.cu:216-233)
// and folds the roots into the window reduction in the same kernel.
#include <algorithm>

#include "queues.h"

namespace gpq {
namespace {

constexpr int kLhqBlock = 256;

__global__ __launch_bounds__(kLhqBlock) void lhq_post_kernel(LhqView q, uint32_t* partial) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t root = kInfBits;
  uint32_t inserted = 0, moves = 0;
  if (lp < q.num_lps) {
    const uint32_t staged = q.ins[lp];
    root = q.head[lp];
    if (staged) {
      const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
      uint32_t n = q.len[lp];
      if (n + k > q.cap) {
        atomicOr(&q.ws->error, kErrQueueOverflow);
      } else {
        for (uint32_t m = 0; m < k; ++m) {
          const Event x = load_event(&q.stage[(size_t)lp * q.stage_cap + m]);
          uint32_t i = n;
          while (i > 0) {
            const uint32_t p = (i - 1u) >> 1;
            const Event ep = load_event(q.slot(lp, p));
            if (!key_less(x, ep)) break;
            store_event(q.slot(lp, i), ep);
            i = p;
            ++moves;
          }
          store_event(q.slot(lp, i), x);
          if (i == 0) root = x.ts;
          ++n;
        }
        q.len[lp] = n;
        q.head[lp] = root;
        inserted = k;
      }
      q.ins[lp] = 0;
    }
  }
  if (q.stats) {
    stats_add(q.stats, 0, inserted);
    stats_add(q.stats, 1, moves);
    stats_add(q.stats, 2, inserted ? 1u : 0u);
    stats_add(q.stats, 3, lp < q.num_lps ? 1u : 0u);
  }
  block_min_to_partial(root, partial);
}

__global__ void lhq_init_kernel(LhqView q) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp < q.num_lps) {
    q.len[lp] = 0;
    q.ins[lp] = 0;
    q.head[lp] = kInfBits;
  }
  if (lp == 0) *q.ws = WindowState{kInfBits, 0u, 0u, 0u, 0ull, 0u, 0u};
}

__global__ void lhq_clear_kernel(LhqView q, uint32_t lp) { q.clear((int)lp); }

}  // namespace

gpq_status LhqQueue::init() {
  GPQ_CHECK_ARG(cfg.capacity_per_lp >= 1, "LHQ: capacity_per_lp must be >= 1");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LHQ: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG((uint64_t)cfg.num_lps * cfg.capacity_per_lp < (1ull << 32), "LHQ: too many slots");
  const uint32_t L = cfg.num_lps;
  v_.num_lps = L;
  v_.cap = cfg.capacity_per_lp;
  v_.stage_cap = cfg.stage_per_lp;
  GPQ_TRY(arena.alloc(&v_.heap, (size_t)L * v_.cap));
  GPQ_TRY(arena.alloc(&v_.stage, (size_t)L * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.len, L));
  GPQ_TRY(arena.alloc(&v_.ins, L));
  GPQ_TRY(arena.alloc(&v_.head, L));
  GPQ_TRY(arena.alloc(&sum_, 1));
  v_.ws = ws;
  v_.stats = stats;
  partial_slots = ceil_div(L, kLhqBlock);
  GPQ_TRY(arena.alloc(&partial, partial_slots));
  hipLaunchKernelGGL(lhq_init_kernel, dim3(ceil_div(L, 256)), dim3(256), 0, stream, v_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LhqQueue::post() {
  hipLaunchKernelGGL(lhq_post_kernel, dim3(partial_slots), dim3(kLhqBlock), 0, stream, v_, partial);
  GPQ_HIP(hipGetLastError());
  return publish_window();
}

gpq_status LhqQueue::pending(uint64_t* n) { return sum_u32(v_.len, cfg.num_lps, sum_, n); }

gpq_status LhqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  GPQ_HIP(hipMemcpyAsync(n, v_.len + lp, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status LhqQueue::clear_lp(uint32_t lp) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  hipLaunchKernelGGL(lhq_clear_kernel, dim3(1), dim3(1), 0, stream, v_, lp);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LhqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  uint32_t len = 0;
  GPQ_HIP(hipMemcpyAsync(&len, v_.len + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  std::vector<Event> items(len);
  for (uint32_t i = 0; i < len; ++i)
    GPQ_HIP(hipMemcpyAsync(&items[i], v_.heap + (size_t)i * cfg.num_lps + lp, sizeof(Event), hipMemcpyDeviceToHost,
                           stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  std::sort(items.begin(), items.end(), [](const Event& a, const Event& b) {
    if (a.ts != b.ts) return a.ts < b.ts;
    if (a.src != b.src) return a.src < b.src;
    return a.seq < b.seq;
  });
  *n = len;
  for (uint32_t i = 0; i < len && i < cap; ++i) out[i] = items[i];
  return GPQ_OK;
}

}  // namespace gpq

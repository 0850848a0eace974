// queue_lsq.hip -- local splay queue: queue_post splay-inserts each LP's staged events
// (one thread per LP, as insert_bulk in src/local_splay_queue.cu:435-454) and folds
// the LP minima into the window reduction in the same kernel.
#include <algorithm>

#include "queues.h"

namespace gpq {
namespace {

constexpr int kLsqBlock = 256;

__global__ __launch_bounds__(kLsqBlock) void lsq_post_kernel(LsqView q, uint32_t* partial) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t minimum = kInfBits;
  uint32_t inserted = 0, visits = 0;
  if (lp < q.num_lps) {
    minimum = q.head[lp];
    const uint32_t staged = q.ins[lp];
    if (staged) {
      const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
      LsqView::Tree t = q.tree(lp);
      uint32_t rt = q.root[lp];
      uint32_t fl = q.freelist[lp];
      uint32_t n = q.len[lp];
      for (uint32_t m = 0; m < k; ++m) {
        if (fl == kNil) {                                 // pool exhausted: keep the tree, flag it
          atomicOr(&q.ws->error, kErrQueueOverflow);
          break;
        }
        const Event x = load_event(&q.stage[(size_t)lp * q.stage_cap + m]);
        const uint32_t nn = fl;                           // malloc_ (src/local_splay_queue.cu:72-94)
        fl = t.L(nn);
        store_event(&t.it[nn], x);
        if (rt == kNil) {
          t.setL(nn, kNil);
          t.setR(nn, kNil);
        } else {                                          // insert (src/local_splay_queue.cu:158-202)
          rt = t.splay(x, rt);
          if (!key_less(t.key(rt), x)) {
            t.setL(nn, t.L(rt));
            t.setR(nn, rt);
            t.setL(rt, kNil);
          } else {
            t.setR(nn, t.R(rt));
            t.setL(nn, rt);
            t.setR(rt, kNil);
          }
        }
        rt = nn;
        ++n;
        ++inserted;
        if (x.ts < minimum) minimum = x.ts;
      }
      q.root[lp] = rt;
      q.freelist[lp] = fl;
      q.len[lp] = n;
      q.head[lp] = minimum;
      q.ins[lp] = 0;
      visits = t.visits;
    }
  }
  if (q.stats) {
    stats_add(q.stats, 0, inserted);
    stats_add(q.stats, 1, visits);
    stats_add(q.stats, 2, inserted ? 1u : 0u);
    stats_add(q.stats, 3, lp < q.num_lps ? 1u : 0u);
  }
  block_min_to_partial(minimum, partial);
}

// Rebuilds LP pools: every slot free, threaded in index order.
__global__ void lsq_reset_kernel(LsqView q, uint32_t lp_first, uint32_t lp_count) {
  const uint64_t total = (uint64_t)lp_count * q.cap;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t lp = lp_first + (uint32_t)(i / q.cap);
    const uint32_t node = (uint32_t)(i % q.cap);
    q.link[(size_t)lp * q.cap + node] = make_uint2(node + 1u < q.cap ? node + 1u : kNil, kNil);
    if (node == 0) {
      q.root[lp] = kNil;
      q.freelist[lp] = 0;
      q.len[lp] = 0;
      q.ins[lp] = 0;
      q.head[lp] = kInfBits;
    }
  }
}

__global__ void lsq_ws_init_kernel(WindowState* ws) { *ws = WindowState{kInfBits, 0u, 0u, 0u, 0ull, 0u, 0u}; }

}  // namespace

gpq_status LsqQueue::init() {
  GPQ_CHECK_ARG(cfg.capacity_per_lp >= 1 && cfg.capacity_per_lp < kNil, "LSQ: bad capacity_per_lp");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LSQ: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG((uint64_t)cfg.num_lps * cfg.capacity_per_lp < (1ull << 31),
                "LSQ: num_lps * capacity_per_lp must be < 2^31 (int item indices)");
  const uint32_t L = cfg.num_lps;
  v_.num_lps = L;
  v_.cap = cfg.capacity_per_lp;
  v_.stage_cap = cfg.stage_per_lp;
  GPQ_TRY(arena.alloc(&v_.item, (size_t)L * v_.cap));
  GPQ_TRY(arena.alloc(&v_.link, (size_t)L * v_.cap));
  GPQ_TRY(arena.alloc(&v_.stage, (size_t)L * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.root, L));
  GPQ_TRY(arena.alloc(&v_.freelist, L));
  GPQ_TRY(arena.alloc(&v_.len, L));
  GPQ_TRY(arena.alloc(&v_.ins, L));
  GPQ_TRY(arena.alloc(&v_.head, L));
  GPQ_TRY(arena.alloc(&sum_, 1));
  v_.ws = ws;
  v_.stats = stats;
  partial_slots = ceil_div(L, kLsqBlock);
  GPQ_TRY(arena.alloc(&partial, partial_slots));
  hipLaunchKernelGGL(lsq_reset_kernel, dim3(4096), dim3(256), 0, stream, v_, 0u, L);
  hipLaunchKernelGGL(lsq_ws_init_kernel, dim3(1), dim3(1), 0, stream, ws);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LsqQueue::post() {
  hipLaunchKernelGGL(lsq_post_kernel, dim3(partial_slots), dim3(kLsqBlock), 0, stream, v_, partial);
  GPQ_HIP(hipGetLastError());
  return publish_window();
}

gpq_status LsqQueue::pending(uint64_t* n) { return sum_u32(v_.len, cfg.num_lps, sum_, n); }

gpq_status LsqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  GPQ_HIP(hipMemcpyAsync(n, v_.len + lp, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

// queue_clear (src/local_splay_queue.cu:480-494): the LP's whole pool becomes free again.
gpq_status LsqQueue::clear_lp(uint32_t lp) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  hipLaunchKernelGGL(lsq_reset_kernel, dim3(ceil_div(v_.cap, 256)), dim3(256), 0, stream, v_, lp, 1u);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LsqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  uint32_t rt = kNil, len = 0;
  std::vector<Event> items(v_.cap);
  std::vector<uint2> links(v_.cap);
  GPQ_HIP(hipMemcpyAsync(&rt, v_.root + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(&len, v_.len + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(items.data(), v_.item + (size_t)lp * v_.cap, sizeof(Event) * v_.cap, hipMemcpyDeviceToHost,
                         stream));
  GPQ_HIP(hipMemcpyAsync(links.data(), v_.link + (size_t)lp * v_.cap, sizeof(uint2) * v_.cap, hipMemcpyDeviceToHost,
                         stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  std::vector<uint32_t> stack;   // in-order traversal = key order
  uint32_t t = rt, count = 0;
  while ((t != kNil || !stack.empty()) && count <= v_.cap) {
    while (t != kNil && stack.size() <= v_.cap) { stack.push_back(t); t = links[t].x; }
    if (stack.empty()) break;
    t = stack.back();
    stack.pop_back();
    if (count < cap) out[count] = items[t];
    ++count;
    t = links[t].y;
  }
  if (count != len) {
    set_error("LSQ: tree of LP " + std::to_string(lp) + " holds " + std::to_string(count) + " nodes, length " +
              std::to_string(len));
    return GPQ_ERR_STATE;
  }
  *n = len;
  return GPQ_OK;
}

}  // namespace gpq

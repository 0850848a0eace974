// queue_laq.hip -- local array queue: per-LP sorted rings, wave-per-LP merge of
// the staged inserts, fused with the reduction that publishes the next window.
//
// Reference: src/local_array_queue.cu.  The reference restores order after a step
// with a serial per-LP insertion sort (d_insertion_sort :254-299, sort.cu
// insert_sorted :329-380) or a block-per-LP parallel shift (:187-251), and finds
// the window minimum with a per-LP head scan plus a device-launched cub reduce
// (:116-146, :460-478).  Here one kernel does both: a block owns 256 LPs, its
// waves merge the LPs that received inserts (coalesced reads of the displaced
// tail, one shifted write per moved record), then the block folds the 256 heads
// and the last block publishes gmin / window end / done.
#include "queues.h"

namespace gpq {
namespace {

constexpr int kLaqBlock = 256;   // LPs per post block = threads per block (4 waves)

// Merges the staged events of LP `lp` into its sorted ring; one wave, lanes in lockstep.
// Old record j moves to j + c_j where c_j = number of new events with a smaller key;
// new event m lands at rank_m + (number of old records with a smaller key).  The
// displaced tail is processed from the back in 64-record chunks and the walk stops
// at the first chunk whose lowest record does not move, so only displaced records
// are read and written (plus at most one chunk of non-moving records).
// Returns the number of existing records moved (d), for the work counters.
__device__ uint32_t laq_merge_lp(const LaqView& q, uint32_t lp, int lane) {
  const uint32_t staged = q.ins[lp];
  const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
  const uint32_t n = q.len[lp];
  const uint32_t r = q.rd[lp];
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  if (n + k > mask + 1u) {
    if (lane == 0) {
      atomicOr(&q.ws->error, kErrQueueOverflow);
      q.ins[lp] = 0;
    }
    return 0;
  }
  Event e{kInfBits, 0u, 0u, 0u};
  if (lane < (int)k) e = load_event(&q.stage[(size_t)lp * q.stage_cap + lane]);
  uint32_t rank = 0;
  for (uint32_t m = 0; m < k; ++m) {
    const Event o = shfl_event(e, (int)m);
    if (lane < (int)k && (key_less(o, e) || (!key_less(e, o) && (int)m < lane))) ++rank;
  }
  uint32_t greater = 0;   // lane m < k: old records with a key above new event m
  uint32_t moved = 0;
  if (n > 0) {
    for (int base = (int)n - kWave;; base -= kWave) {
      const int j = base + lane;
      const bool valid = j >= 0;
      Event old{0u, 0u, 0u, 0u};
      if (valid) old = load_event(&ring[(r + (uint32_t)j) & mask]);
      uint32_t c = 0;
      for (uint32_t m = 0; m < k; ++m) {
        const Event nm = shfl_event(e, (int)m);
        const bool lt = valid && key_less(nm, old);
        c += lt ? 1u : 0u;
        const uint64_t b = __ballot(lt);
        if (lane == (int)m) greater += (uint32_t)__popcll(b);
      }
      if (valid && c) store_event(&ring[(r + (uint32_t)j + c) & mask], old);
      moved += (uint32_t)__popcll(__ballot(valid && c != 0u));
      const int low = base < 0 ? -base : 0;
      const uint32_t c_low = (uint32_t)__shfl((int)c, low, kWave);
      if (base <= 0 || c_low == 0) break;
    }
  }
  if (lane < (int)k) {
    const uint32_t pos = rank + (n - greater);
    store_event(&ring[(r + pos) & mask], e);
    if (pos == 0) q.head[lp] = e.ts;
  }
  if (lane == 0) {
    q.len[lp] = n + k;
    q.ins[lp] = 0;
  }
  return moved;
}

__global__ __launch_bounds__(kLaqBlock) void laq_post_kernel(LaqView q, uint32_t* partial, float lookahead,
                                                             float horizon) {
  __shared__ uint32_t s_work[kLaqBlock];
  __shared__ uint32_t s_count;
  __shared__ uint32_t s_inserted;
  __shared__ uint32_t s_moved;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
  }
  __syncthreads();
  const uint32_t staged = lp < q.num_lps ? q.ins[lp] : 0u;
  if (staged != 0u) s_work[atomicAdd(&s_count, 1u)] = lp;
  if (q.stats && staged) atomicAdd(&s_inserted, staged < q.stage_cap ? staged : q.stage_cap);
  __syncthreads();
  const int wave = (int)threadIdx.x / kWave;
  const int lane = lane_id();
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)wave; i < s_count; i += kLaqBlock / kWave) moved += laq_merge_lp(q, s_work[i], lane);
  if (q.stats && lane == 0 && moved) atomicAdd(&s_moved, moved);
  __syncthreads();
  if (q.stats && threadIdx.x == 0) {   // work counters: one device atomic per block and counter
    const uint32_t lps_here = min((uint32_t)kLaqBlock, q.num_lps - blockIdx.x * kLaqBlock);
    atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
    atomicAdd(&q.stats[1], (unsigned long long)s_moved);
    atomicAdd(&q.stats[2], (unsigned long long)s_count);
    atomicAdd(&q.stats[3], (unsigned long long)lps_here);
  }
  const uint32_t v = lp < q.num_lps ? q.head[lp] : kInfBits;
  window_reduce_publish(v, partial, q.ws, lookahead, horizon);
}

__global__ void laq_init_kernel(LaqView q) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp < q.num_lps) {
    q.rd[lp] = 0;
    q.len[lp] = 0;
    q.ins[lp] = 0;
    q.head[lp] = kInfBits;
  }
  if (lp == 0) {
    q.ws->gmin = kInfBits;
    q.ws->wend = 0;
    q.ws->done = 0;
    q.ws->blocks_arrived = 0;
    q.ws->windows = 0;
    q.ws->error = 0;
  }
}

__global__ void laq_pending_kernel(LaqView q, unsigned long long* sum) {
  __shared__ unsigned long long s[kLaqBlock / kWave];
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = lp < q.num_lps ? (unsigned long long)q.len[lp] : 0ull;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane_id() == 0) s[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) t += s[i];
    atomicAdd(sum, t);
  }
}

__global__ void laq_clear_kernel(LaqView q, uint32_t lp) { q.clear((int)lp); }

}  // namespace

gpq_status LaqQueue::init() {
  GPQ_CHECK_ARG(is_pow2(cfg.capacity_per_lp) && cfg.capacity_per_lp >= 2,
                "LAQ: capacity_per_lp must be a power of two >= 2");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LAQ: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG((uint64_t)cfg.num_lps * cfg.capacity_per_lp < (1ull << 31),
                "LAQ: num_lps * capacity_per_lp must be < 2^31 (int item indices, src/queue.cuh:51)");
  const uint32_t L = cfg.num_lps;
  v_.num_lps = L;
  v_.cap_log2 = log2u(cfg.capacity_per_lp);
  v_.stage_cap = cfg.stage_per_lp;
  GPQ_TRY(arena.alloc(&v_.ring, (size_t)L << v_.cap_log2));
  GPQ_TRY(arena.alloc(&v_.stage, (size_t)L * v_.stage_cap));
  GPQ_TRY(arena.alloc(&v_.rd, L));
  GPQ_TRY(arena.alloc(&v_.len, L));
  GPQ_TRY(arena.alloc(&v_.ins, L));
  GPQ_TRY(arena.alloc(&v_.head, L));
  GPQ_TRY(arena.alloc(&sum_, 1));
  v_.ws = ws;
  v_.stats = stats;
  partial_slots = ceil_div(L, kLaqBlock);
  GPQ_TRY(arena.alloc(&partial, partial_slots));
  hipLaunchKernelGGL(laq_init_kernel, dim3(ceil_div(L, 256)), dim3(256), 0, stream, v_);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::post() {
  hipLaunchKernelGGL(laq_post_kernel, dim3(partial_slots), dim3(kLaqBlock), 0, stream, v_, partial, cfg.lookahead,
                     cfg.horizon);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::pending(uint64_t* n) {
  GPQ_HIP(hipMemsetAsync(sum_, 0, sizeof(uint64_t), stream));
  hipLaunchKernelGGL(laq_pending_kernel, dim3(partial_slots), dim3(kLaqBlock), 0, stream, v_,
                     reinterpret_cast<unsigned long long*>(sum_));
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(n, sum_, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status LaqQueue::length(uint32_t lp, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  GPQ_HIP(hipMemcpyAsync(n, v_.len + lp, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status LaqQueue::clear_lp(uint32_t lp) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  hipLaunchKernelGGL(laq_clear_kernel, dim3(1), dim3(1), 0, stream, v_, lp);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) {
  GPQ_CHECK_ARG(lp < cfg.num_lps, "lp out of range");
  uint32_t rd = 0, len = 0;
  const uint32_t C = 1u << v_.cap_log2;
  std::vector<Event> ring(C);
  GPQ_HIP(hipMemcpyAsync(&rd, v_.rd + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(&len, v_.len + lp, 4, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(ring.data(), v_.ring + ((size_t)lp << v_.cap_log2), sizeof(Event) * C,
                         hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *n = len;
  for (uint32_t i = 0; i < len && i < cap; ++i) out[i] = ring[(rd + i) & (C - 1)];
  return GPQ_OK;
}

}  // namespace gpq

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
constexpr int kSlots = 8;        // ring records held per lane: a group of G lanes covers 8G records

struct Key3 {
  uint32_t ts, src, seq;
};

__device__ __forceinline__ Key3 shfl_key(const Event& e, int src_lane) {
  return Key3{(uint32_t)__shfl((int)e.ts, src_lane, kWave), (uint32_t)__shfl((int)e.src, src_lane, kWave),
              (uint32_t)__shfl((int)e.seq, src_lane, kWave)};
}

__device__ __forceinline__ bool key_less3(const Key3& a, const Event& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

__device__ __forceinline__ bool key_less3(const Key3& a, const Key3& b) {
  if (a.ts != b.ts) return a.ts < b.ts;
  if (a.src != b.src) return a.src < b.src;
  return a.seq < b.seq;
}

// Orders the LDS histogram updates of the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Merges the k staged events of one LP into its sorted run of n records at ring slot r,
// using a group of G lanes (lane g of the group, group starting at wave lane gbase).
// The run is loaded at once (record j in lane j % G, register slot j / G); staged events
// are taken G - 1 at a time (one per lane).  Old record j ends at final index j + c_j,
// c_j = number of new events with a smaller key, and c is nondecreasing in j; the new
// event of rank a lands at a + #{j : c_j <= a}.  A histogram of c over [0, k] (LDS, one
// bin per lane) gives both: its inclusive prefix places the new events, and bin a counts
// the old records that stay put when the run grows by a slots at the front and k - a at
// the back.  The a with the largest bin is taken, so the fewest records move: a = 0 is
// the reference's backward insertion (src/sort.cu:329-356), a = k its forward variant
// (src/sort.cu:358-379), chosen here per LP and step.  Returns the records moved.
template <int G>
__device__ uint32_t laq_merge_group(const LaqView& q, uint32_t lp, uint32_t k_total, uint32_t n, uint32_t r, int g,
                                    int gbase, uint32_t* hist) {
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  const Event* stg = q.stage + (size_t)lp * q.stage_cap;
  uint32_t moved = 0;
  for (uint32_t done = 0; done < k_total;) {
    if (done) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");   // previous pass's ring writes
    const uint32_t k = min(k_total - done, (uint32_t)G - 1u);
    Event e{kInfBits, 0u, 0u, 0u};
    if (g < (int)k) e = load_event(&stg[done + g]);
    const uint32_t ns = (n + G - 1) / G;
    Event old[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const uint32_t j = (uint32_t)(s * G + g);
      old[s] = Event{0u, 0u, 0u, 0u};
      if ((uint32_t)s < ns && j < n) old[s] = load_event(&ring[(r + j) & mask]);
    }
    // rank of each new event among the new events (keys are unique)
    uint32_t rank = 0;
    uint32_t c[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) c[s] = 0;
    for (uint32_t m = 0; m < k; ++m) {
      const Key3 nm = shfl_key(e, gbase + (int)m);
      rank += (g < (int)k && key_less3(nm, Key3{e.ts, e.src, e.seq})) ? 1u : 0u;
#pragma unroll
      for (int s = 0; s < kSlots; ++s)
        if ((uint32_t)s < ns) c[s] += ((uint32_t)(s * G + g) < n && key_less3(nm, old[s])) ? 1u : 0u;
    }
    // histogram of c over [0, k], k < G: bin b in lane b
    hist[g] = 0;
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if ((uint32_t)s < ns && (uint32_t)(s * G + g) < n) atomicAdd(&hist[c[s]], 1u);
    wave_lds_sync();
    const uint32_t h = g <= (int)k ? hist[g] : 0u;
    wave_lds_sync();
    // a = argmax over bins (smallest a on ties); inclusive prefix of the bins
    uint32_t best = g <= (int)k ? ((h << 8) | (255u - (uint32_t)g)) : 0u;
    uint32_t pre = h;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)best, off, kWave);
      best = o > best ? o : best;
      const uint32_t t = (uint32_t)__shfl((int)pre, gbase + (g >= off ? g - off : g), kWave);
      if (g >= off) pre += t;
    }
    const uint32_t best_a = 255u - (best & 255u), best_stay = best >> 8;
    const uint32_t below = (uint32_t)__shfl((int)pre, gbase + (int)(rank < k ? rank : 0u), kWave);
    const uint32_t r_new = (r - best_a) & mask;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const uint32_t j = (uint32_t)(s * G + g);
      if ((uint32_t)s < ns && j < n && c[s] != best_a) store_event(&ring[(r_new + j + c[s]) & mask], old[s]);
    }
    if (g < (int)k) {
      const uint32_t f = rank + below;
      store_event(&ring[(r_new + f) & mask], e);
      if (f == 0) q.head[lp] = e.ts;
    }
    moved += n - best_stay;
    n += k;
    r = r_new;
    done += k;
  }
  if (g == 0) {
    q.rd[lp] = r;
    q.len[lp] = n;
    q.ins[lp] = 0;
  }
  return moved;
}

// One block owns kLaqBlock LPs.  Phase 1 reads their counters coalesced into LDS and
// lists the LPs that received inserts; phase 2 merges them, one LP per group of G
// lanes (kLaqBlock / G LPs in flight per block); phase 3 folds the block's heads into
// the window reduction.
template <int G>
__global__ __launch_bounds__(kLaqBlock, 5) void laq_post_kernel(LaqView q, uint32_t* partial) {
  __shared__ uint32_t s_work[kLaqBlock];
  __shared__ uint32_t s_k[kLaqBlock], s_n[kLaqBlock], s_r[kLaqBlock];
  __shared__ uint32_t s_hist[kLaqBlock];
  __shared__ uint32_t s_count;
  __shared__ uint32_t s_inserted;
  __shared__ uint32_t s_moved;
  const uint32_t lp = blockIdx.x * kLaqBlock + threadIdx.x;
  if (q.block_live && !q.block_live[blockIdx.x]) {     // a block of unused LPs: nothing staged, no heads
    if (threadIdx.x == 0) partial[blockIdx.x] = kInfBits;
    return;
  }
  if (threadIdx.x == 0) {
    s_count = 0;
    s_inserted = 0;
    s_moved = 0;
  }
  __syncthreads();
  uint32_t staged = 0;
  if (lp < q.num_lps) {
    staged = q.ins[lp];
    if (staged) {
      const uint32_t k = staged < q.stage_cap ? staged : q.stage_cap;
      const uint32_t n = q.len[lp];
      if (n + k > (1u << q.cap_log2)) {
        atomicOr(&q.ws->error, kErrQueueOverflow);
        q.ins[lp] = 0;
      } else {
        const uint32_t slot = atomicAdd(&s_count, 1u);
        s_work[slot] = threadIdx.x;
        s_k[threadIdx.x] = k;
        s_n[threadIdx.x] = n;
        s_r[threadIdx.x] = q.rd[lp];
        if (q.stats) atomicAdd(&s_inserted, k);
      }
    }
  }
  __syncthreads();
  const int lane = lane_id();
  const int g = lane % G;
  const int gbase = lane - g;
  const int group = (int)threadIdx.x / G;
  uint32_t moved = 0;
  for (uint32_t i = (uint32_t)group; i < s_count; i += kLaqBlock / G) {
    const uint32_t t = s_work[i];
    moved += laq_merge_group<G>(q, blockIdx.x * kLaqBlock + t, s_k[t], s_n[t], s_r[t], g, gbase, s_hist + group * G);
  }
  if (q.stats && g == 0 && moved) atomicAdd(&s_moved, moved);
  __syncthreads();
  if (q.stats && threadIdx.x == 0) {   // work counters: one device atomic per block and counter
    const uint32_t lps_here = min((uint32_t)kLaqBlock, q.num_lps - blockIdx.x * kLaqBlock);
    atomicAdd(&q.stats[0], (unsigned long long)s_inserted);
    atomicAdd(&q.stats[1], (unsigned long long)s_moved);
    atomicAdd(&q.stats[2], (unsigned long long)s_count);
    atomicAdd(&q.stats[3], (unsigned long long)lps_here);
  }
  const uint32_t v = lp < q.num_lps ? q.head[lp] : kInfBits;
  block_min_to_partial(v, partial);
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
    q.ws->reserved = 0;
    q.ws->windows = 0;
    q.ws->error = 0;
  }
}

__global__ void laq_clear_kernel(LaqView q, uint32_t lp) { q.clear((int)lp); }

}  // namespace

gpq_status LaqQueue::init() {
  GPQ_CHECK_ARG(is_pow2(cfg.capacity_per_lp) && cfg.capacity_per_lp >= 2,
                "LAQ: capacity_per_lp must be a power of two >= 2");
  GPQ_CHECK_ARG(cfg.stage_per_lp >= 1 && cfg.stage_per_lp <= 64, "LAQ: stage_per_lp must be in [1, 64]");
  GPQ_CHECK_ARG(cfg.capacity_per_lp <= 8u * 64u, "LAQ: capacity_per_lp must be <= 512");
  GPQ_CHECK_ARG((uint64_t)cfg.num_lps * cfg.capacity_per_lp < (1ull << 31),
                "LAQ: num_lps * capacity_per_lp must be < 2^31 (int item indices, src/queue.cuh:51)");
  const uint32_t L = cfg.num_lps;
  v_.num_lps = L;
  v_.cap_log2 = log2u(cfg.capacity_per_lp);
  v_.stage_cap = cfg.stage_per_lp;
  // lanes per LP merge: enough to hold the run (kSlots records per lane); staged events are
  // merged G - 1 at a time
  group_ = 16;
  while (group_ < 64 && group_ * (uint32_t)kSlots < cfg.capacity_per_lp) group_ *= 2;
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
  GPQ_TRY(merge_only());
  return publish_window();
}

gpq_status LaqQueue::merge_only() {
  static_assert(kLaqBlock == (int)LaqQueue::kBlockLps, "LAQ post block size");
  const dim3 grid(partial_slots), block(kLaqBlock);
  if (group_ == 16)
    hipLaunchKernelGGL(laq_post_kernel<16>, grid, block, 0, stream, v_, partial);
  else if (group_ == 32)
    hipLaunchKernelGGL(laq_post_kernel<32>, grid, block, 0, stream, v_, partial);
  else
    hipLaunchKernelGGL(laq_post_kernel<64>, grid, block, 0, stream, v_, partial);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status LaqQueue::pending(uint64_t* n) { return sum_u32(v_.len, cfg.num_lps, sum_, n); }

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

// sortprim.hip -- record sort (bitonic tile sort in LDS + merge-path merge passes),
// multi-task merge-path merges, and the parallel one-safe-event-per-LP selection.
#include "sortprim.h"

namespace gpq {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = kMergeTile / kThreads;   // 8
constexpr uint32_t kSelectTile = 1024;              // records per selection block (256 x 4)

__device__ __forceinline__ Event sentinel() { return Event{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}; }

// A = a1[0:n1) ++ a2[...]
struct ASrc {
  const Event* a1;
  const Event* a2;
  uint32_t n1;
  __device__ __forceinline__ Event get(uint32_t i) const { return i < n1 ? load_event(a1 + i) : load_event(a2 + (i - n1)); }
};

struct OutMap {
  Event* out[3];
  uint32_t end[3];
  __device__ __forceinline__ Event* at(uint32_t o) const {
    if (o < end[0]) return out[0] + o;
    if (o < end[1]) return out[1] + (o - end[0]);
    return out[2] + (o - end[1]);
  }
};

// Merge-path split: the number of A records among the first d outputs of the stable merge
// (ties: A first) is the smallest i in [max(0, d - nb), min(d, na)] with B[d-1-i] < A[i].
// Found for the block's two diagonals d0 and d1 at once, each by half of the block: every
// round the 128 threads of a half probe 128 evenly spaced points of the remaining range and
// keep the gap where the predicate turns true, so a range of n shrinks to nothing in
// ceil(log_128 n) rounds of one load each, not log_2 n dependent loads.
__device__ __forceinline__ void block_merge_splits(const ASrc& A, uint32_t na, const Event* B, uint32_t nb, uint32_t d0, uint32_t d1,
                                   uint32_t* s_split) {
  constexpr uint32_t kProbes = kThreads / 2;
  __shared__ uint32_t s_lo[2], s_hi[2], s_first[2];
  const uint32_t h = threadIdx.x / kProbes, k = threadIdx.x % kProbes;
  const uint32_t d = h ? d1 : d0;
  if (k == 0) {
    s_lo[h] = d > nb ? d - nb : 0u;
    s_hi[h] = d < na ? d : na;
  }
  __syncthreads();
  while (s_lo[0] < s_hi[0] || s_lo[1] < s_hi[1]) {   // uniform: read after a barrier
    const uint32_t lo = s_lo[h], hi = s_hi[h];
    const uint32_t step = (hi - lo + kProbes - 1) / kProbes;
    if (k == 0) s_first[h] = 0xFFFFFFFFu;
    __syncthreads();
    if (lo < hi) {
      const uint32_t i = lo + k * step;
      if (i < hi && key_less(load_event(B + (d - 1u - i)), A.get(i))) atomicMin(&s_first[h], k);
    }
    __syncthreads();
    if (k == 0 && lo < hi) {
      const uint32_t f = s_first[h];
      if (f == 0xFFFFFFFFu) {
        s_lo[h] = lo + (hi - lo - 1u) / step * step + 1u;   // past the last probe
      } else {
        s_hi[h] = lo + f * step;
        if (f) s_lo[h] = lo + (f - 1u) * step + 1u;
      }
    }
    __syncthreads();
  }
  if (k == 0) s_split[h] = s_lo[h];
  __syncthreads();
}

// Merges output ranks [t0, t0 + kMergeTile) of merge(A, B) through LDS: the block's input
// ranges are located by two merge-path searches, staged in LDS, each thread merges 8
// outputs, and the results leave through LDS in coalesced 16-byte stores.
// LDS slot of tile element i: one pad record after every 8, so the 8-record runs that
// consecutive threads merge start in different banks.
__device__ __forceinline__ uint32_t pad8(uint32_t i) { return i + (i >> 3); }

__device__ __forceinline__ void merge_tile(const ASrc& A, uint32_t na, const Event* B, uint32_t nb, uint32_t t0, const OutMap& out) {
  __shared__ Event s[kMergeTile + kMergeTile / 8];
  __shared__ uint32_t s_split[2];
  const uint32_t total = na + nb;
  const uint32_t t1 = t0 + kMergeTile < total ? t0 + kMergeTile : total;
  block_merge_splits(A, na, B, nb, t0, t1, s_split);
  const uint32_t i0 = s_split[0], i1 = s_split[1];
  const uint32_t j0 = t0 - i0;
  const uint32_t ca = i1 - i0, cb = (t1 - t0) - ca;
  for (uint32_t x = threadIdx.x; x < ca; x += kThreads) s[pad8(x)] = A.get(i0 + x);
  for (uint32_t x = threadIdx.x; x < cb; x += kThreads) s[pad8(ca + x)] = load_event(B + j0 + x);
  __syncthreads();
  const uint32_t len = t1 - t0;
  const uint32_t d = threadIdx.x * kPerThread;
  uint4 res[kPerThread];
  uint32_t got = 0;
  if (d < len) {
    uint32_t lo = d > cb ? d - cb : 0u, hi = d < ca ? d : ca;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (key_less(s[pad8(ca + (d - 1u - mid))], s[pad8(mid)])) hi = mid; else lo = mid + 1u;
    }
    uint32_t ia = lo, ib = d - lo;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      if (d + q < len) {
        const bool take_a = ib >= cb || (ia < ca && !key_less(s[pad8(ca + ib)], s[pad8(ia)]));
        const Event r = take_a ? s[pad8(ia)] : s[pad8(ca + ib)];
        res[q] = make_uint4(r.ts, r.dst, r.src, r.seq);
        ia += take_a ? 1u : 0u;
        ib += take_a ? 0u : 1u;
        ++got;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPerThread; ++q)
    if ((uint32_t)q < got) s[pad8(d + q)] = Event{res[q].x, res[q].y, res[q].z, res[q].w};
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < len; x += kThreads) store_event(out.at(t0 + x), s[pad8(x)]);
}

__global__ __launch_bounds__(kThreads) void tile_sort_kernel(const Event* in, Event* out, const uint32_t* d_n) {
  __shared__ Event s[kSortTile];
  const uint32_t n = *d_n;
  const uint32_t base = blockIdx.x * kSortTile;
  if (base >= n) return;
  for (uint32_t i = threadIdx.x; i < kSortTile; i += kThreads) s[i] = base + i < n ? load_event(in + base + i) : sentinel();
  __syncthreads();
  for (uint32_t k = 2; k <= kSortTile; k <<= 1) {           // bitonic sorting network
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < kSortTile; i += kThreads) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const Event a = s[i], b = s[ixj];
          const bool up = (i & k) == 0;
          if (up ? key_less(b, a) : key_less(a, b)) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < kSortTile; i += kThreads)
    if (base + i < n) store_event(out + base + i, s[i]);
}

// Merges adjacent sorted runs of width w (2w is a multiple of kMergeTile).
__global__ __launch_bounds__(kThreads) void merge_pass_kernel(const Event* in, Event* out, const uint32_t* d_n,
                                                              uint32_t w) {
  const uint32_t n = *d_n;
  const uint32_t t0 = blockIdx.x * kMergeTile;
  if (t0 >= n) return;
  const uint32_t base = t0 / (2u * w) * (2u * w);
  const uint32_t na = n - base < w ? n - base : w;
  const uint32_t rest = n - base - na;
  const uint32_t nb = rest < w ? rest : w;
  const ASrc A{in + base, in + base, na};
  OutMap om{{out + base, out + base, out + base}, {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};
  merge_tile(A, na, in + base + na, nb, t0 - base, om);
}

__global__ __launch_bounds__(kThreads) void merge_tasks_kernel(const MergeTask* tasks) {
  const MergeTask t = tasks[blockIdx.y];
  const uint32_t na = t.n_a1 + t.n_a2;
  const uint32_t total = na + t.n_b;
  const uint32_t t0 = blockIdx.x * kMergeTile;
  if (t0 >= total) return;
  const ASrc A{t.a1, t.a2, t.n_a1};
  const OutMap om{{t.out[0], t.out[1], t.out[2]}, {t.end[0], t.end[1], t.end[2]}};
  merge_tile(A, na, t.b, t.n_b, t0, om);
}

__global__ void copy_records_kernel(const Event* src, Event* dst, const uint32_t* d_n, uint32_t cap) {
  const uint32_t n = *d_n < cap ? *d_n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    store_event(dst + i, load_event(src + i));
}

// --- selection --------------------------------------------------------------------

__device__ __forceinline__ bool in_prefix(const Event& e, uint32_t wend, bool bounded, const Event& bound) {
  return e.ts < wend && (!bounded || !key_less(bound, e));
}

__global__ void select_prefix_kernel(const Event* const* d_s, const uint32_t* d_n, const uint32_t* d_wend, const Event* d_bound,
                                     const uint32_t* d_bound_valid, uint32_t* d_prefix) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Event* s = *d_s;
  const uint32_t n = *d_n, wend = *d_wend;
  const bool bounded = d_bound_valid && *d_bound_valid;
  const Event bound = bounded ? *d_bound : sentinel();
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (in_prefix(load_event(s + mid), wend, bounded, bound)) lo = mid + 1u; else hi = mid;
  }
  *d_prefix = lo;
}

__global__ void select_mark_kernel(const Event* const* d_s, const uint32_t* d_prefix, uint32_t* first) {
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p; i += gridDim.x * blockDim.x)
    atomicMin(&first[s[i].dst], i);
}

__global__ __launch_bounds__(kThreads) void select_count_kernel(const Event* const* d_s, const uint32_t* d_prefix,
                                                                const uint32_t* first, uint32_t* counts) {
  __shared__ uint32_t s_w[kThreads / kWave];
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  const uint32_t base = blockIdx.x * kSelectTile;
  uint32_t c = 0;
  if (base < p) {
    for (uint32_t i = base + threadIdx.x; i < base + kSelectTile && i < p; i += kThreads)
      c += first[s[i].dst] == i ? 1u : 0u;
  }
  c = wave_sum_u32(c);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Exclusive scan of counts[0:nblk) into offsets[0:nblk); totals of marked / rest.  Thread t
// owns a contiguous chunk: it sums the chunk (independent loads, all in flight at once), the
// 1024 sums are scanned across the block, and the chunk's offsets follow from its share.
__global__ __launch_bounds__(1024) void select_scan_kernel(const uint32_t* counts, uint32_t* offsets, uint32_t nblk,
                                                           const uint32_t* d_prefix, uint32_t* d_deq_n,
                                                           uint32_t* d_rest_n) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const uint32_t per = (nblk + blockDim.x - 1) / blockDim.x;
  const uint32_t b0 = min(threadIdx.x * per, nblk), b1 = min(b0 + per, nblk);
  uint32_t v = 0;
#pragma unroll 8
  for (uint32_t i = b0; i < b1; ++i) v += counts[i];
  uint32_t x = v;   // inclusive wave scan of the chunk sums
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
    if (lane_id() >= off) x += y;
  }
  if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = x;
  __syncthreads();
  uint32_t run = x - v;
  for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) run += s_w[w];
  for (uint32_t i = b0; i < b1; ++i) {
    const uint32_t c = counts[i];
    offsets[i] = run;
    run += c;
  }
  if (threadIdx.x == blockDim.x - 1) s_carry = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    *d_deq_n = s_carry;
    *d_rest_n = *d_prefix - s_carry;
  }
}

__global__ __launch_bounds__(kThreads) void select_scatter_kernel(const Event* const* d_s, const uint32_t* d_prefix,
                                                                  uint32_t* first, const uint32_t* offsets,
                                                                  Event* deq, Event* rest) {
  __shared__ uint32_t s_w[kThreads / kWave];
  const Event* s = *d_s;
  const uint32_t p = *d_prefix;
  const uint32_t base = blockIdx.x * kSelectTile;
  if (base >= p) return;
  constexpr uint32_t kItems = kSelectTile / kThreads;   // 4 consecutive records per thread
  const uint32_t i0 = base + threadIdx.x * kItems;
  bool mark[kItems];
  Event e[kItems];
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t q = 0; q < kItems; ++q) {
    const uint32_t i = i0 + q;
    mark[q] = false;
    if (i < p) {
      e[q] = load_event(s + i);
      mark[q] = first[e[q].dst] == i;
      mine += mark[q] ? 1u : 0u;
    }
  }
  uint32_t x = mine;   // inclusive scan over the block, in thread order
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
    if (lane_id() >= off) x += y;
  }
  if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = x;
  __syncthreads();
  uint32_t wave_off = 0;
  for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) wave_off += s_w[w];
  uint32_t before = offsets[blockIdx.x] + wave_off + x - mine;   // marked records before i0
#pragma unroll
  for (uint32_t q = 0; q < kItems; ++q) {
    const uint32_t i = i0 + q;
    if (i < p) {
      if (mark[q]) {
        store_event(deq + before, e[q]);
        first[e[q].dst] = 0xFFFFFFFFu;
        ++before;
      } else {
        store_event(rest + (i - before), e[q]);
      }
    }
  }
}

uint32_t grid_of(uint64_t items, uint32_t per_block) {
  const uint64_t g = (items + per_block - 1) / per_block;
  return g ? (uint32_t)g : 1u;
}

}  // namespace

gpq_status sort_records(const Event* in, Event* buf0, Event* buf1, const uint32_t* d_n, uint32_t cap,
                        hipStream_t stream, Event** sorted) {
  hipLaunchKernelGGL(tile_sort_kernel, dim3(grid_of(cap, kSortTile)), dim3(kThreads), 0, stream, in, buf0, d_n);
  GPQ_HIP(hipGetLastError());
  Event* cur = buf0;
  Event* other = buf1;
  for (uint64_t w = kSortTile; w < cap; w *= 2) {
    hipLaunchKernelGGL(merge_pass_kernel, dim3(grid_of(cap, kMergeTile)), dim3(kThreads), 0, stream, cur, other, d_n,
                       (uint32_t)w);
    GPQ_HIP(hipGetLastError());
    Event* t = cur;
    cur = other;
    other = t;
  }
  *sorted = cur;
  return GPQ_OK;
}

gpq_status merge_tasks(const MergeTask* d_tasks, uint32_t ntasks, uint64_t max_total, hipStream_t stream) {
  if (ntasks == 0) return GPQ_OK;
  hipLaunchKernelGGL(merge_tasks_kernel, dim3(grid_of(max_total, kMergeTile), ntasks), dim3(kThreads), 0, stream,
                     d_tasks);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status select_safe(const Event* const* s, const uint32_t* d_n, uint32_t cap, const uint32_t* d_wend,
                       const Event* d_bound, const uint32_t* d_bound_valid, const SelectBufs& b, Event* deq,
                       Event* rest, hipStream_t stream) {
  const uint32_t nblk = grid_of(cap, kSelectTile);
  uint32_t* counts = b.scratch;
  uint32_t* offsets = b.scratch + nblk;
  hipLaunchKernelGGL(select_prefix_kernel, dim3(1), dim3(64), 0, stream, s, d_n, d_wend, d_bound, d_bound_valid,
                     b.d_prefix);
  hipLaunchKernelGGL(select_mark_kernel, dim3(grid_of(cap, kThreads) < 4096 ? grid_of(cap, kThreads) : 4096),
                     dim3(kThreads), 0, stream, s, b.d_prefix, b.first);
  hipLaunchKernelGGL(select_count_kernel, dim3(nblk), dim3(kThreads), 0, stream, s, b.d_prefix, b.first, counts);
  hipLaunchKernelGGL(select_scan_kernel, dim3(1), dim3(1024), 0, stream, counts, offsets, nblk, b.d_prefix, b.d_deq_n,
                     b.d_rest_n);
  hipLaunchKernelGGL(select_scatter_kernel, dim3(nblk), dim3(kThreads), 0, stream, s, b.d_prefix, b.first, offsets, deq,
                     rest);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status copy_records(const Event* src, Event* dst, const uint32_t* d_n, uint32_t cap, hipStream_t stream) {
  const uint32_t g = grid_of(cap, kThreads);
  hipLaunchKernelGGL(copy_records_kernel, dim3(g < 8192 ? g : 8192), dim3(kThreads), 0, stream, src, dst, d_n, cap);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

}  // namespace gpq

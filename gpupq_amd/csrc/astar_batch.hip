// astar_batch.hip -- batched A*: many independent grid queries, one workgroup per query.
//
// The reference runs one block per agent and pulls agents from an atomic counter
// (get_new_agent, src/a_star.cu:149-164); thread 0 pops up to 8 vertices, the block expands
// them, deduplicates, and thread 0 inserts or updates (src/a_star.cu:325-374,441-569).  Here
// every thread of the workgroup owns one open list (a binary heap of 64-bit keys in HBM,
// stored slot-major: entry i of list t at heap[i * K + t]) and pops its own best entry each
// iteration, so K expansions run in parallel.  Improved neighbours are routed to the list
// hash(cell) mod K through an LDS staging area (own list when the stage is full); g is
// relaxed with atomicMin, so entries made stale by a cheaper path are recognised on pop.
// The search of a query ends when the best goal cost found is <= the smallest f of all
// lists, which makes the result the optimal cost (admissible, consistent heuristic).
//
// Key = f << 38 | h << 22 | cell: order by f, ties to the smaller h (closer to the goal).
// Limits that follow: side <= 2048 (cell < 2^22), f < 2^26.
//
// g entries carry a generation tag so that a slot is not cleared per query: entry =
// tag << 26 | g, where the tag of the slot's current query is smaller than every tag written
// before (tags count down from 62; 63 = never written), so atomicMin still relaxes and an
// entry of an older query compares as infinite.  After tag 0 the slot is cleared once.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "astar.h"

namespace gpq {

namespace {

constexpr uint32_t kMaxSide = 2048;
constexpr uint64_t kNoKey = ~0ull;
constexpr uint32_t kNoCost = 0xFFFFFFFFu;
constexpr int kMaxWaves = 16;

struct BatchQuery {
  uint32_t side, conn, sx, sy, tx, ty;
  uint64_t bit_off;                     // first bitmap word of the query in the pool
};

constexpr uint32_t kTagShift = 26;
constexpr uint32_t kTagNever = 63;

// Per-slot state: the generation tag of the slot's query, and (plain drive) the search
// state between launches.
struct SlotState {
  uint32_t best;
  uint32_t overflow;
  uint32_t tag;
  uint32_t pad;
  unsigned long long expanded;
};

struct BatchDev {
  const BatchQuery* q;
  const uint32_t* bits;                 // pooled obstacle bitmaps
  uint32_t* g;                          // [slot][g_stride]
  uint64_t* heap;                       // [slot][H][K]
  uint32_t* hn;                         // [slot][K] list lengths (plain drive)
  SlotState* slot;
  long long* cost;                      // [query]
  unsigned long long* expanded;         // [query]
  uint32_t* flags;                      // [query] 1 = open-list overflow
  unsigned long long* clk;              // [query][2] device wall clock at pickup and at completion
  const uint32_t* ids;                  // queries of this launch
  uint32_t n;
  uint32_t* counter;                    // work counter of the persistent launch
  uint32_t slot_base;                   // workgroup b of the launch uses slot slot_base + b
  uint64_t g_stride;
  uint32_t heap_cap, stage_cap, log2k;
  uint32_t lds_rows;                    // heap levels held in LDS: entries [0, lds_rows)
};

__device__ __forceinline__ uint64_t make_key(uint32_t f, uint32_t h, uint32_t cell) {
  return ((uint64_t)f << 38) | ((uint64_t)h << 22) | cell;
}
__device__ __forceinline__ uint32_t key_f(uint64_t k) { return (uint32_t)(k >> 38); }
__device__ __forceinline__ uint32_t key_cell(uint64_t k) { return (uint32_t)(k & 0x3FFFFFu); }

__device__ __forceinline__ uint32_t owner_of(uint32_t cell, uint32_t log2k) {
  return log2k ? (cell * 0x9E3779B1u) >> (32 - log2k) : 0u;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// Binary heap of one thread.  Entry i lives in LDS (row i of a slot-major [T][K] array)
// for i < T -- the levels every push and pop touches -- and in HBM (row i of the slot's
// [H][K] array) below; HBM rows < T only hold the LDS rows between plain-drive launches.
struct ThreadHeap {
  uint64_t* lds;      // lds_rows + tid
  uint64_t* hbm;      // slot heap + tid
  uint32_t K, T;
  __device__ __forceinline__ uint64_t get(uint32_t i) const { return i < T ? lds[i * K] : hbm[(uint64_t)i * K]; }
  __device__ __forceinline__ void set(uint32_t i, uint64_t v) const {
    if (i < T) lds[i * K] = v;
    else hbm[(uint64_t)i * K] = v;
  }
};

// Returns the new top.
__device__ __forceinline__ uint64_t heap_push(const ThreadHeap& hp, uint32_t& hn, uint64_t key, uint64_t top) {
  uint32_t i = hn++;
  while (i > 0) {
    const uint32_t p = (i - 1) >> 1;
    const uint64_t pv = hp.get(p);
    if (pv <= key) break;
    hp.set(i, pv);
    i = p;
  }
  hp.set(i, key);
  return key < top ? key : top;
}

__device__ __forceinline__ uint64_t heap_pop(const ThreadHeap& hp, uint32_t& hn) {
  --hn;
  if (hn == 0) return kNoKey;
  const uint64_t last = hp.get(hn);
  uint32_t i = 0;
  uint64_t top = last;
  for (;;) {
    uint32_t c = 2 * i + 1;
    if (c >= hn) break;
    uint64_t cv = hp.get(c);
    if (c + 1 < hn) {
      const uint64_t c2 = hp.get(c + 1);
      if (c2 < cv) { cv = c2; ++c; }
    }
    if (last <= cv) break;
    hp.set(i, cv);
    if (i == 0) top = cv;
    i = c;
  }
  hp.set(i, last);
  return top;
}

struct Shared {
  uint32_t best;
  uint32_t overflow;
  uint32_t qi;
  uint32_t pad;
  unsigned long long red[kMaxWaves];
  unsigned long long expanded[kMaxWaves];
};

// Searches query `qidx` in `slot`.  max_iters = 0: to the end; otherwise at most that many
// iterations, keeping the state in the slot (resume continues it).  Returns true when done.
__device__ bool search_query(const BatchDev& d, Shared& sh, uint32_t* cnt, uint64_t* stage, uint64_t* lds_rows,
                             uint32_t qidx, uint32_t slot, uint32_t max_iters, bool resume) {
  const uint32_t tid = threadIdx.x, K = blockDim.x, S = d.stage_cap;
  const BatchQuery Q = d.q[qidx];
  const uint32_t W = Q.side;
  const uint32_t* bits = d.bits + Q.bit_off;
  uint32_t* g = d.g + (uint64_t)slot * d.g_stride;
  const ThreadHeap hp{lds_rows + tid, d.heap + (uint64_t)slot * d.heap_cap * K + tid, K, d.lds_rows};
  const uint32_t source = Q.sy * W + Q.sx, target = Q.ty * W + Q.tx;

  uint32_t hn;
  uint64_t top;
  unsigned long long expanded = 0;
  uint32_t tag;
  if (!resume) {
    if (tid == 0) d.clk[2 * qidx] = wall_clock64();
    const uint32_t prev = d.slot[slot].tag;
    if (prev == 0) {                                   // tags exhausted: clear the slot once
      // the whole slot, not only this query's cells: a later, larger query in the slot would
      // otherwise meet entries of older queries whose tags are below its own
      for (uint64_t c = tid; c < d.g_stride; c += K) g[c] = 0xFFFFFFFFu;
    }
    tag = prev == 0 ? kTagNever - 1 : prev - 1;
    hn = 0;
    top = kNoKey;
    if (tid == 0) {
      sh.best = source == target ? 0u : kNoCost;
      sh.overflow = 0;
    }
    __syncthreads();
    if (tid == 0) d.slot[slot].tag = tag;
    if (source != target && tid == owner_of(source, d.log2k)) {
      __hip_atomic_store(&g[source], tag << kTagShift, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t h0 = grid_heur(Q.conn, (int)Q.sx, (int)Q.sy, (int)Q.tx, (int)Q.ty);
      top = heap_push(hp, hn, make_key(h0, h0, source), top);
    }
  } else {
    hn = d.hn[(uint64_t)slot * K + tid];
    for (uint32_t i = 0; i < hn && i < hp.T; ++i) hp.lds[i * K] = hp.hbm[(uint64_t)i * K];   // saved LDS rows
    top = hn ? hp.get(0) : kNoKey;
    tag = d.slot[slot].tag;
    if (tid == 0) {
      sh.best = d.slot[slot].best;
      sh.overflow = d.slot[slot].overflow;
    }
  }
  const uint32_t tag_bits = tag << kTagShift;
  cnt[tid] = 0;
  __syncthreads();

  bool finished = false;
  uint32_t it = 0;
  for (;;) {
    // --- pop and expand -------------------------------------------------------------------
    const uint32_t best_now = *(volatile uint32_t*)&sh.best;
    if (hn && key_f(top) < best_now) {
      const uint64_t key = top;
      top = heap_pop(hp, hn);
      const uint32_t cell = key_cell(key), f = key_f(key);
      const int x = (int)(cell % W), y = (int)(cell / W);
      const uint32_t g0 = f - grid_heur(Q.conn, x, y, (int)Q.tx, (int)Q.ty);
      if ((tag_bits | g0) == __hip_atomic_load(&g[cell], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ++expanded;
        for (uint32_t k = 0; k < Q.conn; ++k) {
          const int nx = x + nbr_dx(k), ny = y + nbr_dy(k);
          if (nx < 0 || ny < 0 || nx >= (int)W || ny >= (int)W) continue;
          const uint32_t v = (uint32_t)ny * W + (uint32_t)nx;
          if (bit_get(bits, v)) continue;
          const uint32_t g2 = g0 + nbr_cost(Q.conn, k);
          if (v == target) {
            atomicMin(&sh.best, g2);
            continue;
          }
          if (atomicMin(&g[v], tag_bits | g2) <= (tag_bits | g2)) continue;
          const uint32_t h2 = grid_heur(Q.conn, nx, ny, (int)Q.tx, (int)Q.ty);
          if (g2 + h2 >= *(volatile uint32_t*)&sh.best) continue;   // cannot beat the goal found
          const uint64_t k2 = make_key(g2 + h2, h2, v);
          const uint32_t o = owner_of(v, d.log2k);
          const uint32_t pos = atomicAdd(&cnt[o], 1u);
          if (pos < S) {
            stage[o * S + pos] = k2;
          } else if (hn < d.heap_cap) {
            top = heap_push(hp, hn, k2, top);
          } else {
            sh.overflow = 1;
          }
        }
      }
    }
    __syncthreads();
    // --- fold the staged inserts into the own list, reduce the smallest key --------------
    const uint32_t mine = cnt[tid] < S ? cnt[tid] : S;
    for (uint32_t i = 0; i < mine; ++i) {
      if (hn < d.heap_cap) top = heap_push(hp, hn, stage[tid * S + i], top);
      else sh.overflow = 1;
    }
    cnt[tid] = 0;
    const uint32_t best_snap = sh.best;        // only changes while expanding
    const uint64_t wmin = wave_min_u64(top);
    if (lane_id() == 0) sh.red[tid / kWave] = wmin;
    __syncthreads();
    uint64_t m = sh.red[0];
    for (uint32_t w = 1; w < K / kWave; ++w) m = sh.red[w] < m ? sh.red[w] : m;
    ++it;
    if (m == kNoKey || key_f(m) >= best_snap) { finished = true; break; }
    if (max_iters && it >= max_iters) break;
  }

  // --- results or saved state ---------------------------------------------------------------
  unsigned long long e = expanded;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, kWave);
  if (lane_id() == 0) sh.expanded[tid / kWave] = e;
  __syncthreads();
  if (tid == 0) {
    unsigned long long tot = 0;
    for (uint32_t w = 0; w < K / kWave; ++w) tot += sh.expanded[w];
    if (resume) tot += d.slot[slot].expanded;
    if (finished) {
      d.clk[2 * qidx + 1] = wall_clock64();
      d.cost[qidx] = sh.best == kNoCost ? -1ll : (long long)sh.best;
      d.expanded[qidx] = tot;
      d.flags[qidx] = sh.overflow;
    } else {
      d.slot[slot] = SlotState{sh.best, sh.overflow, tag, 0u, tot};
    }
  }
  if (!finished) {
    d.hn[(uint64_t)slot * K + tid] = hn;
    for (uint32_t i = 0; i < hn && i < hp.T; ++i) hp.hbm[(uint64_t)i * K] = hp.lds[i * K];   // save LDS rows
  }
  __syncthreads();
  return finished;
}

// Persistent drive: each workgroup owns slot blockIdx.x and pulls queries until none is left.
// Dynamic LDS: stage [K][S] u64, heap rows [T][K] u64, counts [K] u32.
__global__ __launch_bounds__(1024) void astar_batch_kernel(BatchDev d) {
  extern __shared__ uint64_t lds[];
  __shared__ Shared sh;
  uint64_t* stage = lds;
  uint64_t* rows = lds + (size_t)blockDim.x * d.stage_cap;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(rows + (size_t)blockDim.x * d.lds_rows);
  for (;;) {
    if (threadIdx.x == 0) sh.qi = atomicAdd(d.counter, 1u);
    __syncthreads();
    const uint32_t qi = sh.qi;
    __syncthreads();
    if (qi >= d.n) break;
    search_query(d, sh, cnt, stage, rows, d.ids[qi], d.slot_base + blockIdx.x, 0, false);
  }
}

// Plain drive: one iteration of query ids[0] in slot slot_base per launch.
__global__ __launch_bounds__(1024) void astar_batch_step_kernel(BatchDev d, uint32_t resume, uint32_t* done) {
  extern __shared__ uint64_t lds[];
  __shared__ Shared sh;
  uint64_t* stage = lds;
  uint64_t* rows = lds + (size_t)blockDim.x * d.stage_cap;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(rows + (size_t)blockDim.x * d.lds_rows);
  const bool fin = search_query(d, sh, cnt, stage, rows, d.ids[0], d.slot_base, 1, resume != 0);
  if (threadIdx.x == 0) *done = fin ? 1u : 0u;
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0; }

constexpr size_t kLdsBudget = 64 * 1024;

size_t lds_bytes(uint32_t K, uint32_t S, uint32_t rows) { return (size_t)K * (S * 8 + rows * 8 + 4); }

// Heap levels kept in LDS: the largest 2^d - 1 (d <= 5) that fits beside the staging area.
uint32_t heap_lds_rows(uint32_t K, uint32_t S) {
  for (uint32_t rows : {31u, 15u, 7u, 3u, 1u})
    if (lds_bytes(K, S, rows) <= kLdsBudget) return rows;
  return 0;
}

__global__ void init_slots_kernel(SlotState* s, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = SlotState{kNoCost, 0u, kTagNever, 0u, 0ull};
}

uint32_t gen_blocks(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(g, 1), 65536);
}

bool host_blocked(uint32_t seed, uint32_t tag, uint32_t query, uint32_t threshold, uint64_t c) {
  const uint4 r = philox4x32_10(make_uint4((uint32_t)(c >> 2), query, tag, 2u), make_uint2(seed, 0u));
  const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
  return rr[c & 3] < threshold;
}

}  // namespace

struct AstarBatch::Dev : BatchDev {};

// Query q of the synthetic batch: shape from Philox({q, 0, 3, 0}), source and target from
// the first free candidates of Philox({a, q, 4 | 5, 0}), obstacles from stream tag 1.
gpq_status query_synthetic(uint32_t seed, uint32_t q, gpq_query_spec* out) {
  const uint4 r = philox4x32_10(make_uint4(q, 0u, 3u, 0u), make_uint2(seed, 0u));
  const double u0 = (double)r.x * 0x1p-32, u1 = (double)r.y * 0x1p-32;
  gpq_query_spec s{};
  s.side = (uint32_t)std::floor(256.0 * std::exp2(3.0 * u0) + 0.5);
  s.threshold = (uint32_t)((0.1 + 0.2 * u1) * 4294967296.0);
  s.connectivity = 8;
  s.seed = seed;
  s.stream_tag = 1;
  s.query = q;
  for (int which = 0; which < 2; ++which) {
    for (uint32_t a = 0;; ++a) {
      const uint4 p = philox4x32_10(make_uint4(a, q, 4u + (uint32_t)which, 0u), make_uint2(seed, 0u));
      const uint32_t x = (uint32_t)(((uint64_t)p.x * s.side) >> 32), y = (uint32_t)(((uint64_t)p.y * s.side) >> 32);
      if (!host_blocked(seed, s.stream_tag, q, s.threshold, (uint64_t)y * s.side + x)) {
        if (which == 0) { s.sx = x; s.sy = y; } else { s.tx = x; s.ty = y; }
        break;
      }
      if (a > (1u << 20)) {
        set_error("query_synthetic: no free cell found");
        return GPQ_ERR_INVALID;
      }
    }
  }
  *out = s;
  return GPQ_OK;
}

gpq_status AstarBatch::create(const gpq_query_spec* sp, uint32_t count, const gpq_batch_tuning& t, AstarBatch** out) {
  GPQ_CHECK_ARG(sp && count >= 1, "batch: no queries");
  gpq_batch_tuning tu = t;
  if (!tu.threads_per_query) tu.threads_per_query = 256;
  if (!tu.heap_capacity) tu.heap_capacity = 2048;
  if (!tu.stage_capacity) tu.stage_capacity = 8;
  GPQ_CHECK_ARG(tu.threads_per_query >= 64 && tu.threads_per_query <= 1024 && is_pow2(tu.threads_per_query),
                "batch: threads_per_query must be a power of two in [64, 1024]");
  GPQ_CHECK_ARG(tu.heap_capacity >= 16 && tu.heap_capacity <= (1u << 20), "batch: heap_capacity out of range");
  GPQ_CHECK_ARG(tu.stage_capacity >= 1 && tu.stage_capacity <= 64, "batch: stage_capacity must be in [1, 64]");
  GPQ_CHECK_ARG((size_t)tu.threads_per_query * (tu.stage_capacity * 8 + 4) <= kLdsBudget,
                "batch: LDS staging (threads * (8 * stage + 4) bytes) must fit in 64 KiB");
  uint64_t max_cells = 0, words = 0;
  std::vector<BatchQuery> hq(count);
  std::vector<uint64_t> off(count);
  for (uint32_t i = 0; i < count; ++i) {
    const gpq_query_spec& s = sp[i];
    GPQ_CHECK_ARG(s.side >= 1 && s.side <= kMaxSide, "batch: side must be in [1, 2048]");
    GPQ_CHECK_ARG(s.connectivity == 4 || s.connectivity == 8, "batch: connectivity must be 4 or 8");
    GPQ_CHECK_ARG(s.sx < s.side && s.sy < s.side && s.tx < s.side && s.ty < s.side, "batch: endpoint outside grid");
    const uint64_t cells = (uint64_t)s.side * s.side;
    max_cells = std::max(max_cells, cells);
    off[i] = words;
    hq[i] = BatchQuery{s.side, s.connectivity, s.sx, s.sy, s.tx, s.ty, words};
    words += ((cells + 31) / 32 + 3) & ~3ull;
  }
  int dev = 0, cus = 0;
  GPQ_HIP(hipGetDevice(&dev));
  GPQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t rows = heap_lds_rows(tu.threads_per_query, tu.stage_capacity);
  const size_t lds = lds_bytes(tu.threads_per_query, tu.stage_capacity, rows);
  int per_cu = 0;
  GPQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, astar_batch_kernel, (int)tu.threads_per_query, lds));
  uint32_t resident = tu.max_resident ? tu.max_resident : (uint32_t)std::max(1, cus * std::max(per_cu, 1));
  resident = std::min(resident, count);
  tu.max_resident = resident;

  AstarBatch* b = new AstarBatch();
  b->specs.assign(sp, sp + count);
  b->tuning = tu;
  b->resident = resident;
  b->cell_off = off;
  auto fail = [&](gpq_status s) { delete b; return s; };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(GPQ_ERR_HIP);
  }
  Dev* d = new Dev();
  b->dev = d;
  BatchQuery* dq = nullptr;
  uint32_t* bits = nullptr;
  gpq_status s;
  const uint64_t slots = std::max<uint32_t>(resident, 1);
  if ((s = b->arena.alloc(&dq, count)) != GPQ_OK || (s = b->arena.alloc(&bits, words ? words : 1)) != GPQ_OK ||
      (s = b->arena.alloc(&d->g, slots * max_cells)) != GPQ_OK ||
      (s = b->arena.alloc(&d->heap, slots * tu.heap_capacity * tu.threads_per_query)) != GPQ_OK ||
      (s = b->arena.alloc(&d->hn, slots * tu.threads_per_query)) != GPQ_OK ||
      (s = b->arena.alloc(&d->slot, slots)) != GPQ_OK || (s = b->arena.alloc(&d->cost, count)) != GPQ_OK ||
      (s = b->arena.alloc(&d->expanded, count)) != GPQ_OK || (s = b->arena.alloc(&d->flags, count)) != GPQ_OK ||
      (s = b->arena.alloc(&d->clk, 2 * (size_t)count)) != GPQ_OK ||
      (s = b->arena.alloc(&b->d_ids_all, count)) != GPQ_OK || (s = b->arena.alloc(&b->counter, 2)) != GPQ_OK)
    return fail(s);
  d->q = dq;
  d->bits = bits;
  d->g_stride = max_cells;
  d->heap_cap = tu.heap_capacity;
  d->stage_cap = tu.stage_capacity;
  d->log2k = log2u(tu.threads_per_query);
  d->lds_rows = rows;
  d->counter = b->counter;
  std::vector<uint32_t> ids(count);
  for (uint32_t i = 0; i < count; ++i) ids[i] = i;
  if (hipMemcpyAsync(dq, hq.data(), sizeof(BatchQuery) * count, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_ids_all, ids.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice, b->stream) !=
          hipSuccess) {
    set_error("batch: upload failed");
    return fail(GPQ_ERR_HIP);
  }
  hipLaunchKernelGGL(fill_u32_kernel, dim3(gen_blocks(slots * max_cells)), dim3(256), 0, b->stream, d->g,
                     slots * max_cells, 0xFFFFFFFFu);
  hipLaunchKernelGGL(init_slots_kernel, dim3(gen_blocks(slots)), dim3(256), 0, b->stream, d->slot, (uint32_t)slots);
  for (uint32_t i = 0; i < count; ++i) {
    const gpq_query_spec& q = sp[i];
    const uint64_t cells = (uint64_t)q.side * q.side;
    const uint64_t src = (uint64_t)q.sy * q.side + q.sx, tgt = (uint64_t)q.ty * q.side + q.tx;
    hipLaunchKernelGGL(grid_generate_kernel, dim3(gen_blocks((cells + 31) / 32)), dim3(256), 0, b->stream,
                       bits + off[i], cells, q.seed, q.stream_tag, q.query, q.threshold, src, tgt);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
    set_error("batch: grid generation failed");
    return fail(GPQ_ERR_HIP);
  }
  *out = b;
  return GPQ_OK;
}

AstarBatch::~AstarBatch() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  delete static_cast<Dev*>(dev);
  if (stream) (void)hipStreamDestroy(stream);
}

gpq_status AstarBatch::launch(const uint32_t* d_ids, uint32_t n, uint32_t slot_base, uint32_t slots,
                              uint32_t* d_counter, hipStream_t s) {
  GPQ_CHECK_ARG(slots >= 1 && slot_base + slots <= resident, "batch: slot range outside the resident slots");
  Dev d = *static_cast<Dev*>(dev);
  d.ids = d_ids;
  d.n = n;
  d.counter = d_counter;
  d.slot_base = slot_base;
  const uint32_t K = tuning.threads_per_query;
  const size_t lds = lds_bytes(K, tuning.stage_capacity, d.lds_rows);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, d_counter);
  hipLaunchKernelGGL(astar_batch_kernel, dim3(std::max<uint32_t>(1, std::min(slots, n))), dim3(K), lds, s,
                     static_cast<BatchDev>(d));
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

// The one-query-per-launch path: query ids[0] (device) alone in `slot`, one launch per
// iteration, each followed by a host synchronization that reads the done flag.
gpq_status AstarBatch::plain_one(const uint32_t* d_id, uint32_t spec_index, uint32_t slot, uint32_t* d_done,
                                 uint32_t* h_done, hipStream_t s) {
  GPQ_CHECK_ARG(slot < resident, "batch: slot outside the resident slots");
  Dev d = *static_cast<Dev*>(dev);
  d.ids = d_id;
  d.n = 1;
  d.slot_base = slot;
  const uint32_t K = tuning.threads_per_query;
  const size_t lds = lds_bytes(K, tuning.stage_capacity, d.lds_rows);
  const uint64_t step_cap = 64ull * specs[spec_index].side * specs[spec_index].side + 1024;
  for (uint64_t step = 0;; ++step) {
    if (step > step_cap) {
      set_error("batch plain: step bound exceeded");
      return GPQ_ERR_STATE;
    }
    hipLaunchKernelGGL(astar_batch_step_kernel, dim3(1), dim3(K), lds, s, static_cast<BatchDev>(d), step ? 1u : 0u,
                       d_done);
    GPQ_HIP(hipGetLastError());
    GPQ_HIP(hipMemcpyAsync(h_done, d_done, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GPQ_HIP(hipStreamSynchronize(s));
    if (*h_done) return GPQ_OK;
  }
}

gpq_status AstarBatch::run(gpq_drive_mode mode, double* seconds) {
  const uint32_t n = (uint32_t)specs.size();
  GPQ_HIP(hipStreamSynchronize(stream));
  const auto t0 = std::chrono::steady_clock::now();
  if (mode == GPQ_DRIVE_GRAPH) {
    if (!graph_exec) {
      GPQ_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      const gpq_status st = launch(d_ids_all, n, 0, resident, counter, stream);
      hipGraph_t gph = nullptr;
      const hipError_t e = hipStreamEndCapture(stream, &gph);
      if (st != GPQ_OK) return st;
      GPQ_HIP(e);
      graph = gph;
      GPQ_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    }
    GPQ_HIP(hipGraphLaunch(graph_exec, stream));
    GPQ_HIP(hipStreamSynchronize(stream));
  } else {
    uint32_t* h_done = nullptr;
    GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_done), sizeof(uint32_t), hipHostMallocDefault));
    gpq_status st = GPQ_OK;
    for (uint32_t qi = 0; qi < n && st == GPQ_OK; ++qi) st = plain_one(d_ids_all + qi, qi, 0, counter + 1, h_done, stream);
    (void)hipHostFree(h_done);
    if (st != GPQ_OK) return st;
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return GPQ_OK;
}

// Times every candidate (threads per query x stage capacity) on the first `sample` queries
// (two persistent runs each, the second timed) and returns the fastest; heap_capacity and
// max_resident are kept from `base`.  Candidates whose open lists overflow are skipped.
gpq_status batch_autotune(const gpq_query_spec* sp, uint32_t count, uint32_t sample, const gpq_batch_tuning& base,
                          gpq_batch_tuning* out, double* best_seconds) {
  GPQ_CHECK_ARG(sp && count >= 1, "autotune: no queries");
  const uint32_t n = sample ? std::min(sample, count) : count;
  static const uint32_t kThreads[] = {64, 128, 256, 512, 1024};
  static const uint32_t kStage[] = {8, 16, 32};
  double best = 1e300;
  gpq_batch_tuning pick{};
  for (uint32_t k : kThreads)
    for (uint32_t st : kStage) {
      if ((size_t)k * (st * 8 + 4) > kLdsBudget) continue;
      gpq_batch_tuning t = base;
      t.threads_per_query = k;
      t.stage_capacity = st;
      AstarBatch* b = nullptr;
      if (AstarBatch::create(sp, n, t, &b) != GPQ_OK) continue;
      double sec = 0.0;
      gpq_status s = b->run(GPQ_DRIVE_GRAPH, &sec);
      if (s == GPQ_OK) s = b->run(GPQ_DRIVE_GRAPH, &sec);
      if (s == GPQ_OK) s = b->results(nullptr, nullptr);
      const gpq_batch_tuning used = b->tuning;
      delete b;
      if (s == GPQ_OK && sec < best) {
        best = sec;
        pick = used;
        pick.max_resident = base.max_resident;
      }
    }
  if (best == 1e300) {
    set_error("autotune: no candidate completed");
    return GPQ_ERR_CAPACITY;
  }
  *out = pick;
  if (best_seconds) *best_seconds = best;
  return GPQ_OK;
}

gpq_status AstarBatch::results(int64_t* costs, uint64_t* expanded) {
  const Dev* d = static_cast<const Dev*>(dev);
  const size_t n = specs.size();
  std::vector<uint32_t> flags(n);
  if (costs) GPQ_HIP(hipMemcpyAsync(costs, d->cost, sizeof(int64_t) * n, hipMemcpyDeviceToHost, stream));
  if (expanded) GPQ_HIP(hipMemcpyAsync(expanded, d->expanded, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(flags.data(), d->flags, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  for (size_t i = 0; i < n; ++i)
    if (flags[i]) {
      set_error("batch: query " + std::to_string(i) + " overflowed its open lists (raise heap_capacity)");
      return GPQ_ERR_CAPACITY;
    }
  return GPQ_OK;
}

gpq_status AstarBatch::timing(double* start_us, double* finish_us) {
  const Dev* d = static_cast<const Dev*>(dev);
  const size_t n = specs.size();
  std::vector<unsigned long long> clk(2 * n);
  GPQ_HIP(hipMemcpyAsync(clk.data(), d->clk, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  int dev_id = 0, khz = 0;
  GPQ_HIP(hipGetDevice(&dev_id));
  GPQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev_id));
  GPQ_CHECK_ARG(khz > 0, "batch: device wall clock rate unknown");
  unsigned long long t0 = ~0ull;
  for (size_t i = 0; i < n; ++i) t0 = std::min(t0, clk[2 * i]);
  const double us_per_tick = 1000.0 / khz;
  for (size_t i = 0; i < n; ++i) {
    if (start_us) start_us[i] = (double)(clk[2 * i] - t0) * us_per_tick;
    if (finish_us) finish_us[i] = (double)(clk[2 * i + 1] - t0) * us_per_tick;
  }
  return GPQ_OK;
}

// An optimal path of query q: q searched again alone in slot 0 (one persistent launch of one
// workgroup), then the closed-list backtrack of src/a_star.cu:90-94,417-438 over that slot's
// g values on the host: from the target (whose g is the query's cost), a free neighbour whose g
// (of the slot's current tag) plus the step cost equals the current cell's g.  Along such a
// chain every g is the optimal distance (the target's is; a neighbour meeting the equality
// cannot be above its optimum), so the neighbour that relaxed a cell last always qualifies and
// the chain reaches the source.
gpq_status AstarBatch::path(uint32_t q, uint32_t* cells, uint32_t cap, uint32_t* len) {
  GPQ_CHECK_ARG(q < specs.size(), "batch: query index out of range");
  GPQ_CHECK_ARG(len && (cells || cap == 0), "batch path: null output");
  const Dev* d = static_cast<const Dev*>(dev);
  const gpq_query_spec& sp = specs[q];
  const uint32_t side = sp.side;
  const uint64_t n = (uint64_t)side * side;
  GPQ_HIP(hipStreamSynchronize(stream));
  GPQ_TRY(launch(d_ids_all + q, 1, 0, 1, counter, stream));
  long long cost = -1;
  SlotState st{};
  std::vector<uint32_t> g(n);
  GPQ_HIP(hipMemcpyAsync(&cost, d->cost + q, sizeof(cost), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(&st, d->slot, sizeof(st), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipMemcpyAsync(g.data(), d->g, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *len = 0;
  if (cost < 0) return GPQ_OK;
  std::vector<uint8_t> blocked(n);
  GPQ_TRY(grid_bytes(q, blocked.data()));
  const uint32_t mask = (1u << kTagShift) - 1u;
  auto gval = [&](uint64_t c, uint32_t* out) {
    if ((g[c] >> kTagShift) != st.tag) return false;
    *out = g[c] & mask;
    return true;
  };
  static const int dx[8] = {1, -1, 0, 0, 1, -1, 1, -1}, dy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
  const uint64_t source = (uint64_t)sp.sy * side + sp.sx;
  uint64_t cur = (uint64_t)sp.ty * side + sp.tx;
  uint32_t gc = (uint32_t)cost;          // the search records the target's cost, not its g entry
  std::vector<uint32_t> rev{(uint32_t)cur};
  while (cur != source) {
    const int x = (int)(cur % side), y = (int)(cur / side);
    bool found = false;
    for (uint32_t k = 0; k < sp.connectivity && !found; ++k) {
      const int nx = x + dx[k], ny = y + dy[k];
      if (nx < 0 || ny < 0 || nx >= (int)side || ny >= (int)side) continue;
      const uint64_t nc = (uint64_t)ny * side + (uint64_t)nx;
      const uint32_t step = sp.connectivity == 4 ? 1u : (k < 4 ? 10u : 14u);
      uint32_t gn = 0;
      if (blocked[nc] || !gval(nc, &gn) || gn + step != gc) continue;
      cur = nc;
      gc = gn;
      found = true;
    }
    if (!found || rev.size() > n) {
      set_error("batch path: no predecessor with a matching g");
      return GPQ_ERR_STATE;
    }
    rev.push_back((uint32_t)cur);
  }
  *len = (uint32_t)rev.size();
  for (uint32_t i = 0; i < *len && i < cap; ++i) cells[i] = rev[*len - 1 - i];
  return GPQ_OK;
}

gpq_status AstarBatch::grid_bytes(uint32_t q, uint8_t* out) {
  GPQ_CHECK_ARG(q < specs.size(), "batch: query index out of range");
  const Dev* d = static_cast<const Dev*>(dev);
  const uint64_t cells = (uint64_t)specs[q].side * specs[q].side;
  uint8_t* tmp = nullptr;
  GPQ_HIP(hipMalloc(&tmp, cells));
  hipLaunchKernelGGL(grid_to_bytes_kernel, dim3(gen_blocks(cells)), dim3(256), 0, stream, d->bits + cell_off[q], tmp,
                     cells);
  hipError_t e = hipMemcpyAsync(out, tmp, cells, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(tmp);
  GPQ_HIP(e);
  return GPQ_OK;
}

}  // namespace gpq

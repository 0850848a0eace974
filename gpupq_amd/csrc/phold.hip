// phold.hip -- PHOLD model (src/phold.cu) and the DES step loop (src/des.cu:134-221)
// over any queue kind, plus the host-driven insert/peek kernels of the C ABI.
//
// Model kernels are templates over the queue view, so the queue.cuh verbs
// (insert / peek / set_done) inline into them; the reference links them across
// translation units with nvcc -rdc (src/Makefile:21).
//
// Random numbers: LP lp's n-th committed event uses Philox4x32-10 block
// {n, lp, 0, 0} under key {seed, 0}: word 0 -> Exp(1), word 1 -> self test,
// word 2 -> uniform destination.  Initial event j of LP lp uses block {j, lp, 1, 0}.
// The reference keeps one cuRAND XORWOW state per LP (src/phold.cu:53-61,93-106).
#include <chrono>
#include <cstring>

#include "laq_calendar.cuh"
#include "phold.h"
#include "phold_fused.cuh"

namespace gpq {
namespace {

constexpr int kStepBlock = 256;

struct PholdDev {
  uint32_t* count;
  uint64_t* hash;
  uint32_t* last_ts;     // nullptr: checks off
  WindowState* ws;
  uint32_t num_lps;
  uint32_t init_per_lp;
  uint32_t seed;
  uint32_t p_self;
  uint32_t drain;
  float lookahead;
};

// handle_next_event (src/phold.cu:81-117): every peek slot whose item is inside the
// window commits it, counts it, extends the LP trace hash and schedules one new event.
// In drain mode the slot's owner keeps committing its LP's next events while they are
// inside the window.
template <class Q>
__device__ __forceinline__ bool phold_commit_one(const Q& q, const PholdDev& m, uint32_t slot) {
  Event* ev;
  const int idx = q.peek(&ev, (int)slot);
  if (idx < 0) return false;
  const Event e = load_event(ev);
  const uint32_t lp = e.dst;
  const uint32_t n = m.count[lp];
  m.count[lp] = n + 1u;
  m.hash[lp] = trace_step(m.hash[lp], e.ts, e.src, e.seq);
  if (m.last_ts) {                                   // src/des.cu:52-64
    if (e.ts < m.last_ts[lp]) atomicOr(&m.ws->error, kErrBadState);
    m.last_ts[lp] = e.ts;
  }
  const uint4 r = philox4x32_10(make_uint4(n, lp, 0u, 0u), make_uint2(m.seed, 0u));
  const float ex = exp_from_u32(r.x);
  const uint32_t dst = r.y < m.p_self ? lp : (uint32_t)(((uint64_t)r.z * m.num_lps) >> 32);
  const float t_new = (bits_to_f(e.ts) + m.lookahead) + ex;
  q.set_done(idx);
  q.insert(Event{f_to_bits(t_new), dst, lp, m.init_per_lp + n});
  return true;
}

// Drain mode, any queue: commit the slot's events while its head is inside the window.
template <class Q>
__device__ __forceinline__ void phold_drain(const Q& q, const PholdDev& m, uint32_t slot) {
  while (phold_commit_one(q, m, slot)) {
  }
}

// Drain mode on the local array queue: the same commits in the same order as repeated
// peek / set_done, with the LP's ring cursor, count and hash held in registers and the
// next record loaded while the current one is processed; written back once.
__device__ __forceinline__ void phold_drain(const LaqView& q, const PholdDev& m, uint32_t lp) {
  const uint32_t wend = q.ws->wend;
  if (!(q.head[lp] < wend)) return;
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  const Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  uint32_t rd = q.rd[lp], len = q.len[lp];
  uint32_t n = m.count[lp];
  uint64_t h = m.hash[lp];
  uint32_t last = m.last_ts ? m.last_ts[lp] : 0u;
  Event cur = load_event(&ring[rd]);
  while (len && cur.ts < wend) {
    const Event nxt = len > 1 ? load_event(&ring[(rd + 1) & mask]) : Event{kInfBits, 0u, 0u, 0u};
    h = trace_step(h, cur.ts, cur.src, cur.seq);
    if (m.last_ts) {
      if (cur.ts < last) atomicOr(&m.ws->error, kErrBadState);
      last = cur.ts;
    }
    const uint4 r = philox4x32_10(make_uint4(n, lp, 0u, 0u), make_uint2(m.seed, 0u));
    const float ex = exp_from_u32(r.x);
    const uint32_t dst = r.y < m.p_self ? lp : (uint32_t)(((uint64_t)r.z * m.num_lps) >> 32);
    const float t_new = (bits_to_f(cur.ts) + m.lookahead) + ex;
    q.insert(Event{f_to_bits(t_new), dst, lp, m.init_per_lp + n});
    ++n;
    rd = (rd + 1) & mask;
    --len;
    cur = nxt;
  }
  q.rd[lp] = rd;
  q.len[lp] = len;
  q.head[lp] = len ? cur.ts : kInfBits;
  m.count[lp] = n;
  m.hash[lp] = h;
  if (m.last_ts) m.last_ts[lp] = last;
}

template <class Q>
__global__ __launch_bounds__(kStepBlock) void phold_step_kernel(Q q, PholdDev m) {
  const uint32_t slots = q.slots();
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += gridDim.x * blockDim.x) {
    if (m.drain) {
      phold_drain(q, m, s);
    } else {
      (void)phold_commit_one(q, m, s);
    }
  }
}

// Merge-and-commit step (drain mode 2), phold_fused.cuh: the whole queue is one instance.
template <int G, int SL>
__global__ __launch_bounds__(kFusedBlock, 4) void phold_fused_kernel(LaqView q, PholdDev m, uint32_t settle) {
  const FusedModel fm{q.ws, m.count, m.hash, m.last_ts, 0u, m.num_lps, m.init_per_lp, m.seed, m.p_self, m.lookahead};
  fused_block<G, SL>(q, fm, settle);
}

// The step on a calendar-mode LAQ (laq_calendar.cuh): blocks [0, list_blocks) move the events
// beyond the calendar that it now reaches into their logs, the others commit window cur, one
// block of LPs each; the last block to finish moves the window (queue_post's work).  With `inst`
// null the whole queue is one instance; otherwise commit block b takes the model parameters and
// the end time of instance inst[block_inst[b]] (several instances in one queue, every commit
// block in one instance).  m.lookahead and end_time are the queue's lookahead and horizon.
static_assert(cal::kBlock == (int)LaqQueue::kCalBlockLps, "calendar commit block size");
template <bool kStage>
__global__ __launch_bounds__(cal::kBlock) void phold_calendar_kernel(LaqView q, PholdDev m, float end_time,
                                                                     uint32_t list_blocks, const PoolInst* inst,
                                                                     const uint32_t* block_inst) {
  if (q.ws->done) return;
  const uint32_t cur = q.cal_st->cur;
  if (blockIdx.x < list_blocks) {
    cal::list_block(q, end_time, cur + 1u, blockIdx.x, list_blocks);
  } else {
    const uint32_t b = blockIdx.x - list_blocks;
    FusedModel fm{q.ws, m.count, m.hash, m.last_ts, 0u, m.num_lps, m.init_per_lp, m.seed, m.p_self, m.lookahead};
    float et = end_time;
    if (inst) {
      const PoolInst in = inst[block_inst[b]];
      fm.base = in.base;
      fm.lps = in.lps;
      fm.init_per_lp = in.init_per_lp;
      fm.seed = in.seed;
      fm.p_self = in.p_self;
      fm.lookahead = in.lookahead;
      et = in.end_time;
    }
    cal::commit_block<kStage>(q, fm, et, cur, b);
  }
  __syncthreads();
  if (threadIdx.x == 0) cal::block_done(q, m.lookahead, end_time);
}

// Initial event j of every LP of several instances in one calendar-mode LAQ (the populate of
// one instance alone, with the LP's index inside its instance): staged for queue_post_init,
// or parked when it falls at or past the instance's end time.
__global__ void populate_instances_kernel(LaqView q, const PoolInst* inst, const uint32_t* block_inst, uint32_t j) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= q.num_lps) return;
  const PoolInst in = inst[block_inst[lp / cal::kBlock]];
  const uint32_t local = lp - in.base;
  if (local >= in.lps || j >= in.init_per_lp) return;
  const uint4 r = philox4x32_10(make_uint4(j, local, 1u, 0u), make_uint2(in.seed, 0u));
  const Event e{f_to_bits(exp_from_u32(r.x)), lp, local, j};
  if (e.ts < f_to_bits(in.end_time)) q.insert(e);
  else cal::append_one(q, 0u, 0u, cal::Route::kParked, e);
}

// phold_populate (src/phold.cu:63-79): initial event j of every LP.
template <class Q>
__global__ void phold_populate_kernel(Q q, PholdDev m, uint32_t j) {
  for (uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x; lp < m.num_lps; lp += gridDim.x * blockDim.x) {
    const uint4 r = philox4x32_10(make_uint4(j, lp, 1u, 0u), make_uint2(m.seed, 0u));
    q.insert(Event{f_to_bits(exp_from_u32(r.x)), lp, lp, j});
  }
}

template <class Q>
__global__ void insert_items_kernel(Q q, const Event* items, uint32_t n, uint32_t update) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (update) q.insert_or_update(items[i], (int)items[i].dst);
  else q.insert(items[i]);
}

template <class Q>
__global__ void peek_items_kernel(Q q, int set_done, Event* out, int32_t* index, uint32_t* found, uint32_t max) {
  const uint32_t slots = q.slots();
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += gridDim.x * blockDim.x) {
    Event* ev;
    const int idx = q.peek(&ev, (int)s);
    if (idx < 0) continue;
    const uint32_t pos = atomicAdd(found, 1u);
    if (pos < max) {
      out[pos] = load_event(ev);
      index[pos] = idx;
    }
    if (set_done) q.set_done(idx);
  }
}

__global__ void phold_zero_kernel(PholdDev m) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp < m.num_lps) {
    m.count[lp] = 0;
    m.hash[lp] = 0;
    if (m.last_ts) m.last_ts[lp] = 0;
  }
}

__global__ void phold_sum_kernel(const uint32_t* count, uint32_t n, unsigned long long* sum) {
  __shared__ unsigned long long s[16];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = i < n ? (unsigned long long)count[i] : 0ull;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane_id() == 0) s[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s[w];
    atomicAdd(sum, t);
  }
}

PholdDev dev_of(const PholdModel* m) {
  return PholdDev{m->count, m->hash, m->last_ts, m->q->ws, m->cfg.num_lps, m->cfg.init_per_lp, m->cfg.seed,
                  m->cfg.p_self_u32, m->cfg.drain, m->cfg.lookahead};
}

uint32_t grid_for(uint32_t items, uint32_t block) {
  uint32_t g = ceil_div(items, block);
  return g ? g : 1u;
}

uint64_t host_mix64(uint64_t x) {
  x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27; x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

}  // namespace

// ---------------------------------------------------------------------------------
// generic launches over the queue kind
// ---------------------------------------------------------------------------------
gpq_status launch_insert_items(QueueBase* q, const Event* d_items, uint32_t n, bool update) {
  return with_view(q, [&](auto v) -> gpq_status {
    hipLaunchKernelGGL(insert_items_kernel<decltype(v)>, dim3(grid_for(n, 256)), dim3(256), 0, q->stream, v, d_items, n,
                       update ? 1u : 0u);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  });
}

gpq_status launch_peek_items(QueueBase* q, int set_done, Event* d_out, int32_t* d_index, uint32_t* d_found,
                             uint32_t max) {
  return with_view(q, [&](auto v) -> gpq_status {
    hipLaunchKernelGGL(peek_items_kernel<decltype(v)>, dim3(grid_for(q->peek_slots(), 256)), dim3(256), 0,
                       q->stream, v, set_done, d_out, d_index, d_found, max);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  });
}

gpq_status PholdModel::create(QueueBase* queue, const gpq_phold_config& c, PholdModel** out) {
  GPQ_CHECK_ARG(queue != nullptr, "null queue");
  GPQ_CHECK_ARG(c.num_lps == queue->cfg.num_lps, "phold num_lps must equal the queue's num_lps");
  GPQ_CHECK_ARG(c.num_lps > 0 && c.init_per_lp > 0, "phold: num_lps and init_per_lp must be > 0");
  GPQ_CHECK_ARG(c.lookahead > 0.0f && std::isfinite(c.lookahead), "phold: lookahead must be finite and > 0 "
                "(conservative window)");
  GPQ_CHECK_ARG(c.lookahead == queue->cfg.lookahead, "phold lookahead must equal the queue's lookahead");
  GPQ_CHECK_ARG(c.end_time == queue->cfg.horizon, "phold end_time must equal the queue's horizon");
  GPQ_CHECK_ARG(!c.drain || queue->cfg.kind == GPQ_LAQ || queue->cfg.kind == GPQ_LHQ || queue->cfg.kind == GPQ_LSQ,
                "phold drain mode needs a per-LP queue (peek by LP)");
  GPQ_CHECK_ARG(c.drain <= 2, "phold drain must be 0, 1 or 2");
  const bool calendar = queue->cfg.kind == GPQ_LAQ && static_cast<LaqQueue*>(queue)->calendar();
  GPQ_CHECK_ARG(!calendar || c.drain == 2, "a calendar-mode LAQ serves the merge-and-commit step (drain 2) only");
  if (c.drain == 2 && !calendar) {
    GPQ_CHECK_ARG(queue->cfg.kind == GPQ_LAQ && queue->cfg.far_per_lp > 0 && queue->cfg.capacity_per_lp <= 64,
                  "phold drain mode 2 (merge and commit in one step) needs the two-tier LAQ with capacity_per_lp <= 64");
    GPQ_TRY(static_cast<LaqQueue*>(queue)->enable_alt_stage());
  }
  PholdModel* m = new PholdModel();
  m->cfg = c;
  m->q = queue;
  gpq_status s = GPQ_OK;
  if ((s = m->arena.alloc(&m->count, c.num_lps)) != GPQ_OK ||
      (s = m->arena.alloc(&m->hash, c.num_lps)) != GPQ_OK ||
      (s = m->arena.alloc(&m->sum, 1)) != GPQ_OK ||
      (c.checks && (s = m->arena.alloc(&m->last_ts, c.num_lps)) != GPQ_OK)) {
    delete m;
    return s;
  }
  const PholdDev d = dev_of(m);
  hipLaunchKernelGGL(phold_zero_kernel, dim3(grid_for(c.num_lps, 256)), dim3(256), 0, queue->stream, d);
  if (hipGetLastError() != hipSuccess) {
    delete m;
    set_error("phold_zero_kernel launch failed");
    return GPQ_ERR_HIP;
  }
  *out = m;
  return GPQ_OK;
}

PholdModel::~PholdModel() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
}

// model_init (src/phold.cu:124-163): one initial event per LP per chunk, each chunk
// followed by queue_post_init, so a chunk never exceeds one insert per LP.
gpq_status PholdModel::init() {
  const PholdDev d = dev_of(this);
  for (uint32_t j = 0; j < cfg.init_per_lp; ++j) {
    GPQ_TRY(with_view(q, [&](auto v) -> gpq_status {
      hipLaunchKernelGGL(phold_populate_kernel<decltype(v)>, dim3(grid_for(cfg.num_lps, 256)), dim3(256), 0,
                         q->stream, v, d, j);
      GPQ_HIP(hipGetLastError());
      return GPQ_OK;
    }));
    GPQ_TRY(q->post_init());
  }
  if (is_calendar()) GPQ_TRY(static_cast<LaqQueue*>(q)->calendar_start());
  GPQ_HIP(hipStreamSynchronize(q->stream));
  uint32_t flags = 0;
  GPQ_TRY(q->error_flags(&flags));
  if (flags) {
    set_error("queue overflow while populating (flags " + std::to_string(flags) + ")");
    return GPQ_ERR_CAPACITY;
  }
  return GPQ_OK;
}

bool PholdModel::is_calendar() const {
  return q->cfg.kind == GPQ_LAQ && static_cast<const LaqQueue*>(q)->calendar();
}

gpq_status PholdModel::launch_fused(uint32_t settle) {
  auto* laq = static_cast<LaqQueue*>(q);
  if (laq->calendar()) {     // nothing is staged between calendar steps: no closing fold
    if (settle) return GPQ_OK;
    const uint32_t g = cal::kStepListBlocks;
    hipLaunchKernelGGL(phold_calendar_kernel<false>, dim3(g + ceil_div(cfg.num_lps, cal::kBlock)), dim3(cal::kBlock), 0,
                       q->stream, laq->view(), dev_of(this), cfg.end_time, g, nullptr, nullptr);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }
  hipLaunchKernelGGL((phold_fused_kernel<8, laq::kSlots>), dim3(laq->partial_slots), dim3(kFusedBlock), 0, q->stream,
                     laq->view(), dev_of(this), settle);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

// Folds the last step's staged events into the rings and far stores, so the queue holds
// every pending event as after a queue_post.
gpq_status PholdModel::settle() {
  if (cfg.drain != 2 || is_calendar()) return GPQ_OK;
  GPQ_TRY(launch_fused(1));
  return static_cast<LaqQueue*>(q)->reset_stage_select();
}

gpq_status PholdModel::launch_step() {
  if (cfg.drain == 2) return launch_fused(0);
  const PholdDev d = dev_of(this);
  return with_view(q, [&](auto v) -> gpq_status {
    hipLaunchKernelGGL(phold_step_kernel<decltype(v)>, dim3(grid_for(q->peek_slots(), kStepBlock)),
                       dim3(kStepBlock), 0, q->stream, v, d);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  });
}

// One des_run iteration without the host min-ts read: queue_pre, model step, queue_post.
gpq_status PholdModel::step() {
  GPQ_TRY(q->pre());
  GPQ_TRY(launch_step());
  GPQ_TRY(q->post());
  return GPQ_OK;
}

gpq_status PholdModel::events(uint64_t* out) {
  GPQ_HIP(hipMemsetAsync(sum, 0, sizeof(uint64_t), q->stream));
  hipLaunchKernelGGL(phold_sum_kernel, dim3(grid_for(cfg.num_lps, 256)), dim3(256), 0, q->stream, count,
                     cfg.num_lps, reinterpret_cast<unsigned long long*>(sum));
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(out, sum, sizeof(uint64_t), hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipStreamSynchronize(q->stream));
  return GPQ_OK;
}

gpq_status PholdModel::lp_state(uint32_t* counts, uint64_t* hashes) {
  if (counts)
    GPQ_HIP(hipMemcpyAsync(counts, count, sizeof(uint32_t) * cfg.num_lps, hipMemcpyDeviceToHost, q->stream));
  if (hashes)
    GPQ_HIP(hipMemcpyAsync(hashes, hash, sizeof(uint64_t) * cfg.num_lps, hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipStreamSynchronize(q->stream));
  return GPQ_OK;
}

// Combined hash over LPs in order: the same fold the CPU reference applies.
gpq_status PholdModel::trace_hash(uint64_t* out) {
  std::vector<uint32_t> c(cfg.num_lps);
  std::vector<uint64_t> h(cfg.num_lps);
  GPQ_TRY(lp_state(c.data(), h.data()));
  uint64_t x = 0x6A09E667F3BCC908ull;
  for (uint32_t i = 0; i < cfg.num_lps; ++i) {
    x = host_mix64(x ^ h[i]);
    x = host_mix64(x ^ (((uint64_t)c[i] << 32) | i));
  }
  *out = x;
  return GPQ_OK;
}

// Captures `batch` steps (queue_pre, model step, queue_post) into a graph, once per batch size.
gpq_status PholdModel::prepare_graph(uint32_t batch) {
  GPQ_CHECK_ARG(batch >= 1, "batch must be >= 1");
  if (graph_exec && graph_batch == batch) return GPQ_OK;
  if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
  if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  GPQ_HIP(hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
  gpq_status s = GPQ_OK;
  for (uint32_t i = 0; i < batch && s == GPQ_OK; ++i) s = step();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(q->stream, &g);
  if (s != GPQ_OK) return s;
  GPQ_HIP(e);
  graph = g;
  GPQ_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
  graph_batch = batch;
  return GPQ_OK;
}

// Non-blocking drive for callers that advance many models at once: one graph replay, the
// window state copied to the queue's pinned mirror, and `done` recorded behind both.
gpq_status PholdModel::launch_async(hipEvent_t done) {
  GPQ_CHECK_ARG(graph_exec, "launch_async before prepare_graph");
  GPQ_HIP(hipGraphLaunch(graph_exec, q->stream));
  GPQ_HIP(hipMemcpyAsync(q->host_ws, q->ws, sizeof(WindowState), hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipEventRecord(done, q->stream));
  return GPQ_OK;
}

// des_run (src/des.cu:134-221).  Plain: the reference's pattern, one host round trip
// per step for the minimum timestamp.  Graph: `batch` steps captured once and replayed;
// after the last window, remaining steps of a batch find nothing to commit.
gpq_status PholdModel::run(gpq_drive_mode mode, uint32_t batch, gpq_des_report* rep) {
  GPQ_CHECK_ARG(batch >= 1, "batch must be >= 1");
  GPQ_HIP(hipStreamSynchronize(q->stream));
  uint64_t steps = 0;
  WindowState w{};
  const auto t0 = std::chrono::steady_clock::now();
  if (mode == GPQ_DRIVE_PLAIN) {
    for (;;) {
      GPQ_TRY(q->window(&w));          // queue_get_min_ts: D2H + sync every step
      if (w.done || w.error) break;
      GPQ_TRY(step());
      ++steps;
    }
  } else {
    GPQ_TRY(prepare_graph(batch));
    for (;;) {
      GPQ_TRY(q->window(&w));
      if (w.done || w.error) break;
      GPQ_HIP(hipGraphLaunch(graph_exec, q->stream));
      steps += batch;
    }
  }
  GPQ_HIP(hipStreamSynchronize(q->stream));
  const auto t1 = std::chrono::steady_clock::now();
  GPQ_TRY(settle());
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->steps = steps;
    rep->windows = w.windows;
    rep->sim_seconds = std::chrono::duration<double>(t1 - t0).count();
    GPQ_TRY(events(&rep->events));
    rep->events_per_second = rep->sim_seconds > 0 ? (double)rep->events / rep->sim_seconds : 0.0;
    std::memcpy(&rep->final_min_ts, &w.gmin, sizeof(float));
    rep->error_flags = w.error;
  }
  if (w.error) {
    set_error("device error flags " + std::to_string(w.error) + " (1 staging, 2 queue, 4 heap, 8 order)");
    return (w.error & kErrBadState) ? GPQ_ERR_STATE : GPQ_ERR_CAPACITY;
  }
  return GPQ_OK;
}

// Per-kernel timing: steps launched eagerly in groups of `batch`, five events per step
// (before queue_pre, before the model step kernel, after it, before the window publication
// of a per-LP queue's post, after queue_post).
gpq_status PholdModel::profile(uint32_t batch, gpq_kernel_profile* out) {
  GPQ_CHECK_ARG(batch >= 1 && batch <= 4096, "profile batch must be in [1, 4096]");
  std::vector<hipEvent_t> ev(5 * (size_t)batch, nullptr);
  for (auto& e : ev) GPQ_HIP(hipEventCreate(&e));
  struct Release {
    std::vector<hipEvent_t>& v;
    QueueBase* q;
    ~Release() {
      q->profile_mark = nullptr;
      for (auto e : v) if (e) (void)hipEventDestroy(e);
    }
  } release{ev, q};
  const bool marked = q->cfg.kind == GPQ_LAQ || q->cfg.kind == GPQ_LHQ || q->cfg.kind == GPQ_LSQ;
  GPQ_HIP(hipStreamSynchronize(q->stream));
  double step_ms = 0.0, queue_ms = 0.0, window_ms = 0.0;
  uint64_t steps = 0;
  WindowState w{};
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    GPQ_TRY(q->window(&w));
    if (w.done || w.error) break;
    for (uint32_t i = 0; i < batch; ++i) {
      hipEvent_t* e = &ev[5 * (size_t)i];
      GPQ_HIP(hipEventRecord(e[0], q->stream));
      GPQ_TRY(q->pre());
      GPQ_HIP(hipEventRecord(e[1], q->stream));
      GPQ_TRY(launch_step());
      GPQ_HIP(hipEventRecord(e[2], q->stream));
      q->profile_mark = marked ? e[3] : nullptr;
      GPQ_TRY(q->post());
      q->profile_mark = nullptr;
      GPQ_HIP(hipEventRecord(e[4], q->stream));
    }
    GPQ_HIP(hipEventSynchronize(ev[5 * (size_t)(batch - 1) + 4]));
    for (uint32_t i = 0; i < batch; ++i) {
      hipEvent_t* e = &ev[5 * (size_t)i];
      float a = 0, b = 0, c = 0, d = 0;
      GPQ_HIP(hipEventElapsedTime(&a, e[0], e[1]));
      GPQ_HIP(hipEventElapsedTime(&b, e[1], e[2]));
      if (marked) {
        GPQ_HIP(hipEventElapsedTime(&c, e[2], e[3]));
        GPQ_HIP(hipEventElapsedTime(&d, e[3], e[4]));
      } else {
        GPQ_HIP(hipEventElapsedTime(&c, e[2], e[4]));
      }
      step_ms += b;
      queue_ms += a + c;
      window_ms += d;
    }
    steps += batch;
  }
  const auto t1 = std::chrono::steady_clock::now();
  std::memset(out, 0, sizeof(*out));
  out->steps = steps;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  out->step_kernel_us = steps ? 1000.0 * step_ms / (double)steps : 0.0;
  out->post_kernel_us = steps ? 1000.0 * queue_ms / (double)steps : 0.0;
  out->window_kernel_us = steps ? 1000.0 * window_ms / (double)steps : 0.0;
  GPQ_TRY(settle());
  GPQ_TRY(events(&out->events));
  if (w.error) {
    set_error("device error flags " + std::to_string(w.error));
    return GPQ_ERR_CAPACITY;
  }
  return GPQ_OK;
}

gpq_status launch_calendar_instances(LaqQueue* q, const PoolInst* inst, const uint32_t* block_inst, uint32_t* count,
                                     uint64_t* hash) {
  GPQ_CHECK_ARG(q->calendar() && inst && block_inst, "calendar step over instances: needs a calendar-mode LAQ");
  PholdDev d{};
  d.count = count;
  d.hash = hash;
  d.ws = q->ws;
  d.num_lps = q->cfg.num_lps;
  d.drain = 2;
  d.lookahead = q->cfg.lookahead;
  const uint32_t g = cal::kStepListBlocks;
  hipLaunchKernelGGL(phold_calendar_kernel<true>, dim3(g + ceil_div(q->cfg.num_lps, cal::kBlock)), dim3(cal::kBlock), 0,
                     q->stream, q->view(), d, q->cfg.horizon, g, inst, block_inst);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status launch_populate_instances(LaqQueue* q, const PoolInst* inst, const uint32_t* block_inst, uint32_t j) {
  GPQ_CHECK_ARG(q->calendar() && inst && block_inst, "populate over instances: needs a calendar-mode LAQ");
  hipLaunchKernelGGL(populate_instances_kernel, dim3(ceil_div(q->cfg.num_lps, 256)), dim3(256), 0, q->stream,
                     q->view(), inst, block_inst, j);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

}  // namespace gpq

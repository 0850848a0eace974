// phold_batch.hip -- the PHOLD pool: many instances in one LAQ, advanced together.
//
// Per step, over the whole pool:
//   pool_step_kernel     every LP of an active instance drains its safe window (the
//                        instance's own window end), exactly as one instance alone would:
//                        Philox counters, destinations and the (ts, src, seq) identity use
//                        the LP's index inside its instance, so each instance's trace hash
//                        equals the single-instance run's;
//   LaqQueue::merge_only the LAQ merge of the staged inserts, one partial minimum per block;
//   pool_publish_kernel  one workgroup per instance folds the partials of its blocks into
//                        its window (gmin, window end, done).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <memory>

#include "phold_batch.h"

namespace gpq {

namespace {

constexpr uint32_t kBlk = LaqQueue::kBlockLps;

uint64_t host_mix64_b(uint64_t x) {
  x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27; x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// 256 threads = one post block of LPs; an instance owns whole blocks, so its slot, window and
// parameters are block-uniform (block_win, filled at admission).
__global__ __launch_bounds__(256) void pool_step_kernel(LaqView q, const PoolInst* inst, const WindowState* ws,
                                                        const uint32_t* block_win, uint32_t* count, uint64_t* hash,
                                                        uint32_t capacity, uint32_t* error) {
  if (!q.block_live[blockIdx.x]) return;
  const uint32_t i = block_win[blockIdx.x];
  if (ws[i].done) return;
  const uint32_t wend = ws[i].wend;
  const PoolInst in = inst[i];
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t local = lp - in.base;
  if (lp >= capacity || local >= in.lps) return;         // the instance's last block may be partly unused
  if (!(q.head[lp] < wend)) return;
  const uint32_t mask = (1u << q.cap_log2) - 1u;
  const Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  uint32_t rd = q.rd[lp], len = q.len[lp];
  uint32_t n = count[lp];
  uint64_t h = hash[lp];
  Event cur = load_event(&ring[rd]);
  while (len && cur.ts < wend) {
    const Event nxt = len > 1 ? load_event(&ring[(rd + 1) & mask]) : Event{kInfBits, 0u, 0u, 0u};
    h = trace_step(h, cur.ts, cur.src, cur.seq);
    const uint4 r = philox4x32_10(make_uint4(n, local, 0u, 0u), make_uint2(in.seed, 0u));
    const float ex = exp_from_u32(r.x);
    const uint32_t dst_local = r.y < in.p_self ? local : (uint32_t)(((uint64_t)r.z * in.lps) >> 32);
    const float t_new = (bits_to_f(cur.ts) + in.lookahead) + ex;
    q.insert(Event{f_to_bits(t_new), in.base + dst_local, local, in.init_per_lp + n});
    ++n;
    rd = (rd + 1) & mask;
    --len;
    cur = nxt;
  }
  q.rd[lp] = rd;
  q.len[lp] = len;
  q.head[lp] = len ? cur.ts : kInfBits;
  count[lp] = n;
  hash[lp] = h;
  (void)error;
}

// One workgroup per instance slot: the minimum over its post blocks' partials.
__global__ void pool_publish_kernel(const uint32_t* partial, const PoolInst* inst, WindowState* ws, uint32_t slots) {
  const uint32_t i = blockIdx.x;
  if (i >= slots) return;
  const PoolInst in = inst[i];
  if (!in.active || ws[i].done) return;
  __shared__ uint32_t s_w[4];
  uint32_t m = kInfBits;
  for (uint32_t b = in.block_begin + threadIdx.x; b < in.block_end; b += blockDim.x) m = partial[b] < m ? partial[b] : m;
  m = wave_min_u32(m);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t g = s_w[0];
    for (uint32_t k = 1; k < blockDim.x / kWave; ++k) g = s_w[k] < g ? s_w[k] : g;
    float wend = bits_to_f(g) + in.lookahead;
    if (in.end_time < wend) wend = in.end_time;
    WindowState w = ws[i];
    w.gmin = g;
    w.wend = f_to_bits(wend);
    w.done = (g == kInfBits || !(bits_to_f(g) < in.end_time)) ? 1u : 0u;
    w.windows += 1;
    ws[i] = w;
  }
}

// Two-tier pool: one workgroup per instance; gmin = min(ring heads, far minima) over its
// blocks, the window is capped at the instance's near horizon, and when the next window could
// reach it, the move to gmin + 2 lookahead + 4 lookahead is scheduled for the next post.
__global__ void pool_far_publish_kernel(const uint32_t* partial, uint32_t nb, const PoolInst* inst, WindowState* ws,
                                        uint32_t slots) {
  const uint32_t i = blockIdx.x;
  if (i >= slots) return;
  const PoolInst in = inst[i];
  if (!in.active || ws[i].done) return;
  __shared__ uint32_t s_a[4], s_b[4];
  uint32_t a = kInfBits, b = kInfBits;
  for (uint32_t k = in.block_begin + threadIdx.x; k < in.block_end; k += blockDim.x) {
    a = partial[k] < a ? partial[k] : a;
    b = partial[nb + k] < b ? partial[nb + k] : b;
  }
  a = wave_min_u32(a);
  b = wave_min_u32(b);
  if (lane_id() == 0) {
    s_a[threadIdx.x / kWave] = a;
    s_b[threadIdx.x / kWave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t k = 0; k < blockDim.x / kWave; ++k) {
      a = s_a[k] < a ? s_a[k] : a;
      b = s_b[k] < b ? s_b[k] : b;
    }
    WindowState w = ws[i];
    if (w.near_next) {
      w.near_end = w.near_next;
      w.near_next = 0;
    }
    const uint32_t g = a < b ? a : b;
    const float gf = bits_to_f(g);
    float wend = gf + in.lookahead;
    if (in.end_time < wend) wend = in.end_time;
    const float nearf = bits_to_f(w.near_end);
    if (nearf < wend) wend = nearf;
    if (g != kInfBits && gf + 2.0f * in.lookahead > nearf) {   // one window ahead, as the single queue does
      float next = gf + 6.0f * in.lookahead;
      if (!(next > nearf)) next = nearf + 4.0f * in.lookahead;
      w.near_next = f_to_bits(next);
    }
    w.gmin = g;
    w.wend = f_to_bits(wend);
    w.done = (g == kInfBits || !(gf < in.end_time)) ? 1u : 0u;
    w.windows += 1;
    ws[i] = w;
  }
}

__global__ void pool_fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Admission: the instance's LPs get their initial events directly in sorted order (the same
// ring contents as init_per_lp rounds of populate + queue_post_init), padding LPs of its
// last block are emptied, and the counters and hashes start at zero.
__global__ void pool_populate_kernel(LaqView q, PoolInst in, uint32_t* count, uint64_t* hash, uint32_t padded) {
  const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= padded) return;
  const uint32_t lp = in.base + local;
  Event* ring = q.ring + ((size_t)lp << q.cap_log2);
  q.rd[lp] = 0;
  q.ins[lp] = 0;
  if (q.far_cap) {
    q.nfar[lp] = 0;
    q.far_min[lp] = kInfBits;
  }
  count[lp] = 0;
  hash[lp] = 0;
  if (local >= in.lps) {
    q.len[lp] = 0;
    q.head[lp] = kInfBits;
    return;
  }
  const uint32_t k = in.init_per_lp;
  for (uint32_t j = 0; j < k; ++j) {
    const uint4 r = philox4x32_10(make_uint4(j, local, 1u, 0u), make_uint2(in.seed, 0u));
    Event e{f_to_bits(exp_from_u32(r.x)), lp, local, j};
    uint32_t p = j;                        // insertion into the sorted prefix
    while (p > 0) {
      const Event prev = load_event(&ring[p - 1]);
      if (!key_less(e, prev)) break;
      store_event(&ring[p], prev);
      --p;
    }
    store_event(&ring[p], e);
  }
  q.len[lp] = k;
  q.head[lp] = k ? ring[0].ts : kInfBits;
}

// The first window of a freshly admitted instance: the minimum head over its LPs.
__global__ void pool_first_window_kernel(const LaqView q, PoolInst in, WindowState* ws, uint32_t slot) {
  __shared__ uint32_t s_w[16];
  uint32_t m = kInfBits;
  for (uint32_t l = threadIdx.x; l < in.lps; l += blockDim.x) {
    const uint32_t h = q.head[in.base + l];
    m = h < m ? h : m;
  }
  m = wave_min_u32(m);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t g = s_w[0];
    for (uint32_t k = 1; k < blockDim.x / kWave; ++k) g = s_w[k] < g ? s_w[k] : g;
    float wend = bits_to_f(g) + in.lookahead;
    if (in.end_time < wend) wend = in.end_time;
    // two-tier: every record is in the ring, so the near horizon starts past the first window
    const uint32_t near_end = q.far_cap ? f_to_bits(bits_to_f(g) + 5.0f * in.lookahead) : 0u;
    ws[slot] = WindowState{g, f_to_bits(wend), (g == kInfBits || !(bits_to_f(g) < in.end_time)) ? 1u : 0u, near_end,
                           1ull, 0u, 0u};
  }
}

__global__ void pool_release_kernel(uint32_t base, uint32_t n, LaqView q) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < n) {
    q.len[base + l] = 0;
    q.ins[base + l] = 0;
    q.head[base + l] = kInfBits;
    if (q.far_cap) {
      q.nfar[base + l] = 0;
      q.far_min[base + l] = kInfBits;
    }
  }
}

__global__ void pool_init_kernel(WindowState* ws, uint32_t slots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < slots) ws[i] = WindowState{kInfBits, 0u, 1u, 0u, 0ull, 0u, 0u};
}

void fold_instances(const std::vector<PoolInst>& in, const uint32_t* counts, const uint64_t* hashes,
                    uint64_t* trace_hash, uint64_t* events);

}  // namespace

uint64_t PholdPool::bytes_per_lp(uint32_t capacity_per_lp, uint32_t stage_per_lp, uint32_t far_per_lp) {
  return (uint64_t)(capacity_per_lp + stage_per_lp + far_per_lp) * sizeof(Event) + 4ull * 4 + 4 + 8 +
         (far_per_lp ? 8ull : 0ull);
}

uint64_t PholdPool::device_bytes() const { return (q_ ? q_->arena.bytes() : 0) + arena_.bytes(); }

gpq_status PholdPool::create(uint32_t capacity_lps, uint32_t capacity_per_lp, uint32_t stage_per_lp,
                             uint32_t max_instances, PholdPool** out, uint32_t far_per_lp) {
  GPQ_CHECK_ARG(capacity_lps >= 1 && max_instances >= 1, "pool: empty");
  const uint32_t cap = (uint32_t)ceil_div(capacity_lps, kBlk) * kBlk;
  gpq_queue_config qc{};
  qc.kind = GPQ_LAQ;
  qc.num_lps = cap;
  qc.capacity_per_lp = capacity_per_lp;
  qc.stage_per_lp = stage_per_lp;
  qc.lookahead = 0.1f;
  qc.horizon = 1.0f;
  qc.block_threads = 256;
  qc.far_per_lp = far_per_lp;
  auto p = std::make_unique<PholdPool>();
  p->q_ = new LaqQueue(qc);
  GPQ_TRY(p->q_->init_common());
  GPQ_TRY(p->q_->init());
  p->capacity_ = cap;
  p->max_inst_ = max_instances;
  GPQ_TRY(p->arena_.alloc(&p->d_inst_, max_instances));
  GPQ_TRY(p->arena_.alloc(&p->d_ws_, max_instances));
  GPQ_TRY(p->arena_.alloc(&p->d_block_live_, cap / kBlk));
  GPQ_HIP(hipMemsetAsync(p->d_block_live_, 0, cap / kBlk, p->q_->stream));
  p->q_->set_block_live(p->d_block_live_);
  p->far_ = far_per_lp != 0;
  GPQ_TRY(p->arena_.alloc(&p->d_block_win_, cap / kBlk));
  GPQ_HIP(hipMemsetAsync(p->d_block_win_, 0, sizeof(uint32_t) * (cap / kBlk), p->q_->stream));
  if (p->far_) p->q_->set_block_windows(p->d_block_win_, p->d_ws_);
  GPQ_TRY(p->arena_.alloc(&p->d_count_, cap));
  GPQ_TRY(p->arena_.alloc(&p->d_hash_, cap));
  GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_ws_), sizeof(WindowState) * max_instances,
                        hipHostMallocDefault));
  GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_count_), sizeof(uint32_t) * cap, hipHostMallocDefault));
  GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_hash_), sizeof(uint64_t) * cap, hipHostMallocDefault));
  GPQ_HIP(hipEventCreateWithFlags(&p->done_, hipEventDisableTiming));
  p->inst_.assign(max_instances, PoolInst{});
  p->block_used_.assign(cap / kBlk, 0);
  GPQ_HIP(hipMemcpyAsync(p->d_inst_, p->inst_.data(), sizeof(PoolInst) * max_instances, hipMemcpyHostToDevice,
                         p->q_->stream));
  hipLaunchKernelGGL(pool_init_kernel, dim3(ceil_div(max_instances, 256)), dim3(256), 0, p->q_->stream, p->d_ws_,
                     max_instances);
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(p->h_ws_, p->d_ws_, sizeof(WindowState) * max_instances, hipMemcpyDeviceToHost,
                         p->q_->stream));
  GPQ_HIP(hipStreamSynchronize(p->q_->stream));
  *out = p.release();
  return GPQ_OK;
}

PholdPool::~PholdPool() {
  if (q_) (void)hipStreamSynchronize(q_->stream);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (done_) (void)hipEventDestroy(done_);
  if (h_ws_) (void)hipHostFree(h_ws_);
  if (h_count_) (void)hipHostFree(h_count_);
  if (h_hash_) (void)hipHostFree(h_hash_);
  arena_.release();
  delete q_;
}

uint32_t PholdPool::free_lps() const {
  uint32_t f = 0;
  for (uint8_t u : block_used_) f += u ? 0u : kBlk;
  return f;
}

gpq_status PholdPool::admit(const gpq_phold_instance& c, uint32_t* slot, bool* ok) {
  *ok = false;
  GPQ_CHECK_ARG(c.num_lps >= 1, "pool: instance without LPs");
  GPQ_CHECK_ARG(c.init_per_lp <= (1u << q_->view().cap_log2), "pool: init_per_lp exceeds the LP capacity");
  GPQ_CHECK_ARG(c.lookahead > 0.0f && c.end_time > 0.0f, "pool: lookahead and end_time must be > 0");
  GPQ_CHECK_ARG(!pending_, "pool: admit while a replay is in flight");
  uint32_t s = 0;
  while (s < max_inst_ && inst_[s].active) ++s;
  if (s == max_inst_) return GPQ_OK;
  const uint32_t need = (uint32_t)ceil_div(c.num_lps, kBlk);
  const uint32_t nb = (uint32_t)block_used_.size();
  uint32_t run = 0, start = 0;
  bool found = false;
  for (uint32_t b = 0; b < nb; ++b) {
    if (block_used_[b]) { run = 0; continue; }
    if (run == 0) start = b;
    if (++run == need) { found = true; break; }
  }
  if (!found) return GPQ_OK;
  for (uint32_t b = start; b < start + need; ++b) block_used_[b] = 1;
  GPQ_HIP(hipMemsetAsync(d_block_live_ + start, 1, need, q_->stream));
  hipLaunchKernelGGL(pool_fill_u32_kernel, dim3(ceil_div(need, 256)), dim3(256), 0, q_->stream, d_block_win_ + start,
                     need, s);
  GPQ_HIP(hipGetLastError());
  PoolInst in{};
  in.base = start * kBlk;
  in.lps = c.num_lps;
  in.init_per_lp = c.init_per_lp;
  in.seed = c.seed;
  in.end_time = c.end_time;
  in.lookahead = c.lookahead;
  in.p_self = c.p_self_u32;
  in.active = 1;
  in.block_begin = start;
  in.block_end = start + need;
  inst_[s] = in;
  const LaqView v = q_->view();
  const uint32_t padded = need * kBlk;
  hipLaunchKernelGGL(pool_populate_kernel, dim3(ceil_div(padded, 256)), dim3(256), 0, q_->stream, v, in, d_count_,
                     d_hash_, padded);
  hipLaunchKernelGGL(pool_first_window_kernel, dim3(1), dim3(1024), 0, q_->stream, v, in, d_ws_, s);
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(d_inst_ + s, &inst_[s], sizeof(PoolInst), hipMemcpyHostToDevice, q_->stream));
  GPQ_HIP(hipMemcpyAsync(h_ws_ + s, d_ws_ + s, sizeof(WindowState), hipMemcpyDeviceToHost, q_->stream));
  GPQ_HIP(hipStreamSynchronize(q_->stream));
  *slot = s;
  *ok = true;
  return GPQ_OK;
}

gpq_status PholdPool::prepare(uint32_t batch) {
  GPQ_CHECK_ARG(batch >= 1, "pool: batch must be >= 1");
  if (graph_exec_ && batch_ == batch) return GPQ_OK;
  if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
  if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
  const LaqView v = q_->view();
  GPQ_HIP(hipStreamBeginCapture(q_->stream, hipStreamCaptureModeThreadLocal));
  gpq_status st = GPQ_OK;
  for (uint32_t i = 0; i < batch && st == GPQ_OK; ++i) {
    hipLaunchKernelGGL(pool_step_kernel, dim3(ceil_div(capacity_, 256)), dim3(256), 0, q_->stream, v, d_inst_, d_ws_,
                       d_block_win_, d_count_, d_hash_, capacity_, &q_->ws->error);
    st = q_->merge_only();
    if (st == GPQ_OK && far_)
      hipLaunchKernelGGL(pool_far_publish_kernel, dim3(max_inst_), dim3(256), 0, q_->stream, q_->partial,
                         capacity_ / kBlk, d_inst_, d_ws_, max_inst_);
    else if (st == GPQ_OK)
      hipLaunchKernelGGL(pool_publish_kernel, dim3(max_inst_), dim3(256), 0, q_->stream, q_->partial, d_inst_, d_ws_,
                         max_inst_);
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(q_->stream, &g);
  if (st != GPQ_OK) return st;
  GPQ_HIP(e);
  graph_ = g;
  GPQ_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  batch_ = batch;
  return GPQ_OK;
}

gpq_status PholdPool::launch_async() {
  GPQ_CHECK_ARG(graph_exec_, "pool: launch before prepare");
  GPQ_HIP(hipGraphLaunch(graph_exec_, q_->stream));
  GPQ_HIP(hipMemcpyAsync(h_ws_, d_ws_, sizeof(WindowState) * max_inst_, hipMemcpyDeviceToHost, q_->stream));
  GPQ_HIP(hipMemcpyAsync(q_->host_ws, q_->ws, sizeof(WindowState), hipMemcpyDeviceToHost, q_->stream));
  GPQ_HIP(hipEventRecord(done_, q_->stream));
  pending_ = true;
  return GPQ_OK;
}

bool PholdPool::replay_done() {
  if (!pending_) return true;
  if (hipEventQuery(done_) != hipSuccess) return false;
  pending_ = false;
  return true;
}

void PholdPool::finished(std::vector<uint32_t>* slots) const {
  slots->clear();
  for (uint32_t s = 0; s < max_inst_; ++s)
    if (inst_[s].active && h_ws_[s].done) slots->push_back(s);
}

bool PholdPool::any_active() const {
  for (const auto& in : inst_)
    if (in.active) return true;
  return false;
}

uint32_t PholdPool::error_flags_host() const { return q_->host_ws->error; }

gpq_status PholdPool::errors(uint32_t* flags) { return q_->error_flags(flags); }

gpq_status PholdPool::retire(uint32_t s, uint64_t* trace_hash, uint64_t* events) {
  return retire_many(std::vector<uint32_t>{s}, trace_hash, events);
}

gpq_status PholdPool::retire_many(const std::vector<uint32_t>& slots, uint64_t* trace_hash, uint64_t* events) {
  GPQ_CHECK_ARG(!pending_, "pool: retire while a replay is in flight");
  for (size_t k = 0; k < slots.size(); ++k) {
    GPQ_CHECK_ARG(slots[k] < max_inst_ && inst_[slots[k]].active, "pool: retire of an inactive slot");
    for (size_t j = 0; j < k; ++j) GPQ_CHECK_ARG(slots[j] != slots[k], "pool: slot retired twice");
  }
  if (slots.empty()) return GPQ_OK;
  std::vector<PoolInst> in(slots.size());
  // Every instance's counters and hashes come back in one round trip (pinned staging at the
  // instance's own LP offset), then its range is released on the device.
  for (size_t k = 0; k < slots.size(); ++k) {
    const uint32_t s = slots[k];
    in[k] = inst_[s];
    GPQ_HIP(hipMemcpyAsync(h_count_ + in[k].base, d_count_ + in[k].base, sizeof(uint32_t) * in[k].lps,
                           hipMemcpyDeviceToHost, q_->stream));
    GPQ_HIP(hipMemcpyAsync(h_hash_ + in[k].base, d_hash_ + in[k].base, sizeof(uint64_t) * in[k].lps,
                           hipMemcpyDeviceToHost, q_->stream));
    inst_[s].active = 0;
    GPQ_HIP(hipMemcpyAsync(d_inst_ + s, &inst_[s], sizeof(PoolInst), hipMemcpyHostToDevice, q_->stream));
    const uint32_t padded = (in[k].block_end - in[k].block_begin) * kBlk;
    hipLaunchKernelGGL(pool_release_kernel, dim3(ceil_div(padded, 256)), dim3(256), 0, q_->stream, in[k].base,
                       padded, q_->view());
    GPQ_HIP(hipGetLastError());
    GPQ_HIP(hipMemsetAsync(d_block_live_ + in[k].block_begin, 0, in[k].block_end - in[k].block_begin, q_->stream));
  }
  GPQ_HIP(hipStreamSynchronize(q_->stream));
  for (const auto& r : in)
    for (uint32_t b = r.block_begin; b < r.block_end; ++b) block_used_[b] = 0;
  fold_instances(in, h_count_, h_hash_, trace_hash, events);
  return GPQ_OK;
}

namespace {
// The trace hash of one instance is a sequential fold over its LPs (the single-instance
// definition); instances fold on separate host threads, largest first.  counts and hashes are
// indexed by the LP's index in the shared queue.
void fold_instances(const std::vector<PoolInst>& in, const uint32_t* counts, const uint64_t* hashes,
                    uint64_t* trace_hash, uint64_t* events) {
  auto fold = [&](size_t k) {
    const uint32_t* c = counts + in[k].base;
    const uint64_t* h = hashes + in[k].base;
    uint64_t x = 0x6A09E667F3BCC908ull, ev = 0;
    for (uint32_t i = 0; i < in[k].lps; ++i) {
      x = host_mix64_b(x ^ h[i]);
      x = host_mix64_b(x ^ (((uint64_t)c[i] << 32) | i));
      ev += c[i];
    }
    if (trace_hash) trace_hash[k] = x;
    if (events) events[k] = ev;
  };
  uint64_t total = 0;
  for (const auto& r : in) total += r.lps;
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const unsigned nthreads = (unsigned)std::min<uint64_t>({(uint64_t)hw, (uint64_t)in.size(), total / 16384 + 1});
  if (nthreads <= 1) {
    for (size_t k = 0; k < in.size(); ++k) fold(k);
    return;
  }
  std::vector<size_t> order(in.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return in[a].lps > in[b].lps; });
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  pool.reserve(nthreads);
  for (unsigned t = 0; t < nthreads; ++t)
    pool.emplace_back([&]() {
      for (size_t i; (i = next.fetch_add(1)) < order.size();) fold(order[i]);
    });
  for (auto& t : pool) t.join();
}
}  // namespace

// ---------------------------------------------------------------------------------------
// PholdBatch
// ---------------------------------------------------------------------------------------
gpq_status PholdBatch::create(const gpq_phold_instance* in, uint32_t count, uint32_t cap, uint32_t stage,
                              uint32_t far, PholdBatch** out) {
  GPQ_CHECK_ARG(in && count >= 1, "phold batch: no instances");
  for (uint32_t i = 0; i < count; ++i) {
    GPQ_CHECK_ARG(in[i].num_lps >= 1 && in[i].init_per_lp >= 1, "phold batch: empty instance");
    GPQ_CHECK_ARG(in[i].lookahead > 0.0f && in[i].end_time > 0.0f, "phold batch: lookahead/end_time must be > 0");
  }
  auto b = std::make_unique<PholdBatch>();
  b->inst.assign(in, in + count);
  b->capacity_per_lp = cap ? cap : 128;
  b->stage_per_lp = stage ? stage : 32;
  b->far_per_lp = far;
  b->hash.assign(count, 0);
  b->events.assign(count, 0);
  *out = b.release();
  return GPQ_OK;
}

gpq_status PholdBatch::run(gpq_drive_mode mode, uint32_t batch, double* seconds) {
  const uint32_t n = (uint32_t)inst.size();
  double sec = 0.0;
  if (!calendar_windows)      // the sorted ring holds every initial event of an LP
    for (const auto& c : inst)
      GPQ_CHECK_ARG(c.init_per_lp <= capacity_per_lp, "phold batch: init_per_lp exceeds capacity_per_lp");
  if (mode != GPQ_DRIVE_PLAIN && calendar_windows) return run_calendar_union(batch, seconds);
  if (mode == GPQ_DRIVE_PLAIN) {
    for (uint32_t i = 0; i < n; ++i) {
      gpq_queue_config qc{};
      qc.kind = GPQ_LAQ;
      qc.num_lps = inst[i].num_lps;
      qc.capacity_per_lp = capacity_per_lp;
      qc.stage_per_lp = stage_per_lp;
      qc.far_per_lp = far_per_lp;
      qc.lookahead = inst[i].lookahead;
      qc.horizon = inst[i].end_time;
      qc.block_threads = 256;
      qc.calendar_windows = calendar_windows;
      std::unique_ptr<QueueBase> q(make_queue(qc));
      GPQ_TRY(q->init_common());
      GPQ_TRY(q->init());
      gpq_phold_config pc{inst[i].num_lps, inst[i].init_per_lp, inst[i].seed, inst[i].lookahead,
                          inst[i].end_time, inst[i].p_self_u32, 0u, calendar_windows ? 2u : 1u};
      PholdModel* raw = nullptr;
      GPQ_TRY(PholdModel::create(q.get(), pc, &raw));
      std::unique_ptr<PholdModel> m(raw);
      GPQ_TRY(m->init());
      gpq_des_report rep;
      GPQ_TRY(m->run(GPQ_DRIVE_PLAIN, 1, &rep));
      sec += rep.sim_seconds;
      GPQ_TRY(m->trace_hash(&hash[i]));
      events[i] = rep.events;
    }
  } else {
    uint64_t total = 0;
    for (const auto& c : inst) total += ceil_div(c.num_lps, kBlk) * (uint64_t)kBlk;
    GPQ_CHECK_ARG(total < (1ull << 31) / capacity_per_lp, "phold batch: too many LPs for one pool");
    PholdPool* raw = nullptr;
    GPQ_TRY(PholdPool::create((uint32_t)total, capacity_per_lp, stage_per_lp, n, &raw, far_per_lp));
    std::unique_ptr<PholdPool> pool(raw);
    std::vector<uint32_t> slot_of(n), req_of(n);
    for (uint32_t i = 0; i < n; ++i) {
      bool ok = false;
      GPQ_TRY(pool->admit(inst[i], &slot_of[i], &ok));
      GPQ_CHECK_ARG(ok, "phold batch: pool admission failed");
      req_of[slot_of[i]] = i;
    }
    GPQ_TRY(pool->prepare(batch ? batch : 16));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> fin;
    for (;;) {
      pool->finished(&fin);
      if (pool->error_flags_host()) {
        set_error("phold batch: device error flags " + std::to_string(pool->error_flags_host()));
        return GPQ_ERR_CAPACITY;
      }
      if (!fin.empty()) {
        std::vector<uint64_t> fh(fin.size()), fe(fin.size());
        GPQ_TRY(pool->retire_many(fin, fh.data(), fe.data()));
        for (size_t k = 0; k < fin.size(); ++k) {
          hash[req_of[fin[k]]] = fh[k];
          events[req_of[fin[k]]] = fe[k];
        }
      }
      if (!pool->any_active()) break;
      GPQ_TRY(pool->launch_async());
      while (!pool->replay_done()) {
      }
    }
    sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (seconds) *seconds = sec;
  return GPQ_OK;
}

// All instances in one calendar-mode LAQ on one window sequence.  The timed part is the step
// loop, as for each instance run alone (population and graph capture come before it, reading
// back and folding the instances' counters and hashes after it).
gpq_status PholdBatch::run_calendar_union(uint32_t batch, double* seconds) {
  const uint32_t n = (uint32_t)inst.size();
  constexpr uint32_t kCalBlk = LaqQueue::kCalBlockLps;   // an instance owns whole commit blocks
  std::vector<PoolInst> tab(n);
  uint64_t total = 0;
  uint32_t max_init = 0;
  float la = inst[0].lookahead, horizon = inst[0].end_time;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t blocks = (uint32_t)ceil_div(inst[i].num_lps, kCalBlk);
    PoolInst& t = tab[i];
    t = PoolInst{};
    t.base = (uint32_t)total;
    t.lps = inst[i].num_lps;
    t.init_per_lp = inst[i].init_per_lp;
    t.seed = inst[i].seed;
    t.end_time = inst[i].end_time;
    t.lookahead = inst[i].lookahead;
    t.p_self = inst[i].p_self_u32;
    t.active = 1;
    t.block_begin = (uint32_t)(total / kCalBlk);
    t.block_end = t.block_begin + blocks;
    total += (uint64_t)blocks * kCalBlk;
    max_init = std::max(max_init, inst[i].init_per_lp);
    la = std::min(la, inst[i].lookahead);
    horizon = std::max(horizon, inst[i].end_time);
  }
  GPQ_CHECK_ARG(total < (1ull << 31), "phold batch: too many LPs for one queue");
  GPQ_CHECK_ARG(far_per_lp >= max_init, "phold batch (calendar): far_per_lp must hold every initial event of an LP "
                "(the parked list takes an LP's events past its end time)");
  gpq_queue_config qc{};
  qc.kind = GPQ_LAQ;
  qc.num_lps = (uint32_t)total;
  qc.capacity_per_lp = capacity_per_lp;
  qc.stage_per_lp = stage_per_lp;
  qc.far_per_lp = far_per_lp;
  qc.lookahead = la;
  qc.horizon = horizon;
  qc.block_threads = 256;
  qc.calendar_windows = calendar_windows;
  std::unique_ptr<LaqQueue> q(new LaqQueue(qc));
  GPQ_TRY(q->init_common());
  GPQ_TRY(q->init());
  DeviceArena arena;
  PoolInst* d_inst = nullptr;
  uint32_t* d_block = nullptr;
  uint32_t* d_count = nullptr;
  uint64_t* d_hash = nullptr;
  const uint32_t nblocks = (uint32_t)(total / kCalBlk);
  GPQ_TRY(arena.alloc(&d_inst, n));
  GPQ_TRY(arena.alloc(&d_block, nblocks));
  GPQ_TRY(arena.alloc(&d_count, (size_t)total));
  GPQ_TRY(arena.alloc(&d_hash, (size_t)total));
  std::vector<uint32_t> block_of(nblocks);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t b = tab[i].block_begin; b < tab[i].block_end; ++b) block_of[b] = i;
  GPQ_HIP(hipMemcpyAsync(d_inst, tab.data(), sizeof(PoolInst) * n, hipMemcpyHostToDevice, q->stream));
  GPQ_HIP(hipMemcpyAsync(d_block, block_of.data(), sizeof(uint32_t) * nblocks, hipMemcpyHostToDevice, q->stream));
  GPQ_HIP(hipMemsetAsync(d_count, 0, sizeof(uint32_t) * total, q->stream));
  GPQ_HIP(hipMemsetAsync(d_hash, 0, sizeof(uint64_t) * total, q->stream));
  for (uint32_t j = 0; j < max_init; ++j) {       // one initial event per LP per chunk, as one instance alone
    GPQ_TRY(launch_populate_instances(q.get(), d_inst, d_block, j));
    GPQ_TRY(q->post_init());
  }
  GPQ_TRY(q->calendar_start());
  GPQ_HIP(hipStreamSynchronize(q->stream));
  uint32_t flags = 0;
  GPQ_TRY(q->error_flags(&flags));
  if (flags) {
    set_error("phold batch (calendar): queue overflow while populating (flags " + std::to_string(flags) + ")");
    return GPQ_ERR_CAPACITY;
  }
  // `batch` steps captured once
  struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphGuard() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } gg;
  const uint32_t steps = batch ? batch : 16;
  GPQ_HIP(hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
  gpq_status st = GPQ_OK;
  for (uint32_t i = 0; i < steps && st == GPQ_OK; ++i) {
    st = launch_calendar_instances(q.get(), d_inst, d_block, d_count, d_hash);
    if (st == GPQ_OK) st = q->post();
  }
  const hipError_t e = hipStreamEndCapture(q->stream, &gg.g);
  GPQ_TRY(st);
  GPQ_HIP(e);
  GPQ_HIP(hipGraphInstantiate(&gg.x, gg.g, nullptr, nullptr, 0));
  struct Pinned {              // host staging of the counters and hashes
    void* p = nullptr;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
  } pc, ph;
  GPQ_HIP(hipHostMalloc(&pc.p, sizeof(uint32_t) * total, hipHostMallocDefault));
  GPQ_HIP(hipHostMalloc(&ph.p, sizeof(uint64_t) * total, hipHostMallocDefault));
  uint32_t* h_count = static_cast<uint32_t*>(pc.p);
  uint64_t* h_hash = static_cast<uint64_t*>(ph.p);
  const auto t0 = std::chrono::steady_clock::now();
  WindowState w{};
  for (;;) {
    GPQ_TRY(q->window(&w));
    if (w.done || w.error) break;
    GPQ_HIP(hipGraphLaunch(gg.x, q->stream));
  }
  GPQ_HIP(hipStreamSynchronize(q->stream));
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  GPQ_HIP(hipMemcpyAsync(h_count, d_count, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipMemcpyAsync(h_hash, d_hash, sizeof(uint64_t) * total, hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipStreamSynchronize(q->stream));
  fold_instances(tab, h_count, h_hash, hash.data(), events.data());
  if (w.error) {
    set_error("phold batch (calendar): device error flags " + std::to_string(w.error));
    return GPQ_ERR_CAPACITY;
  }
  if (seconds) *seconds = sec;
  return GPQ_OK;
}

gpq_status PholdBatch::results(uint64_t* hashes, uint64_t* ev) {
  for (size_t i = 0; i < inst.size(); ++i) {
    if (hashes) hashes[i] = hash[i];
    if (ev) ev[i] = events[i];
  }
  return GPQ_OK;
}

}  // namespace gpq

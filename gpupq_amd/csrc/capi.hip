// capi.hip -- the extern "C" boundary (include/gpupq.h): argument checks, handle
// lifetime, and translation of device error flags into status codes.
#include <cstring>
#include <string>

#include "phold.h"

namespace gpq {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

gpq_status QueueBase::init_common() {
  if (cfg.block_threads == 0) cfg.block_threads = 256;
  GPQ_CHECK_ARG(cfg.block_threads % 64 == 0 && cfg.block_threads <= 1024,
                "block_threads must be a multiple of 64 and <= 1024");
  threads = cfg.block_threads;
  GPQ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  GPQ_TRY(arena.alloc(&ws, 1));
  GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_ws), sizeof(WindowState), hipHostMallocDefault));
  std::memset(host_ws, 0, sizeof(WindowState));
  GPQ_HIP(hipMemcpyAsync(ws, host_ws, sizeof(WindowState), hipMemcpyHostToDevice, stream));
  if (cfg.collect_stats) {
    GPQ_TRY(arena.alloc(&stats, kStatSlots));
    GPQ_HIP(hipMemsetAsync(stats, 0, sizeof(unsigned long long) * kStatSlots, stream));
  }
  return GPQ_OK;
}

gpq_status QueueBase::stats_get(gpq_queue_stats* out) {
  std::memset(out, 0, sizeof(*out));
  WindowState w;
  GPQ_TRY(window(&w));
  out->post_launches = w.windows;     // every post publishes exactly one window
  if (!stats) return GPQ_OK;
  unsigned long long h[kStatSlots];
  GPQ_HIP(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  out->inserted = h[0];
  out->displaced = h[1];
  out->merged_lps = h[2];
  out->scanned_lps = h[3];
  out->aux_bytes = h[4];
  out->sorted = h[kStatSorted];
  out->merged = h[kStatMerged];
  out->selected = h[kStatSelected];
  return GPQ_OK;
}

__global__ void window_publish_kernel(const uint32_t* partial, uint32_t count, WindowState* ws, float lookahead,
                                      float horizon) {
  __shared__ uint32_t s_w[16];
  uint32_t m = kInfBits;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) m = partial[i] < m ? partial[i] : m;
  m = wave_min_u32(m);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t g = s_w[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) g = s_w[i] < g ? s_w[i] : g;
    float wend = bits_to_f(g) + lookahead;   // inf + lookahead stays inf
    if (horizon < wend) wend = horizon;
    ws->gmin = g;
    ws->wend = f_to_bits(wend);
    ws->done = (g == kInfBits || !(bits_to_f(g) < horizon)) ? 1u : 0u;
    ws->windows += 1;
  }
}

__global__ void sum_u32_kernel(const uint32_t* v, uint32_t count, unsigned long long* sum) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t x = i < count ? v[i] : 0u;
  unsigned long long w = x;
  for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, kWave);
  if (lane_id() == 0 && w) atomicAdd(sum, w);
}

gpq_status QueueBase::sum_u32(const uint32_t* d_values, uint32_t count, uint64_t* tmp, uint64_t* out) {
  GPQ_HIP(hipMemsetAsync(tmp, 0, sizeof(uint64_t), stream));
  hipLaunchKernelGGL(sum_u32_kernel, dim3(ceil_div(count, 256) ? ceil_div(count, 256) : 1), dim3(256), 0, stream,
                     d_values, count, reinterpret_cast<unsigned long long*>(tmp));
  GPQ_HIP(hipGetLastError());
  GPQ_HIP(hipMemcpyAsync(out, tmp, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  return GPQ_OK;
}

gpq_status QueueBase::publish_window() {
  if (profile_mark) GPQ_HIP(hipEventRecord(profile_mark, stream));
  hipLaunchKernelGGL(window_publish_kernel, dim3(1), dim3(1024), 0, stream, partial, partial_slots, ws,
                     cfg.lookahead, cfg.horizon);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status QueueBase::window(WindowState* out) {
  GPQ_HIP(hipMemcpyAsync(host_ws, ws, sizeof(WindowState), hipMemcpyDeviceToHost, stream));
  GPQ_HIP(hipStreamSynchronize(stream));
  *out = *host_ws;
  return GPQ_OK;
}

gpq_status QueueBase::error_flags(uint32_t* flags) {
  WindowState w;
  GPQ_TRY(window(&w));
  *flags = w.error;
  return GPQ_OK;
}

QueueBase* make_queue(const gpq_queue_config& cfg) {
  switch (cfg.kind) {
    case GPQ_LAQ: return new LaqQueue(cfg);
    case GPQ_LHQ: return new LhqQueue(cfg);
    case GPQ_LSQ: return new LsqQueue(cfg);
    case GPQ_GAQ: return new GaqQueue(cfg);
    case GPQ_HQ: return new HqQueue(cfg);
    default: return nullptr;
  }
}

}  // namespace gpq

using gpq::set_error;

extern "C" {

const char* gpq_status_string(gpq_status s) {
  switch (s) {
    case GPQ_OK: return "ok";
    case GPQ_ERR_INVALID: return "invalid argument";
    case GPQ_ERR_CAPACITY: return "capacity exceeded";
    case GPQ_ERR_HIP: return "HIP runtime error";
    case GPQ_ERR_NOMEM: return "out of device memory";
    case GPQ_ERR_STATE: return "invalid state";
    case GPQ_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char* gpq_last_error(void) { return gpq::last_error().c_str(); }

const char* gpq_version(void) { return "gpupq-amd 0.1 (gfx950)"; }

gpq_status gpq_queue_create(const gpq_queue_config* cfg, gpq_queue** out) {
  if (!cfg || !out) { set_error("null argument"); return GPQ_ERR_INVALID; }
  *out = nullptr;
  if (cfg->num_lps == 0) { set_error("num_lps must be > 0"); return GPQ_ERR_INVALID; }
  if (!(cfg->lookahead > 0.0f)) { set_error("lookahead must be > 0"); return GPQ_ERR_INVALID; }
  if (!(cfg->horizon > 0.0f)) { set_error("horizon must be > 0"); return GPQ_ERR_INVALID; }
  gpq::QueueBase* q = gpq::make_queue(*cfg);
  if (!q) { set_error("queue kind not provided"); return GPQ_ERR_UNSUPPORTED; }
  gpq_status s = q->init_common();
  if (s == GPQ_OK) s = q->init();
  if (s == GPQ_OK && hipStreamSynchronize(q->stream) != hipSuccess) {
    set_error("queue init failed on device");
    s = GPQ_ERR_HIP;
  }
  if (s != GPQ_OK) { delete q; return s; }
  *out = new gpq_queue{q};
  return GPQ_OK;
}

gpq_status gpq_queue_destroy(gpq_queue* q) {
  if (!q) return GPQ_OK;
  if (q->impl) (void)hipStreamSynchronize(q->impl->stream);
  delete q->impl;
  delete q;
  return GPQ_OK;
}

#define GPQ_Q(q) do { if (!(q) || !(q)->impl) { set_error("null queue"); return GPQ_ERR_INVALID; } } while (0)
#define GPQ_M(m) do { if (!(m) || !(m)->impl) { set_error("null model"); return GPQ_ERR_INVALID; } } while (0)

gpq_status gpq_queue_post_init(gpq_queue* q) { GPQ_Q(q); return q->impl->post_init(); }
gpq_status gpq_queue_pre(gpq_queue* q) { GPQ_Q(q); return q->impl->pre(); }
gpq_status gpq_queue_post(gpq_queue* q) { GPQ_Q(q); return q->impl->post(); }

gpq_status gpq_queue_get_min_ts(gpq_queue* q, float* out) {
  GPQ_Q(q);
  if (!out) { set_error("null out"); return GPQ_ERR_INVALID; }
  gpq::WindowState w;
  GPQ_TRY(q->impl->window(&w));
  std::memcpy(out, &w.gmin, sizeof(float));
  return GPQ_OK;
}

gpq_status gpq_queue_pending(gpq_queue* q, uint64_t* pending) {
  GPQ_Q(q);
  if (!pending) { set_error("null out"); return GPQ_ERR_INVALID; }
  return q->impl->pending(pending);
}

gpq_status gpq_queue_error_flags(gpq_queue* q, uint32_t* flags) {
  GPQ_Q(q);
  if (!flags) { set_error("null out"); return GPQ_ERR_INVALID; }
  return q->impl->error_flags(flags);
}

gpq_status gpq_queue_stats_get(gpq_queue* q, gpq_queue_stats* out) {
  GPQ_Q(q);
  if (!out) { set_error("null out"); return GPQ_ERR_INVALID; }
  return q->impl->stats_get(out);
}

gpq_status gpq_queue_sync(gpq_queue* q) {
  GPQ_Q(q);
  GPQ_HIP(hipStreamSynchronize(q->impl->stream));
  return GPQ_OK;
}

static gpq_status insert_host(gpq_queue* q, const gpq_event* items, uint32_t n, bool update) {
  GPQ_Q(q);
  if (n == 0) return GPQ_OK;
  if (!items) { set_error("null items"); return GPQ_ERR_INVALID; }
  for (uint32_t i = 0; i < n; ++i)
    if (items[i].dst >= q->impl->cfg.num_lps) { set_error("event dst out of range"); return GPQ_ERR_INVALID; }
  gpq::Event* d = nullptr;
  GPQ_HIP(hipMalloc(&d, sizeof(gpq::Event) * n));
  gpq_status s = GPQ_OK;
  if (hipMemcpyAsync(d, items, sizeof(gpq::Event) * n, hipMemcpyHostToDevice, q->impl->stream) != hipSuccess) {
    set_error("copy of items failed");
    s = GPQ_ERR_HIP;
  }
  if (s == GPQ_OK) s = gpq::launch_insert_items(q->impl, d, n, update);
  if (hipStreamSynchronize(q->impl->stream) != hipSuccess && s == GPQ_OK) s = GPQ_ERR_HIP;
  (void)hipFree(d);
  return s;
}

gpq_status gpq_queue_insert_host(gpq_queue* q, const gpq_event* items, uint32_t n) {
  return insert_host(q, items, n, false);
}

gpq_status gpq_queue_insert_or_update_host(gpq_queue* q, const gpq_event* items, uint32_t n) {
  return insert_host(q, items, n, true);
}

gpq_status gpq_queue_peek_host(gpq_queue* q, int set_done, gpq_event* items, int32_t* index, uint32_t max,
                               uint32_t* found) {
  GPQ_Q(q);
  if (!found || (max && (!items || !index))) { set_error("null out"); return GPQ_ERR_INVALID; }
  const uint32_t cap = max ? max : 1;
  gpq::Event* d_items = nullptr;
  int32_t* d_index = nullptr;
  uint32_t* d_found = nullptr;
  GPQ_HIP(hipMalloc(&d_items, sizeof(gpq::Event) * cap));
  GPQ_HIP(hipMalloc(&d_index, sizeof(int32_t) * cap));
  GPQ_HIP(hipMalloc(&d_found, sizeof(uint32_t)));
  hipStream_t st = q->impl->stream;
  gpq_status s = GPQ_OK;
  if (hipMemsetAsync(d_found, 0, sizeof(uint32_t), st) != hipSuccess) s = GPQ_ERR_HIP;
  if (s == GPQ_OK) s = gpq::launch_peek_items(q->impl, set_done, d_items, d_index, d_found, max);
  uint32_t h_found = 0;
  if (s == GPQ_OK && hipMemcpyAsync(&h_found, d_found, 4, hipMemcpyDeviceToHost, st) != hipSuccess) s = GPQ_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess && s == GPQ_OK) s = GPQ_ERR_HIP;
  const uint32_t ncopy = h_found < max ? h_found : max;
  if (s == GPQ_OK && ncopy) {
    if (hipMemcpy(items, d_items, sizeof(gpq::Event) * ncopy, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(index, d_index, sizeof(int32_t) * ncopy, hipMemcpyDeviceToHost) != hipSuccess)
      s = GPQ_ERR_HIP;
  }
  *found = h_found;
  (void)hipFree(d_items);
  (void)hipFree(d_index);
  (void)hipFree(d_found);
  if (s == GPQ_ERR_HIP) set_error("peek_host: HIP call failed");
  return s;
}

gpq_status gpq_queue_lp_contents(gpq_queue* q, uint32_t lp, gpq_event* out, uint32_t cap, uint32_t* n) {
  GPQ_Q(q);
  if (!n || (cap && !out)) { set_error("null out"); return GPQ_ERR_INVALID; }
  return q->impl->lp_contents(lp, reinterpret_cast<gpq::Event*>(out), cap, n);
}

gpq_status gpq_heap_node_contents(gpq_queue* q, uint32_t node, gpq_event* out, uint32_t cap, uint32_t* n) {
  GPQ_Q(q);
  if (q->impl->cfg.kind != GPQ_HQ) { set_error("not a heap queue"); return GPQ_ERR_UNSUPPORTED; }
  if (!n || (cap && !out)) { set_error("null out"); return GPQ_ERR_INVALID; }
  return static_cast<gpq::HqQueue*>(q->impl)->node_contents(node, reinterpret_cast<gpq::Event*>(out), cap, n);
}

gpq_status gpq_queue_length(gpq_queue* q, uint32_t lp, uint32_t* n) {
  GPQ_Q(q);
  if (!n) { set_error("null out"); return GPQ_ERR_INVALID; }
  return q->impl->length(lp, n);
}

gpq_status gpq_queue_clear_lp(gpq_queue* q, uint32_t lp) { GPQ_Q(q); return q->impl->clear_lp(lp); }

gpq_status gpq_queue_device_view(gpq_queue* q, gpq_queue_kind* kind, void* view, uint64_t view_bytes) {
  GPQ_Q(q);
  if (!view) { set_error("null view"); return GPQ_ERR_INVALID; }
  if (q->impl->cfg.kind == GPQ_LAQ) {
    const gpq::LaqView v = static_cast<gpq::LaqQueue*>(q->impl)->view();
    if (v.wlog) {
      set_error("device_view: a calendar-mode LAQ serves the library's PHOLD step only");
      return GPQ_ERR_STATE;
    }
    if (v.stage_alt) {
      set_error("device_view: the queue is driven by a PHOLD model that merges and commits in one kernel; "
                "its staging buffers alternate per step");
      return GPQ_ERR_STATE;
    }
  }
  return gpq::with_view(q->impl, [&](const auto& v) -> gpq_status {
    if (view_bytes != sizeof(v)) {
      set_error("device_view: view_bytes " + std::to_string(view_bytes) + " differs from the view size " +
                std::to_string(sizeof(v)) + " of the queue's kind");
      return GPQ_ERR_INVALID;
    }
    std::memcpy(view, &v, sizeof(v));
    if (kind) *kind = q->impl->cfg.kind;
    return GPQ_OK;
  });
}

gpq_status gpq_queue_stream(gpq_queue* q, void** stream) {
  GPQ_Q(q);
  if (!stream) { set_error("null out"); return GPQ_ERR_INVALID; }
  *stream = static_cast<void*>(q->impl->stream);
  return GPQ_OK;
}

gpq_status gpq_phold_create(gpq_queue* q, const gpq_phold_config* cfg, gpq_phold** out) {
  GPQ_Q(q);
  if (!cfg || !out) { set_error("null argument"); return GPQ_ERR_INVALID; }
  gpq::PholdModel* m = nullptr;
  GPQ_TRY(gpq::PholdModel::create(q->impl, *cfg, &m));
  *out = new gpq_phold{m};
  return GPQ_OK;
}

gpq_status gpq_phold_destroy(gpq_phold* m) {
  if (!m) return GPQ_OK;
  if (m->impl && m->impl->q) (void)hipStreamSynchronize(m->impl->q->stream);
  delete m->impl;
  delete m;
  return GPQ_OK;
}

gpq_status gpq_phold_init(gpq_phold* m) { GPQ_M(m); return m->impl->init(); }
gpq_status gpq_phold_handle_next(gpq_phold* m) { GPQ_M(m); return m->impl->launch_step(); }

gpq_status gpq_phold_get_events(gpq_phold* m, uint64_t* events) {
  GPQ_M(m);
  if (!events) { set_error("null out"); return GPQ_ERR_INVALID; }
  return m->impl->events(events);
}

gpq_status gpq_phold_finish(gpq_phold* m) {
  GPQ_M(m);
  GPQ_HIP(hipStreamSynchronize(m->impl->q->stream));
  return GPQ_OK;
}

gpq_status gpq_phold_lp_state(gpq_phold* m, uint32_t* counts, uint64_t* hashes) {
  GPQ_M(m);
  return m->impl->lp_state(counts, hashes);
}

gpq_status gpq_phold_trace_hash(gpq_phold* m, uint64_t* hash) {
  GPQ_M(m);
  if (!hash) { set_error("null out"); return GPQ_ERR_INVALID; }
  return m->impl->trace_hash(hash);
}

gpq_status gpq_des_profile(gpq_phold* m, uint32_t batch, gpq_kernel_profile* out) {
  GPQ_M(m);
  if (!out) { set_error("null out"); return GPQ_ERR_INVALID; }
  return m->impl->profile(batch, out);
}

gpq_status gpq_des_run(gpq_phold* m, gpq_drive_mode mode, uint32_t batch, gpq_des_report* report) {
  GPQ_M(m);
  if (mode != GPQ_DRIVE_PLAIN && mode != GPQ_DRIVE_GRAPH) { set_error("bad drive mode"); return GPQ_ERR_INVALID; }
  return m->impl->run(mode, batch, report);
}

}  // extern "C"

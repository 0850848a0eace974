// serve.hip -- serving a stream of mixed requests: PHOLD simulations (each its own per-LP
// queue and model) and A* grid queries (searched by the persistent batch kernel).
//
// Served drive (GPQ_DRIVE_GRAPH), continuous batching:
//   * admission in arrival order against an HBM budget: a PHOLD request is admitted when
//     the bytes of its queue and model fit beside the requests already running; A* queries
//     need no new memory (their grids and slots are allocated up front) and bypass a PHOLD
//     request that waits for memory;
//   * every admitted PHOLD instance advances on its own stream by graph replays of `batch`
//     steps; the host polls an event per replay and never blocks on one instance;
//   * admitted A* queries are gathered into launches of the persistent kernel on one of
//     `astar_lanes` streams, each lane owning a fixed range of workgroup slots; a lane takes
//     every query waiting when it becomes free.
// Plain drive (GPQ_DRIVE_PLAIN): one request at a time in arrival order, a host
// synchronization per step -- the reference's des_run pattern (src/des.cu:134-221).
//
// Results do not depend on the drive: a PHOLD request's trace hash and an A* query's cost
// are the same alone, in the mix, plain or served.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "serve.h"

namespace gpq {

namespace {

constexpr uint32_t kPholdCapacity = 64;    // records per LP ring (two-tier LAQ: below the near horizon)
constexpr uint32_t kPholdStage = 32;       // staged inserts per LP and step
constexpr uint32_t kPholdFar = 128;        // far-store records per LP

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

gpq_serve_config with_defaults(const gpq_serve_config& c) {
  gpq_serve_config s = c;
  if (!s.hbm_budget) s.hbm_budget = 4ull << 30;
  if (!s.phold_min_log2) s.phold_min_log2 = 12;
  if (!s.phold_max_log2) s.phold_max_log2 = 18;
  if (!s.phold_init_per_lp) s.phold_init_per_lp = 16;
  if (!(s.phold_end_time > 0.0f)) s.phold_end_time = 20.0f;
  if (!s.phold_batch) s.phold_batch = 16;
  if (!s.astar_lanes) s.astar_lanes = 4;
  return s;
}

gpq_queue_config phold_queue_config(uint32_t lps, float end_time) {
  gpq_queue_config q{};
  q.kind = GPQ_LAQ;
  q.num_lps = lps;
  q.capacity_per_lp = kPholdCapacity;
  q.stage_per_lp = kPholdStage;
  q.far_per_lp = kPholdFar;
  q.lookahead = 0.1f;
  q.horizon = end_time;
  q.block_threads = 256;
  return q;
}

gpq_phold_config phold_model_config(uint32_t lps, const gpq_serve_config& c, uint32_t seed) {
  gpq_phold_config p{};
  p.num_lps = lps;
  p.init_per_lp = c.phold_init_per_lp;
  p.seed = seed;
  p.lookahead = 0.1f;
  p.end_time = c.phold_end_time;
  p.p_self_u32 = 0x80000000u;
  p.checks = 0;
  p.drain = 1;
  return p;
}

// HBM of one PHOLD request: ring + staging + far records, per-LP queue words, model state.
uint64_t phold_bytes(uint32_t lps) {
  return (uint64_t)lps * ((kPholdCapacity + kPholdStage + kPholdFar) * 16ull + 40ull + 16ull) + (1u << 20);
}

struct Instance {
  std::unique_ptr<QueueBase> q;
  std::unique_ptr<PholdModel> m;
  hipEvent_t ev = nullptr;
  uint32_t req = 0;
  uint64_t bytes = 0;
  ~Instance() {
    m.reset();
    if (q) (void)hipStreamSynchronize(q->stream);
    if (ev) (void)hipEventDestroy(ev);
    q.reset();
  }
};

// Creates a PHOLD request's queue and model; populates it (model_init) when `populate` is set.
gpq_status start_phold(const gpq_serve_config& c, const gpq_request& r, uint32_t seed, uint32_t batch,
                       std::unique_ptr<Instance>* out, bool populate = true) {
  auto inst = std::make_unique<Instance>();
  const gpq_queue_config qc = phold_queue_config(r.size, c.phold_end_time);
  inst->q.reset(make_queue(qc));
  GPQ_CHECK_ARG(inst->q, "serve: queue kind not provided");
  GPQ_TRY(inst->q->init_common());
  GPQ_TRY(inst->q->init());
  PholdModel* m = nullptr;
  GPQ_TRY(PholdModel::create(inst->q.get(), phold_model_config(r.size, c, seed), &m));
  inst->m.reset(m);
  if (populate) GPQ_TRY(m->init());
  if (batch) {
    GPQ_TRY(m->prepare_graph(batch));
    GPQ_HIP(hipEventCreateWithFlags(&inst->ev, hipEventDisableTiming));
  }
  inst->bytes = inst->q->arena.bytes() + m->arena.bytes();
  *out = std::move(inst);
  return GPQ_OK;
}

gpq_status finish_phold(Instance& inst, gpq_request* r) {
  GPQ_HIP(hipStreamSynchronize(inst.q->stream));
  WindowState w;
  GPQ_TRY(inst.q->window(&w));
  if (w.error) {
    set_error("serve: PHOLD request overflowed its queue (flags " + std::to_string(w.error) + ")");
    return GPQ_ERR_CAPACITY;
  }
  GPQ_TRY(inst.m->trace_hash(&r->result));
  GPQ_TRY(inst.m->events(&r->work));
  r->bytes = inst.bytes;
  return GPQ_OK;
}

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  uint32_t* d_ids = nullptr;
  uint32_t* h_ids = nullptr;
  uint32_t* d_counter = nullptr;
  uint32_t slot_base = 0, slots = 0;
  std::vector<uint32_t> reqs;     // request indices of the launch in flight
  bool busy = false;
};

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p * (double)v.size());
  k = k ? k - 1 : 0;
  return v[std::min(k, v.size() - 1)];
}

}  // namespace

// Request i of the synthetic trace: Philox({i, 0, 6, 0}, {seed, 0}) gives the inter-arrival
// gap (Exp with mean 1/rate), the kind (word 1 < phold_fraction * 2^32 -> PHOLD), the PHOLD
// size 2^U[min, max] rounded to an integer, and the PHOLD seed (word 3).  A* request i is
// query i of the synthetic batched workload under the same master seed.
gpq_status serve_trace(const gpq_serve_config& cin, gpq_request* out, uint32_t* phold_seeds) {
  const gpq_serve_config c = with_defaults(cin);
  GPQ_CHECK_ARG(c.requests >= 1, "serve: no requests");
  GPQ_CHECK_ARG(c.phold_min_log2 <= c.phold_max_log2 && c.phold_max_log2 <= 24, "serve: bad PHOLD size range");
  GPQ_CHECK_ARG(c.phold_fraction >= 0.0 && c.phold_fraction <= 1.0, "serve: phold_fraction must be in [0, 1]");
  GPQ_CHECK_ARG(c.rate >= 0.0, "serve: rate must be >= 0");
  double t = 0.0;
  const double thr = c.phold_fraction * 4294967296.0;
  for (uint32_t i = 0; i < c.requests; ++i) {
    const uint4 r = philox4x32_10(make_uint4(i, 0u, 6u, 0u), make_uint2(c.seed, 0u));
    if (c.rate > 0.0) t += -std::log(((double)r.x + 0.5) * 0x1p-32) / c.rate * 1000.0;
    gpq_request q{};
    q.arrival_ms = t;
    q.kind = (double)r.y < thr ? 0u : 1u;
    if (q.kind == 0) {
      const double u = (double)r.z * 0x1p-32;
      const double e = c.phold_min_log2 + (c.phold_max_log2 - c.phold_min_log2) * u;
      q.size = (uint32_t)std::floor(std::exp2(e) + 0.5);
    } else {
      gpq_query_spec s;
      GPQ_TRY(query_synthetic(c.seed, i, &s));
      q.size = s.side;
    }
    out[i] = q;
    if (phold_seeds) phold_seeds[i] = r.w;
  }
  return GPQ_OK;
}

gpq_status serve_run(const gpq_serve_config& cin, gpq_drive_mode mode, gpq_serve_report* rep, gpq_request* out) {
  const gpq_serve_config c = with_defaults(cin);
  const uint32_t n = c.requests;
  std::vector<gpq_request> req(n);
  std::vector<uint32_t> seeds(n);
  GPQ_TRY(serve_trace(c, req.data(), seeds.data()));

  // A* queries: one batch object over all of them, grids generated before the clock starts.
  std::vector<uint32_t> astar_index(n, 0xFFFFFFFFu), astar_req;
  std::vector<gpq_query_spec> specs;
  for (uint32_t i = 0; i < n; ++i)
    if (req[i].kind == 1) {
      gpq_query_spec s;
      GPQ_TRY(query_synthetic(c.seed, i, &s));
      astar_index[i] = (uint32_t)specs.size();
      astar_req.push_back(i);
      specs.push_back(s);
    }
  // Every PHOLD request must fit the HBM budget on its own (the pool's per-LP layout, whole
  // 256-LP blocks); a request that cannot is rejected before anything runs.
  const uint64_t per_lp_bytes = PholdPool::bytes_per_lp(kPholdCapacity, kPholdStage, kPholdFar);
  for (uint32_t i = 0; i < n; ++i)
    if (req[i].kind == 0) {
      const uint64_t need = ((uint64_t)req[i].size + 255) / 256 * 256 * per_lp_bytes + (1ull << 20);
      if (need > c.hbm_budget) {
        set_error("serve: PHOLD request " + std::to_string(i) + " (" + std::to_string(req[i].size) +
                  " LPs) needs " + std::to_string(need) + " B of HBM, above the budget of " +
                  std::to_string(c.hbm_budget) + " B");
        return GPQ_ERR_CAPACITY;
      }
    }
  std::unique_ptr<AstarBatch> batch;
  const uint32_t lanes_n = mode == GPQ_DRIVE_GRAPH ? std::max<uint32_t>(1, c.astar_lanes) : 1;
  if (!specs.empty()) {
    gpq_batch_tuning t{};
    t.threads_per_query = 256;
    AstarBatch* b = nullptr;
    GPQ_TRY(AstarBatch::create(specs.data(), (uint32_t)specs.size(), t, &b));
    batch.reset(b);
  }
  const uint64_t astar_bytes_each =
      batch ? (uint64_t)batch->tuning.heap_capacity * batch->tuning.threads_per_query * 8ull : 0ull;

  // Lanes for A* launches (served) / one stream for the plain drive.
  DeviceArena arena;
  std::vector<Lane> lanes(lanes_n);
  struct LaneRelease {
    std::vector<Lane>& l;
    ~LaneRelease() {
      for (auto& x : l) {
        if (x.stream) (void)hipStreamSynchronize(x.stream);
        if (x.ev) (void)hipEventDestroy(x.ev);
        if (x.h_ids) (void)hipHostFree(x.h_ids);
        if (x.stream) (void)hipStreamDestroy(x.stream);
      }
    }
  } lane_release{lanes};
  if (batch) {
    const uint32_t R = batch->resident_slots();
    const uint32_t per = std::max<uint32_t>(1, R / lanes_n);
    for (uint32_t j = 0; j < lanes_n; ++j) {
      Lane& L = lanes[j];
      GPQ_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
      GPQ_HIP(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
      GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_ids), sizeof(uint32_t) * specs.size(),
                            hipHostMallocDefault));
      GPQ_TRY(arena.alloc(&L.d_ids, specs.size()));
      GPQ_TRY(arena.alloc(&L.d_counter, 2));
      L.slot_base = std::min(j * per, R - 1);
      L.slots = j + 1 == lanes_n ? R - L.slot_base : std::min(per, R - L.slot_base);
      if (L.slots == 0) L.slots = 1;
    }
  }

  // The PHOLD pool of the served drive: sized by the HBM budget, its graph captured, before
  // the clock starts (a server allocates it once).
  std::unique_ptr<PholdPool> pool;
  if (mode != GPQ_DRIVE_PLAIN) {
    uint32_t max_lps = 1;
    uint32_t phold_n = 0;
    for (const auto& r : req)
      if (r.kind == 0) { max_lps = std::max(max_lps, r.size); ++phold_n; }
    if (phold_n) {
      // as many whole 256-LP blocks as the budget holds beside 1 MiB of fixed tables (instance
      // table, windows, block maps, partial minima); every request fits (checked above)
      const uint64_t fixed = 1ull << 20;
      uint64_t cap = c.hbm_budget > fixed ? (c.hbm_budget - fixed) / per_lp_bytes / 256 * 256 : 0;
      cap = std::min<uint64_t>(cap, ((1ull << 31) / kPholdFar) - 256);
      if (cap < ((uint64_t)max_lps + 255) / 256 * 256) {
        set_error("serve: the PHOLD pool's LP limit is below the largest request");
        return GPQ_ERR_CAPACITY;
      }
      PholdPool* raw = nullptr;
      GPQ_TRY(PholdPool::create((uint32_t)cap, kPholdCapacity, kPholdStage, std::min<uint32_t>(phold_n, 1024), &raw,
                                kPholdFar));
      pool.reset(raw);
      GPQ_TRY(pool->prepare(c.phold_batch));
      GPQ_HIP(hipStreamSynchronize(pool->queue()->stream));
    }
  }

  // Plain drive with preallocation: every PHOLD request's queue and model exist before the
  // clock starts; inside it a request is populated, run and read back.
  std::vector<std::unique_ptr<Instance>> pre(mode == GPQ_DRIVE_PLAIN && c.plain_preallocated ? n : 0);
  uint64_t pre_bytes = 0;
  for (uint32_t i = 0; i < (uint32_t)pre.size(); ++i)
    if (req[i].kind == 0) {
      GPQ_TRY(start_phold(c, req[i], seeds[i], 0, &pre[i], false));
      pre_bytes += pre[i]->bytes;
    }
  if (!pre.empty()) {
    for (auto& p : pre)
      if (p) GPQ_HIP(hipStreamSynchronize(p->q->stream));
  }

  uint64_t inflight_bytes = 0, peak_bytes = 0, phold_bytes_now = 0, phold_peak = 0;
  uint32_t inflight = 0, peak_inflight = 0, completed = 0;
  const auto t0 = std::chrono::steady_clock::now();

  if (mode == GPQ_DRIVE_PLAIN) {
    uint32_t* h_done = nullptr;
    GPQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_done), sizeof(uint32_t), hipHostMallocDefault));
    std::unique_ptr<uint32_t, void (*)(uint32_t*)> done_guard(h_done, [](uint32_t* p) { (void)hipHostFree(p); });
    if (!pre.empty()) phold_peak = peak_bytes = pre_bytes;     // all of them are held throughout
    for (uint32_t i = 0; i < n; ++i) {
      while (now_ms(t0) < req[i].arrival_ms) std::this_thread::sleep_for(std::chrono::microseconds(50));
      req[i].start_ms = now_ms(t0);
      if (req[i].kind == 0) {
        std::unique_ptr<Instance> own;
        Instance* inst = nullptr;
        if (!pre.empty()) {
          inst = pre[i].get();
          GPQ_TRY(inst->m->init());
        } else {
          GPQ_TRY(start_phold(c, req[i], seeds[i], 0, &own));
          inst = own.get();
          peak_bytes = std::max(peak_bytes, inst->bytes);
          phold_peak = std::max(phold_peak, inst->bytes);
        }
        gpq_des_report dr;
        GPQ_TRY(inst->m->run(GPQ_DRIVE_PLAIN, 1, &dr));
        GPQ_TRY(finish_phold(*inst, &req[i]));
      } else {
        const uint32_t a = astar_index[i];
        GPQ_TRY(batch->plain_one(batch->device_ids() + a, a, 0, lanes[0].d_counter + 1, h_done, lanes[0].stream));
        req[i].bytes = astar_bytes_each;
      }
      req[i].finish_ms = now_ms(t0);
      ++completed;
    }
    peak_inflight = 1;
  } else {
    // PHOLD requests share one pool sized by the HBM budget (created before the clock starts);
    // a request is admitted when a free LP range of its size exists, between graph replays.
    std::vector<uint32_t> slot_req(pool ? 1024 : 0, 0xFFFFFFFFu);
    std::vector<uint32_t> fin;
    std::vector<uint64_t> fin_hash, fin_events;
    std::deque<uint32_t> waiting;          // arrived, not admitted
    std::vector<uint32_t> astar_ready;     // admitted A* queries not yet launched
    uint32_t next_arrival = 0;
    while (completed < n) {
      bool progress = false;
      const double now = now_ms(t0);
      while (next_arrival < n && req[next_arrival].arrival_ms <= now) waiting.push_back(next_arrival++);
      const bool pool_idle = pool && pool->replay_done();
      // PHOLD: retire finished instances once the pool's replay is complete
      if (pool_idle) {
        if (pool->error_flags_host()) {
          set_error("serve: PHOLD pool overflowed (flags " + std::to_string(pool->error_flags_host()) + ")");
          return GPQ_ERR_CAPACITY;
        }
        pool->finished(&fin);
        fin_hash.resize(fin.size());
        fin_events.resize(fin.size());
        GPQ_TRY(pool->retire_many(fin, fin_hash.data(), fin_events.data()));
        const double done_ms = now_ms(t0);
        for (size_t k = 0; k < fin.size(); ++k) {
          const uint32_t i = slot_req[fin[k]];
          req[i].result = fin_hash[k];
          req[i].work = fin_events[k];
          req[i].finish_ms = done_ms;
          inflight_bytes -= req[i].bytes;
          phold_bytes_now -= req[i].bytes;
          --inflight;
          ++completed;
          progress = true;
        }
      }
      // admission in arrival order; A* bypasses a PHOLD request that waits for LPs
      bool phold_blocked = false;
      for (auto it = waiting.begin(); it != waiting.end();) {
        const uint32_t i = *it;
        if (req[i].kind == 1) {
          req[i].start_ms = now_ms(t0);
          req[i].bytes = astar_bytes_each;
          astar_ready.push_back(i);
          it = waiting.erase(it);
          progress = true;
          continue;
        }
        if (phold_blocked || !pool_idle) {
          ++it;
          continue;
        }
        gpq_phold_instance pi{req[i].size, c.phold_init_per_lp, seeds[i], c.phold_end_time, 0.1f, 0x80000000u};
        uint32_t slot = 0;
        bool ok = false;
        GPQ_TRY(pool->admit(pi, &slot, &ok));
        if (!ok) {
          phold_blocked = true;
          ++it;
          continue;
        }
        slot_req[slot] = i;
        req[i].start_ms = now_ms(t0);
        req[i].bytes = ((uint64_t)req[i].size + 255) / 256 * 256 * per_lp_bytes;
        inflight_bytes += req[i].bytes;
        phold_bytes_now += req[i].bytes;
        phold_peak = std::max(phold_peak, phold_bytes_now);
        ++inflight;
        it = waiting.erase(it);
        progress = true;
      }
      if (pool_idle && pool->any_active()) {
        GPQ_TRY(pool->launch_async());
        progress = true;
      }
      // A* launches on free lanes
      for (auto& L : lanes) {
        if (astar_ready.empty()) break;
        if (L.busy) continue;
        const uint32_t k = (uint32_t)astar_ready.size();
        for (uint32_t j = 0; j < k; ++j) L.h_ids[j] = astar_index[astar_ready[j]];
        GPQ_HIP(hipMemcpyAsync(L.d_ids, L.h_ids, sizeof(uint32_t) * k, hipMemcpyHostToDevice, L.stream));
        GPQ_TRY(batch->launch(L.d_ids, k, L.slot_base, L.slots, L.d_counter, L.stream));
        GPQ_HIP(hipEventRecord(L.ev, L.stream));
        L.reqs.assign(astar_ready.begin(), astar_ready.end());
        L.busy = true;
        inflight += k;
        inflight_bytes += astar_bytes_each * k;
        astar_ready.clear();
        progress = true;
      }
      peak_inflight = std::max(peak_inflight, inflight);
      peak_bytes = std::max(peak_bytes, inflight_bytes);
      // A* lanes
      for (auto& L : lanes) {
        if (!L.busy) continue;
        const hipError_t e = hipEventQuery(L.ev);
        if (e == hipErrorNotReady) continue;
        GPQ_HIP(e);
        const double t = now_ms(t0);
        for (uint32_t i : L.reqs) req[i].finish_ms = t;
        completed += (uint32_t)L.reqs.size();
        inflight -= (uint32_t)L.reqs.size();
        inflight_bytes -= astar_bytes_each * L.reqs.size();
        L.reqs.clear();
        L.busy = false;
        progress = true;
      }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  const double end_ms = now_ms(t0);

  // A* results (cost, expansions) from the batch object.
  if (batch) {
    std::vector<int64_t> costs(specs.size());
    std::vector<uint64_t> expanded(specs.size());
    GPQ_TRY(batch->results(costs.data(), expanded.data()));
    for (size_t a = 0; a < specs.size(); ++a) {
      gpq_request& r = req[astar_req[a]];
      r.result = costs[a] < 0 ? ~0ull : (uint64_t)costs[a];
      r.work = expanded[a];
    }
  }
  std::vector<double> lat, lat_p, lat_a;
  double sum = 0.0, first = 1e300;
  for (const auto& r : req) {
    const double l = r.finish_ms - r.arrival_ms;
    lat.push_back(l);
    (r.kind == 0 ? lat_p : lat_a).push_back(l);
    sum += l;
    first = std::min(first, r.arrival_ms);
  }
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->completed = completed;
    rep->seconds = (end_ms - first) / 1000.0;
    rep->requests_per_second = rep->seconds > 0 ? completed / rep->seconds : 0.0;
    rep->p50_ms = percentile(lat, 0.50);
    rep->p99_ms = percentile(lat, 0.99);
    rep->mean_ms = sum / n;
    rep->phold_p99_ms = percentile(lat_p, 0.99);
    rep->astar_p99_ms = percentile(lat_a, 0.99);
    rep->peak_bytes = peak_bytes;
    rep->phold_peak_bytes = phold_peak;
    rep->pool_bytes = pool ? pool->device_bytes() : 0;
    rep->peak_inflight = peak_inflight;
  }
  if (out) std::copy(req.begin(), req.end(), out);
  return GPQ_OK;
}

}  // namespace gpq

// astar.hip -- A* on grids (src/a_star.cu), exact integer costs.
//
// Single query (gpq_astar_*): the LPs of a per-LP queue are the open lists.  A step pops
// the best entry of every list, drops entries made stale by a cheaper path (g != g[cell]),
// expands the rest and relaxes neighbours with atomicMin on the g array; improved
// neighbours go to list hash(cell) through the queue's staged insert.  queue_post folds
// the inserts in and reduces the smallest f over all lists; the search ends when the best
// goal cost is <= that minimum (or every list is empty).  The reference pops up to 8
// entries per agent and step with thread 0, keeps f in fp16 and recomputes g from it
// (src/a_star.cu:325-374,470-474), so it does not guarantee the optimum.
//
// Batched (gpq_astar_batch_*): one workgroup per query, pulled from a device counter as
// get_new_agent does (src/a_star.cu:149-164), K = blockDim per-thread binary heaps of
// 64-bit keys (f << 32 | cell) in HBM, successors routed through LDS staging.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "astar.h"

namespace gpq {

// ---------------------------------------------------------------------------------------
// shared grid helpers
// ---------------------------------------------------------------------------------------
// Obstacle bitmap: cell c blocked iff word (c & 3) of Philox({c >> 2, query, tag, 2}, {seed, 0})
// is below threshold.  One thread per 32 cells (one bitmap word).
__global__ void grid_generate_kernel(uint32_t* bits, uint64_t cells, uint32_t seed, uint32_t tag, uint32_t query,
                                     uint32_t threshold, uint64_t force_a, uint64_t force_b) {
  const uint64_t words = (cells + 31) >> 5;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t word = 0;
#pragma unroll
    for (int q4 = 0; q4 < 8; ++q4) {
      const uint64_t base = (w << 5) + 4u * q4;
      const uint4 r = philox4x32_10(make_uint4((uint32_t)(base >> 2), query, tag, 2u), make_uint2(seed, 0u));
      const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t c = base + j;
        const bool blocked = c < cells && rr[j] < threshold && c != force_a && c != force_b;
        word |= (blocked ? 1u : 0u) << (4 * q4 + j);
      }
    }
    bits[w] = word;
  }
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void grid_from_bytes_kernel(uint32_t* bits, const uint8_t* bytes, uint64_t cells) {
  const uint64_t words = (cells + 31) >> 5;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t word = 0;
    for (int j = 0; j < 32; ++j) {
      const uint64_t c = (w << 5) + j;
      if (c < cells && bytes[c]) word |= 1u << j;
    }
    bits[w] = word;
  }
}

__global__ void grid_to_bytes_kernel(const uint32_t* bits, uint8_t* bytes, uint64_t cells) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < cells; c += (uint64_t)gridDim.x * blockDim.x)
    bytes[c] = bit_get(bits, c) ? 1 : 0;
}

namespace {

uint32_t grid_blocks(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(g, 1), 65536);
}

// ---------------------------------------------------------------------------------------
// single query over a per-LP queue
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t list_of(uint32_t cell, uint32_t lists) {
  return (uint32_t)(((uint64_t)(cell * 0x9E3779B1u) * lists) >> 32);
}

template <class Q>
__global__ __launch_bounds__(256) void astar_step_kernel(Q q, AstarDev a) {
  uint32_t expanded = 0, newly_closed = 0;
  const uint32_t best = a.st->best;
  const uint32_t fmin = a.ws->gmin;
  const uint32_t fmax = fmin > 0xFFFFFFFFu - a.window ? 0xFFFFFFFFu : fmin + a.window;
  for (uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x; lp < a.lists; lp += gridDim.x * blockDim.x) {
    Event* ev;
    const int idx = q.peek(&ev, (int)lp);
    if (idx < 0) continue;
    const Event e = load_event(ev);
    if (e.ts > fmax) continue;                      // outside the f window: stays in its list
    q.set_done(idx);
    const uint32_t cell = e.seq, f = e.ts, hv = e.src;
    const uint32_t g = f - hv;
    if (g != __hip_atomic_load(&a.g[cell], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;   // stale
    if (cell == a.target) {
      atomicMin(&a.st->best, g);
      continue;
    }
    if (best != 0xFFFFFFFFu && f >= best) continue;   // cannot improve on the goal found
    ++expanded;
    const uint32_t bit = 1u << (cell & 31);
    if (!(atomicOr(&a.closed[cell >> 5], bit) & bit)) ++newly_closed;
    const int x = (int)(cell % a.width), y = (int)(cell / a.width);
    for (uint32_t k = 0; k < a.conn; ++k) {
      const int nx = x + nbr_dx(k), ny = y + nbr_dy(k);
      if (nx < 0 || ny < 0 || nx >= (int)a.width || ny >= (int)a.height) continue;
      const uint32_t v = (uint32_t)ny * a.width + (uint32_t)nx;
      if (bit_get(a.blocked, v)) continue;
      const uint32_t g2 = g + nbr_cost(a.conn, k);
      if (atomicMin(&a.g[v], g2) <= g2) continue;
      const uint32_t h2 = grid_heur(a.conn, nx, ny, (int)a.tx, (int)a.ty);
      q.insert(Event{g2 + h2, list_of(v, a.lists), h2, v});
    }
  }
  const uint32_t se = wave_sum_u32(expanded), sc = wave_sum_u32(newly_closed);
  if (lane_id() == 0) {
    if (se) atomicAdd(&a.st->expanded, (unsigned long long)se);
    if (sc) atomicAdd(&a.st->closed, (unsigned long long)sc);
  }
}

// Done when every list is empty or the best goal cost is <= the smallest f of any list.
__global__ void astar_check_kernel(AstarState* st, const WindowState* ws) {
  const uint32_t fmin = ws->gmin;
  st->done = (fmin == kInfBits || (st->best != 0xFFFFFFFFu && st->best <= fmin)) ? 1u : 0u;
  st->steps += 1;
}

template <class Q>
__global__ void astar_seed_kernel(Q q, AstarDev a, uint32_t source) {
  if (threadIdx.x || blockIdx.x) return;
  a.g[source] = 0;
  const uint32_t h0 = grid_heur(a.conn, (int)(source % a.width), (int)(source / a.width), (int)a.tx, (int)a.ty);
  q.insert(Event{h0, list_of(source, a.lists), h0, source});
  *a.st = AstarState{0xFFFFFFFFu, 0u, 0ull, 0ull, 0ull};
}

}  // namespace

gpq_status AstarModel::create(QueueBase* queue, const gpq_astar_config& c, AstarModel** out) {
  GPQ_CHECK_ARG(queue, "null queue");
  GPQ_CHECK_ARG(queue->cfg.kind == GPQ_LAQ || queue->cfg.kind == GPQ_LHQ || queue->cfg.kind == GPQ_LSQ,
                "A*: the open lists must be a per-LP queue (LAQ, LHQ or LSQ)");
  GPQ_CHECK_ARG(std::isinf(queue->cfg.lookahead) && std::isinf(queue->cfg.horizon),
                "A*: the open-list queue needs lookahead = horizon = +inf (peek = pop the minimum)");
  GPQ_CHECK_ARG(c.connectivity == 4 || c.connectivity == 8, "A*: connectivity must be 4 or 8");
  GPQ_CHECK_ARG(c.width >= 1 && c.height >= 1 && (uint64_t)c.width * c.height < (1ull << 31), "A*: bad grid size");
  GPQ_CHECK_ARG(c.sx < c.width && c.tx < c.width && c.sy < c.height && c.ty < c.height, "A*: endpoint outside grid");
  AstarModel* m = new AstarModel();
  m->cfg = c;
  m->q = queue;
  const uint64_t cells = (uint64_t)c.width * c.height;
  gpq_status s;
  if ((s = m->arena.alloc(&m->blocked, (cells + 31) / 32)) != GPQ_OK || (s = m->arena.alloc(&m->g, cells)) != GPQ_OK ||
      (s = m->arena.alloc(&m->closed, (cells + 31) / 32)) != GPQ_OK || (s = m->arena.alloc(&m->st, 1)) != GPQ_OK) {
    delete m;
    return s;
  }
  *out = m;
  return GPQ_OK;
}

AstarModel::~AstarModel() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
}

gpq_status AstarModel::generate() {
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  const uint64_t s = (uint64_t)cfg.sy * cfg.width + cfg.sx, t = (uint64_t)cfg.ty * cfg.width + cfg.tx;
  hipLaunchKernelGGL(grid_generate_kernel, dim3(grid_blocks((cells + 31) / 32)), dim3(256), 0, q->stream, blocked,
                     cells, cfg.seed, cfg.stream_tag, cfg.query, cfg.threshold, s, t);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

AstarDev AstarModel::view() const {
  return AstarDev{blocked, g, closed, st, q->ws, cfg.width, cfg.height, cfg.connectivity, cfg.tx, cfg.ty,
                  cfg.ty * cfg.width + cfg.tx, q->cfg.num_lps, cfg.window};
}

// model_init: fresh g / closed arrays, the source in its open list.
gpq_status AstarModel::restart() {
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(cells)), dim3(256), 0, q->stream, g, cells, 0xFFFFFFFFu);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks((cells + 31) / 32)), dim3(256), 0, q->stream, closed,
                     (cells + 31) / 32, 0u);
  GPQ_HIP(hipGetLastError());
  const AstarDev d = view();
  const uint32_t source = cfg.sy * cfg.width + cfg.sx;
  GPQ_TRY(with_view(q, [&](auto qv) -> gpq_status {
    hipLaunchKernelGGL(astar_seed_kernel<decltype(qv)>, dim3(1), dim3(64), 0, q->stream, qv, d, source);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }));
  GPQ_TRY(q->post_init());
  GPQ_HIP(hipStreamSynchronize(q->stream));
  return GPQ_OK;
}

gpq_status AstarModel::init() {
  GPQ_TRY(generate());
  return restart();
}

gpq_status AstarModel::set_grid(const uint8_t* bytes) {
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  uint8_t* d = nullptr;
  GPQ_HIP(hipMalloc(&d, cells));
  hipError_t e = hipMemcpyAsync(d, bytes, cells, hipMemcpyHostToDevice, q->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(grid_from_bytes_kernel, dim3(grid_blocks((cells + 31) / 32)), dim3(256), 0, q->stream, blocked,
                       d, cells);
    e = hipStreamSynchronize(q->stream);
  }
  (void)hipFree(d);
  GPQ_HIP(e);
  return restart();
}

gpq_status AstarModel::grid_bytes(uint8_t* out) {
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  uint8_t* d = nullptr;
  GPQ_HIP(hipMalloc(&d, cells));
  hipLaunchKernelGGL(grid_to_bytes_kernel, dim3(grid_blocks(cells)), dim3(256), 0, q->stream, blocked, d, cells);
  hipError_t e = hipMemcpyAsync(out, d, cells, hipMemcpyDeviceToHost, q->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(q->stream);
  (void)hipFree(d);
  GPQ_HIP(e);
  return GPQ_OK;
}

gpq_status AstarModel::step() {
  const AstarDev d = view();
  GPQ_TRY(q->pre());
  GPQ_TRY(with_view(q, [&](auto qv) -> gpq_status {
    const uint32_t lists = q->cfg.num_lps;
    hipLaunchKernelGGL(astar_step_kernel<decltype(qv)>, dim3((lists + 255) / 256), dim3(256), 0, q->stream, qv, d);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }));
  GPQ_TRY(q->post());
  hipLaunchKernelGGL(astar_check_kernel, dim3(1), dim3(1), 0, q->stream, st, q->ws);
  GPQ_HIP(hipGetLastError());
  return GPQ_OK;
}

gpq_status AstarModel::poll(AstarState* out) {
  GPQ_HIP(hipMemcpyAsync(&host_st, st, sizeof(AstarState), hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipStreamSynchronize(q->stream));
  *out = host_st;
  return GPQ_OK;
}

gpq_status AstarModel::run(gpq_drive_mode mode, uint32_t batch, gpq_astar_report* rep) {
  GPQ_CHECK_ARG(batch >= 1, "batch must be >= 1");
  GPQ_HIP(hipStreamSynchronize(q->stream));
  AstarState s{};
  uint64_t launched = 0;
  // Every step pops at least one entry and an entry is pushed only when a g value drops, so
  // a search that runs past this bound is a bug, reported instead of spinning.
  const uint64_t step_cap = 64ull * cfg.width * cfg.height + 1024;
  const auto t0 = std::chrono::steady_clock::now();
  if (mode == GPQ_DRIVE_PLAIN) {
    for (;;) {
      GPQ_TRY(poll(&s));
      WindowState w;
      GPQ_TRY(q->window(&w));
      if (s.done || w.error) break;
      if (launched > step_cap) { set_error("A*: step bound exceeded"); return GPQ_ERR_STATE; }
      GPQ_TRY(step());
      ++launched;
    }
  } else {
    if (!graph_exec || graph_batch != batch) {
      if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
      if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
      GPQ_HIP(hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
      gpq_status st2 = GPQ_OK;
      for (uint32_t i = 0; i < batch && st2 == GPQ_OK; ++i) st2 = step();
      hipGraph_t gph = nullptr;
      const hipError_t e = hipStreamEndCapture(q->stream, &gph);
      if (st2 != GPQ_OK) return st2;
      GPQ_HIP(e);
      graph = gph;
      GPQ_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
      graph_batch = batch;
    }
    for (;;) {
      GPQ_TRY(poll(&s));
      WindowState w;
      GPQ_TRY(q->window(&w));
      if (s.done || w.error) break;
      if (launched > step_cap) { set_error("A*: step bound exceeded"); return GPQ_ERR_STATE; }
      GPQ_HIP(hipGraphLaunch(graph_exec, q->stream));
      launched += batch;
    }
  }
  GPQ_HIP(hipStreamSynchronize(q->stream));
  const auto t1 = std::chrono::steady_clock::now();
  WindowState w;
  GPQ_TRY(q->window(&w));
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->cost = s.best == 0xFFFFFFFFu ? -1 : (int64_t)s.best;
    rep->expanded = s.expanded;
    rep->closed = s.closed;
    rep->steps = launched;
    rep->seconds = std::chrono::duration<double>(t1 - t0).count();
    rep->error_flags = w.error;
  }
  if (w.error) {
    set_error("A*: device error flags " + std::to_string(w.error) + " (open-list capacity or staging)");
    return GPQ_ERR_CAPACITY;
  }
  return GPQ_OK;
}

// Path of the last search (src/a_star.cu:90-94,417-438 backtrack the closed list): from the
// target, repeatedly step to a free neighbour whose g plus the edge cost equals the cell's g;
// the result runs source -> target and its edge costs sum to the optimal cost.
gpq_status AstarModel::path(uint32_t* cells, uint32_t cap, uint32_t* len) {
  const uint64_t ncell = (uint64_t)cfg.width * cfg.height;
  std::vector<uint32_t> gh(ncell), bits((ncell + 31) / 32);
  GPQ_HIP(hipMemcpyAsync(gh.data(), g, sizeof(uint32_t) * ncell, hipMemcpyDeviceToHost, q->stream));
  GPQ_HIP(hipMemcpyAsync(bits.data(), blocked, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, q->stream));
  AstarState s;
  GPQ_TRY(poll(&s));
  *len = 0;
  if (s.best == 0xFFFFFFFFu) return GPQ_OK;          // unreachable: empty path
  const uint32_t src = cfg.sy * cfg.width + cfg.sx;
  uint32_t cur = cfg.ty * cfg.width + cfg.tx;
  uint32_t gcur = s.best;
  std::vector<uint32_t> rev{cur};
  static const int dx[8] = {1, -1, 0, 0, 1, -1, 1, -1}, dy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
  while (cur != src) {
    const int x = (int)(cur % cfg.width), y = (int)(cur / cfg.width);
    bool stepped = false;
    for (uint32_t k = 0; k < cfg.connectivity && !stepped; ++k) {
      const int nx = x + dx[k], ny = y + dy[k];
      if (nx < 0 || ny < 0 || nx >= (int)cfg.width || ny >= (int)cfg.height) continue;
      const uint32_t v = (uint32_t)ny * cfg.width + (uint32_t)nx;
      if ((bits[v >> 5] >> (v & 31)) & 1u) continue;
      const uint32_t w = cfg.connectivity == 4 ? 1u : (k < 4 ? 10u : 14u);
      if (gh[v] != 0xFFFFFFFFu && gh[v] + w == gcur) {
        cur = v;
        gcur = gh[v];
        rev.push_back(v);
        stepped = true;
      }
    }
    if (!stepped || rev.size() > ncell) {
      set_error("A*: no predecessor chain from the target (search not run to the end?)");
      return GPQ_ERR_STATE;
    }
  }
  *len = (uint32_t)rev.size();
  for (uint32_t i = 0; i < *len && i < cap; ++i) cells[i] = rev[rev.size() - 1 - i];
  return GPQ_OK;
}

}  // namespace gpq

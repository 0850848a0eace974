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
      const uint32_t list = list_of(v, a.lists);
      q.insert_or_update(Event{g2 + h2, list, h2, v}, (int)list);   // decrease-key on a LAQ in that mode
    }
  }
  const uint32_t se = wave_sum_u32(expanded), sc = wave_sum_u32(newly_closed);
  if (lane_id() == 0) {
    if (se) atomicAdd(&a.st->expanded, (unsigned long long)se);
    if (sc) atomicAdd(&a.st->closed, (unsigned long long)sc);
  }
}

// ---------------------------------------------------------------------------------------
// tile search: one workgroup per tile x tile square of the grid
// ---------------------------------------------------------------------------------------
constexpr int kTileBlock = 256;
constexpr int kTileGrid = 1024;   // workgroups of a tile step: four per CU, all resident

// Appends (cell, g) to a tile's inbox of the given parity (a flag when full); the entry that
// opens the inbox also lists the tile among that parity's active tiles.
__device__ __forceinline__ void inbox_push(const AstarTiles& t, uint32_t par, uint32_t tile, uint32_t cell,
                                           uint32_t g) {
  const size_t slot = (size_t)par * t.tiles + tile;
  const uint32_t pos = atomicAdd(&t.count[slot], 1u);
  if (pos < t.cap) t.inbox[slot * t.cap + pos] = make_uint2(cell, g);
  else atomicOr(t.error, kErrQueueOverflow);
  if (pos == 0) t.active[(size_t)par * t.tiles + atomicAdd(&t.n_active[par], 1u)] = tile;
}

// One step of the tile search.  The workgroups walk the list of tiles whose inbox received
// entries in the last step (kTileGrid workgroups, all resident).  For each such tile: load the
// tile's g values and obstacle bits (and those of the cells around it) into LDS, keep the
// inbox entries inside the window (f <= fmax) as the frontier and hold the others back, then
// expand the frontier layer after layer (at most 8 * TILE layers) in LDS only: a neighbour
// inside the tile is relaxed in LDS and joins the next layer if it stays inside the window; a
// neighbour outside keeps its best g in an LDS ring.  When the tile is done, its held-back
// cells go to its own next inbox, the ring's cells are relaxed in HBM (atomicMin) and the
// improved ones sent to their tiles' inboxes, and the tile's lowered g values are folded back
// into HBM.  g only decreases, an entry whose g is no longer its cell's is dropped, and the
// search ends only when the best goal cost is <= the smallest f still pending, so the cost is
// exact.
// The ring of cells around a TILE x TILE tile: hx, hy in [0, TILE + 1] relative to the cell
// left of and above the tile's corner, not both inside.  Top row, bottom row (both with the
// corners), then the left and right columns.
template <int TILE>
__device__ __forceinline__ uint32_t halo_index(uint32_t hx, uint32_t hy) {
  if (hy == 0) return hx;
  if (hy == TILE + 1) return (TILE + 2) + hx;
  return 2 * (TILE + 2) + (hx == 0 ? 0u : (uint32_t)TILE) + (hy - 1);
}
template <int TILE>
__device__ __forceinline__ void halo_cell(uint32_t i, uint32_t* hx, uint32_t* hy) {
  if (i < TILE + 2) { *hx = i; *hy = 0; return; }
  if (i < 2 * (TILE + 2)) { *hx = i - (TILE + 2); *hy = TILE + 1; return; }
  i -= 2 * (TILE + 2);
  *hx = i < TILE ? 0u : (uint32_t)TILE + 1;
  *hy = (i < TILE ? i : i - TILE) + 1;
}

template <int TILE>
__global__ __launch_bounds__(kTileBlock) void astar_tile_kernel(AstarDev a, AstarTiles t) {
  constexpr uint32_t kCells = TILE * TILE, kCap = 2 * kCells, kHalo = 4 * TILE + 4;
  __shared__ uint32_t s_g[kCells];
  __shared__ uint32_t s_chg[kCells / 32];         // g lowered in this step
  __shared__ uint32_t s_def[kCells / 32];         // to be sent to the tile's next inbox
  __shared__ uint32_t s_exp[kCells / 32];         // expanded in this step (folded into the closed set)
  __shared__ uint32_t s_blk[kCells / 32];         // obstacle bits of the tile (past the grid's edge: set)
  __shared__ uint32_t s_out[kHalo];               // best g found for each cell around the tile
  __shared__ uint32_t s_hblk[(kHalo + 31) / 32];  // obstacle bits around the tile (outside the grid: set)
  __shared__ uint2 s_fr[2][kCap];                 // frontier layers: (local cell, g)
  __shared__ uint32_t s_n[2];
  __shared__ uint32_t s_red[kTileBlock / kWave][2];
  AstarState* st = a.st;
  if (st->done) return;
  const uint32_t par = st->parity, nxt = 1u - par;
  const uint32_t fmax = st->fmax, best = st->best;
  const uint32_t n_tiles = t.n_active[par];
  uint32_t fmin_pushed = 0xFFFFFFFFu;
  unsigned long long expanded = 0;
  for (uint32_t j = blockIdx.x; j < n_tiles; j += gridDim.x) {
    __syncthreads();                              // the previous tile's LDS is no longer read
    const uint32_t tile = t.active[(size_t)par * t.tiles + j];
    const uint32_t n_in = min(t.count[(size_t)par * t.tiles + tile], t.cap);
    const uint32_t x0 = (tile % t.tiles_x) * TILE, y0 = (tile / t.tiles_x) * TILE;
    for (uint32_t i = threadIdx.x; i < kCells; i += kTileBlock) {
      const uint32_t x = x0 + i % TILE, y = y0 + i / TILE;
      s_g[i] = (x < a.width && y < a.height) ? a.g[(size_t)y * a.width + x] : 0u;
    }
    for (uint32_t i = threadIdx.x; i < kCells / 32; i += kTileBlock) {
      s_chg[i] = 0;
      s_def[i] = 0;
      s_exp[i] = 0;
      s_blk[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < kHalo; i += kTileBlock) s_out[i] = 0xFFFFFFFFu;
    for (uint32_t i = threadIdx.x; i < (kHalo + 31) / 32; i += kTileBlock) s_hblk[i] = 0;
    if (threadIdx.x == 0) {
      s_n[0] = 0;
      s_n[1] = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kCells; i += kTileBlock) {   // cells past the grid's edge: set
      const uint32_t x = x0 + i % TILE, y = y0 + i / TILE;
      if (x >= a.width || y >= a.height || bit_get(a.blocked, (uint64_t)y * a.width + x))
        atomicOr(&s_blk[i >> 5], 1u << (i & 31));
    }
    for (uint32_t i = threadIdx.x; i < kHalo; i += kTileBlock) {
      uint32_t hx, hy;
      halo_cell<TILE>(i, &hx, &hy);
      const int nx = (int)(x0 + hx) - 1, ny = (int)(y0 + hy) - 1;
      if (nx < 0 || ny < 0 || nx >= (int)a.width || ny >= (int)a.height ||
          bit_get(a.blocked, (uint64_t)ny * a.width + (uint32_t)nx))
        atomicOr(&s_hblk[i >> 5], 1u << (i & 31));
    }
    // the inbox, and the parked entries once the window reaches them: entries inside the
    // window start the frontier, the others are parked again
    const uint2* in = t.inbox + ((size_t)par * t.tiles + tile) * t.cap;
    const uint32_t n_park = t.park_fmin[tile] <= fmax ? min(t.park_count[tile], t.cap) : 0u;
    const uint2* pk = t.park + (size_t)tile * t.cap;
    for (uint32_t i = threadIdx.x; i < n_in + n_park; i += kTileBlock) {
      const uint2 e = i < n_in ? in[i] : pk[i - n_in];
      const uint32_t x = e.x % a.width, y = e.x / a.width;
      if (x - x0 >= (uint32_t)TILE || y - y0 >= (uint32_t)TILE) {   // not this tile's cell: never sent
        atomicOr(t.error, kErrBadState);
        continue;
      }
      const uint32_t li = (y - y0) * TILE + (x - x0);
      if (e.y != s_g[li]) continue;              // superseded
      const uint32_t f = e.y + grid_heur(a.conn, (int)x, (int)y, (int)a.tx, (int)a.ty);
      if (best != 0xFFFFFFFFu && f >= best) continue;
      const uint32_t pos = f <= fmax ? atomicAdd(&s_n[0], 1u) : kCap;
      if (pos < kCap) s_fr[0][pos] = make_uint2(li, e.y);
      else atomicOr(&s_def[li >> 5], 1u << (li & 31));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      t.count[(size_t)par * t.tiles + tile] = 0;  // consumed; refilled two steps on
      if (n_park) {
        t.park_count[tile] = 0;
        t.park_fmin[tile] = 0xFFFFFFFFu;
      }
    }
    __syncthreads();
    // layers: everything below stays in LDS until the tile is done
    uint32_t cur = 0;
    for (uint32_t layer = 0; layer < 8u * TILE; ++layer) {
      const uint32_t n = min(s_n[cur], kCap);
      if (n == 0) break;
      for (uint32_t i = threadIdx.x; i < n; i += kTileBlock) {
        const uint2 e = s_fr[cur][i];
        const uint32_t li = e.x, gv = e.y;
        if (gv != s_g[li]) continue;             // improved since it was queued
        const int x = (int)(x0 + li % TILE), y = (int)(y0 + li / TILE);
        const uint32_t cell = (uint32_t)y * a.width + (uint32_t)x;
        if (cell == a.target) {
          atomicMin(&st->best, gv);
          continue;
        }
        if (best != 0xFFFFFFFFu && gv + grid_heur(a.conn, x, y, (int)a.tx, (int)a.ty) >= best) continue;
        ++expanded;
        atomicOr(&s_exp[li >> 5], 1u << (li & 31));
        for (uint32_t k = 0; k < a.conn; ++k) {
          const int nx = x + nbr_dx(k), ny = y + nbr_dy(k);
          const uint32_t hx = (uint32_t)(nx - (int)x0 + 1), hy = (uint32_t)(ny - (int)y0 + 1);   // in [0, TILE + 1]
          const bool inside = hx - 1u < (uint32_t)TILE && hy - 1u < (uint32_t)TILE;
          const uint32_t lv = inside ? (hy - 1u) * TILE + (hx - 1u) : halo_index<TILE>(hx, hy);
          if (inside ? ((s_blk[lv >> 5] >> (lv & 31)) & 1u) : ((s_hblk[lv >> 5] >> (lv & 31)) & 1u)) continue;
          const uint32_t g2 = gv + nbr_cost(a.conn, k);
          const uint32_t f2 = g2 + grid_heur(a.conn, nx, ny, (int)a.tx, (int)a.ty);
          if (best != 0xFFFFFFFFu && f2 >= best) continue;
          if (!inside) {
            atomicMin(&s_out[lv], g2);            // sent on when the tile is done
            continue;
          }
          if (atomicMin(&s_g[lv], g2) <= g2) continue;
          atomicOr(&s_chg[lv >> 5], 1u << (lv & 31));
          const uint32_t pos = f2 <= fmax ? atomicAdd(&s_n[1u - cur], 1u) : kCap;
          if (pos < kCap) s_fr[1u - cur][pos] = make_uint2(lv, g2);
          else atomicOr(&s_def[lv >> 5], 1u << (lv & 31));
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) s_n[cur] = 0;
      cur = 1u - cur;
      __syncthreads();
    }
    // a layer cap reached: the remaining frontier waits too
    const uint32_t left = min(s_n[cur], kCap);
    for (uint32_t i = threadIdx.x; i < left; i += kTileBlock) {
      const uint2 e = s_fr[cur][i];
      if (e.y == s_g[e.x]) atomicOr(&s_def[e.x >> 5], 1u << (e.x & 31));
    }
    __syncthreads();
    // send on: the waiting cells of the tile, the improved cells around it; g and the closed
    // set back to HBM
    for (uint32_t i = threadIdx.x; i < kCells; i += kTileBlock) {
      if ((s_exp[i >> 5] >> (i & 31)) & 1u) {
        const uint32_t cell = (y0 + i / TILE) * a.width + x0 + i % TILE;
        atomicOr(&a.closed[cell >> 5], 1u << (cell & 31));
      }
      if (!((s_chg[i >> 5] >> (i & 31)) & 1u) && !((s_def[i >> 5] >> (i & 31)) & 1u)) continue;
      const uint32_t x = x0 + i % TILE, y = y0 + i / TILE;
      const uint32_t cell = y * a.width + x, gv = s_g[i];
      if ((s_chg[i >> 5] >> (i & 31)) & 1u) atomicMin(&a.g[cell], gv);
      if ((s_def[i >> 5] >> (i & 31)) & 1u) {    // parked until the window reaches it
        const uint32_t pos = atomicAdd(&t.park_count[tile], 1u);
        if (pos < t.cap) t.park[(size_t)tile * t.cap + pos] = make_uint2(cell, gv);
        else atomicOr(t.error, kErrQueueOverflow);
        atomicMin(&t.park_fmin[tile], gv + grid_heur(a.conn, (int)x, (int)y, (int)a.tx, (int)a.ty));
      }
    }
    for (uint32_t i = threadIdx.x; i < kHalo; i += kTileBlock) {
      const uint32_t g2 = s_out[i];
      if (g2 == 0xFFFFFFFFu) continue;
      uint32_t hx, hy;
      halo_cell<TILE>(i, &hx, &hy);
      const uint32_t nx = x0 + hx - 1u, ny = y0 + hy - 1u;   // in the grid: only such cells are written
      const uint32_t v = ny * a.width + nx;
      if (atomicMin(&a.g[v], g2) <= g2) continue;
      inbox_push(t, nxt, (ny / TILE) * t.tiles_x + nx / TILE, v, g2);
      fmin_pushed = min(fmin_pushed, g2 + grid_heur(a.conn, (int)nx, (int)ny, (int)a.tx, (int)a.ty));
    }
  }
  // one atomic per block for the smallest f sent on and the expansions
  uint32_t m = wave_min_u32(fmin_pushed);
  unsigned long long ex = expanded;
  for (int off = 32; off > 0; off >>= 1) ex += __shfl_xor(ex, off, kWave);
  if (lane_id() == 0) {
    s_red[threadIdx.x / kWave][0] = m;
    s_red[threadIdx.x / kWave][1] = (uint32_t)ex;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t mm = s_red[0][0];
    unsigned long long ee = s_red[0][1];
    for (int w = 1; w < kTileBlock / kWave; ++w) {
      mm = min(mm, s_red[w][0]);
      ee += s_red[w][1];
    }
    if (mm != 0xFFFFFFFFu) atomicMin(&st->fmin_next[nxt], mm);
    if (ee) atomicAdd(&st->expanded, ee);
  }
}

// After a tile step: the smallest f pending bounds the next window; the search is done when
// nothing is pending or the best goal cost is <= it.
// The smallest f parked in any tile (one thread per tile).
__global__ void astar_tile_scan_kernel(AstarState* st, AstarTiles t) {
  __shared__ uint32_t s_m[4];
  const uint32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = 0xFFFFFFFFu;
  if (!st->done && tile < t.tiles && t.park_count[tile]) v = t.park_fmin[tile];
  v = wave_min_u32(v);
  if (lane_id() == 0) s_m[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / kWave; ++w) v = min(v, s_m[w]);
    v = min(v, s_m[0]);
    if (v != 0xFFFFFFFFu) atomicMin(&st->park_min, v);
  }
}

// After the publication: tiles whose parked entries the new window reaches join the next
// step's list (unless an arrival listed them already).
__global__ void astar_tile_wake_kernel(AstarState* st, AstarTiles t) {
  const uint32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (st->done || tile >= t.tiles) return;
  const uint32_t par = st->parity;
  if (t.park_count[tile] && t.park_fmin[tile] <= st->fmax && t.count[(size_t)par * t.tiles + tile] == 0)
    t.active[(size_t)par * t.tiles + atomicAdd(&t.n_active[par], 1u)] = tile;
}

__global__ void astar_tile_publish_kernel(AstarState* st, uint32_t* n_active, uint32_t window) {
  if (st->done) return;
  const uint32_t par = st->parity, nxt = 1u - par;
  const uint32_t fmin = min(st->fmin_next[nxt], st->park_min);
  st->park_min = 0xFFFFFFFFu;                   // recomputed by the next step's scan
  st->fmin_next[par] = 0xFFFFFFFFu;             // filled by the next step
  n_active[par] = 0;                            // this step's list is consumed; refilled by the next step
  st->parity = nxt;
  st->fmax = fmin > 0xFFFFFFFFu - window ? 0xFFFFFFFFu : fmin + window;
  st->done = (fmin == 0xFFFFFFFFu || (st->best != 0xFFFFFFFFu && st->best <= fmin)) ? 1u : 0u;
  st->steps += 1;
}

__global__ void astar_tile_seed_kernel(AstarDev a, AstarTiles t, uint32_t source, uint32_t tile_side) {
  if (threadIdx.x || blockIdx.x) return;
  const uint32_t x = source % a.width, y = source / a.width;
  a.g[source] = 0;
  const uint32_t tile = (y / tile_side) * t.tiles_x + x / tile_side;
  t.inbox[(size_t)tile * t.cap] = make_uint2(source, 0u);
  t.count[tile] = 1;
  t.active[0] = tile;
  t.n_active[0] = 1;
  t.n_active[1] = 0;
  const uint32_t h0 = grid_heur(a.conn, (int)x, (int)y, (int)a.tx, (int)a.ty);
  *a.st = AstarState{0xFFFFFFFFu, 0u, 0ull, 0ull, 0ull, 0u,
                     h0 > 0xFFFFFFFFu - a.window ? 0xFFFFFFFFu : h0 + a.window, {0xFFFFFFFFu, 0xFFFFFFFFu},
                     0xFFFFFFFFu};
}

__global__ void popcount_kernel(const uint32_t* bits, uint64_t words, unsigned long long* out) {
  unsigned long long c = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    c += (unsigned long long)__popc(bits[i]);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
  if (lane_id() == 0 && c) atomicAdd(out, c);
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
  q.insert_or_update(Event{h0, list_of(source, a.lists), h0, source}, (int)list_of(source, a.lists));
  *a.st = AstarState{0xFFFFFFFFu, 0u, 0ull, 0ull, 0ull, 0u, 0u, {0xFFFFFFFFu, 0xFFFFFFFFu}, 0xFFFFFFFFu};
}

}  // namespace

gpq_status AstarModel::create(QueueBase* queue, const gpq_astar_config& c, AstarModel** out) {
  GPQ_CHECK_ARG(queue, "null queue");
  GPQ_CHECK_ARG(queue->cfg.kind == GPQ_LAQ || queue->cfg.kind == GPQ_LHQ || queue->cfg.kind == GPQ_LSQ,
                "A*: the open lists must be a per-LP queue (LAQ, LHQ or LSQ)");
  GPQ_CHECK_ARG(std::isinf(queue->cfg.lookahead) && std::isinf(queue->cfg.horizon),
                "A*: the open-list queue needs lookahead = horizon = +inf (peek = pop the minimum)");
  GPQ_CHECK_ARG(c.connectivity == 4 || c.connectivity == 8, "A*: connectivity must be 4 or 8");
  GPQ_CHECK_ARG(c.tile == 0 || c.tile == 16 || c.tile == 32, "A*: tile must be 0, 16 or 32");
  GPQ_CHECK_ARG(c.width >= 1 && c.height >= 1 && (uint64_t)c.width * c.height < (1ull << 31), "A*: bad grid size");
  GPQ_CHECK_ARG(c.sx < c.width && c.tx < c.width && c.sy < c.height && c.ty < c.height, "A*: endpoint outside grid");
  GPQ_CHECK_ARG(!queue->cfg.update_ids || (queue->cfg.kind == GPQ_LAQ &&
                                            (uint64_t)queue->cfg.update_ids >= (uint64_t)c.width * c.height),
                "A*: a decrease-key LAQ needs update_ids >= the grid's cells");
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
  if (c.tile) {
    AstarTiles& t = m->tiles_;
    t.tiles_x = (c.width + c.tile - 1) / c.tile;
    t.tiles = t.tiles_x * ((c.height + c.tile - 1) / c.tile);
    t.cap = 4u * c.tile * c.tile;                // deferred entries plus arrivals across the borders
    t.error = &queue->ws->error;
    if ((s = m->arena.alloc(&t.inbox, 2ull * t.tiles * t.cap)) != GPQ_OK ||
        (s = m->arena.alloc(&t.count, 2ull * t.tiles)) != GPQ_OK ||
        (s = m->arena.alloc(&t.active, 2ull * t.tiles)) != GPQ_OK || (s = m->arena.alloc(&t.n_active, 2)) != GPQ_OK ||
        (s = m->arena.alloc(&t.park, (size_t)t.tiles * t.cap)) != GPQ_OK ||
        (s = m->arena.alloc(&t.park_count, t.tiles)) != GPQ_OK || (s = m->arena.alloc(&t.park_fmin, t.tiles)) != GPQ_OK) {
      delete m;
      return s;
    }
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
  if (cfg.tile) return restart_tiles();
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(cells)), dim3(256), 0, q->stream, g, cells, 0xFFFFFFFFu);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks((cells + 31) / 32)), dim3(256), 0, q->stream, closed,
                     (cells + 31) / 32, 0u);
  GPQ_HIP(hipGetLastError());
  // entries left by an earlier search are dropped as stale (their g is not the new g[cell]);
  // in decrease-key mode they would hold their ids' marks, so the lists start empty
  if (q->cfg.kind == GPQ_LAQ && q->cfg.update_ids) GPQ_TRY(static_cast<LaqQueue*>(q)->empty_all());
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

// Tile search: fresh g / closed arrays, empty inboxes, the source in its tile's inbox.
gpq_status AstarModel::restart_tiles() {
  const uint64_t cells = (uint64_t)cfg.width * cfg.height;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(cells)), dim3(256), 0, q->stream, g, cells, 0xFFFFFFFFu);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks((cells + 31) / 32)), dim3(256), 0, q->stream, closed,
                     (cells + 31) / 32, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(2ull * tiles_.tiles)), dim3(256), 0, q->stream, tiles_.count,
                     2ull * tiles_.tiles, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(tiles_.tiles)), dim3(256), 0, q->stream, tiles_.park_count,
                     (uint64_t)tiles_.tiles, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_blocks(tiles_.tiles)), dim3(256), 0, q->stream, tiles_.park_fmin,
                     (uint64_t)tiles_.tiles, 0xFFFFFFFFu);
  hipLaunchKernelGGL(astar_tile_seed_kernel, dim3(1), dim3(64), 0, q->stream, view(), tiles_,
                     cfg.sy * cfg.width + cfg.sx, cfg.tile);
  GPQ_HIP(hipGetLastError());
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
  if (cfg.tile) {
    const dim3 grid(std::min<uint32_t>(tiles_.tiles, kTileGrid));
    if (cfg.tile == 16)
      hipLaunchKernelGGL(astar_tile_kernel<16>, grid, dim3(kTileBlock), 0, q->stream, d, tiles_);
    else
      hipLaunchKernelGGL(astar_tile_kernel<32>, grid, dim3(kTileBlock), 0, q->stream, d, tiles_);
    const dim3 all((tiles_.tiles + 255) / 256);
    hipLaunchKernelGGL(astar_tile_scan_kernel, all, dim3(256), 0, q->stream, st, tiles_);
    hipLaunchKernelGGL(astar_tile_publish_kernel, dim3(1), dim3(1), 0, q->stream, st, tiles_.n_active, cfg.window);
    hipLaunchKernelGGL(astar_tile_wake_kernel, all, dim3(256), 0, q->stream, st, tiles_);
    GPQ_HIP(hipGetLastError());
    return GPQ_OK;
  }
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
  if (cfg.tile) {            // distinct expanded cells: the closed bitmap's population
    const uint64_t words = ((uint64_t)cfg.width * cfg.height + 31) / 32;
    hipLaunchKernelGGL(popcount_kernel, dim3(std::min<uint64_t>(grid_blocks(words), 1024)), dim3(256), 0, q->stream,
                       closed, words, &st->closed);
    GPQ_HIP(hipGetLastError());
    GPQ_TRY(poll(&s));
  }
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

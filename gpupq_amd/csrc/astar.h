// astar.h -- host classes of the A* models: one large query over a per-LP queue whose LPs
// are the open lists, and batches of independent queries searched one workgroup each.
#pragma once

#include <vector>

#include "phold.h"

namespace gpq {

// Heuristics (admissible and consistent): Manhattan for 4-connected unit costs, octile
// 10*max + 4*min for 8-connected 10/14 costs (src/a_star.cu:123-135 uses Euclidean floats).
__host__ __device__ __forceinline__ uint32_t grid_heur(uint32_t conn, int x, int y, int tx, int ty) {
  const uint32_t dx = (uint32_t)(x > tx ? x - tx : tx - x), dy = (uint32_t)(y > ty ? y - ty : ty - y);
  if (conn == 4) return dx + dy;
  const uint32_t mx = dx > dy ? dx : dy, mn = dx > dy ? dy : dx;
  return 10u * mx + 4u * mn;
}

__device__ __forceinline__ bool bit_get(const uint32_t* bits, uint64_t i) { return (bits[i >> 5] >> (i & 31)) & 1u; }

// Neighbour k: 0..3 = E, W, S, N (cost 1 or 10), 4..7 = the diagonals (cost 14).
__device__ __forceinline__ int nbr_dx(uint32_t k) { return k < 4 ? (k == 0 ? 1 : (k == 1 ? -1 : 0)) : ((k & 1) ? -1 : 1); }
__device__ __forceinline__ int nbr_dy(uint32_t k) { return k < 4 ? (k == 2 ? 1 : (k == 3 ? -1 : 0)) : ((k & 2) ? -1 : 1); }
__device__ __forceinline__ uint32_t nbr_cost(uint32_t conn, uint32_t k) { return conn == 4 ? 1u : (k < 4 ? 10u : 14u); }

struct AstarState {
  uint32_t best;                 // best goal cost found so far (0xFFFFFFFF: none)
  uint32_t done;
  unsigned long long expanded;
  unsigned long long closed;
  unsigned long long steps;
  // tile search: inbox parity read this step, window bound, smallest f pushed per parity,
  // smallest f held back in the tiles' parking areas
  uint32_t parity, fmax, fmin_next[2], park_min;
};

// Tile search (gpq_astar_config.tile > 0): per tile and parity an inbox of (cell, g) entries.
struct AstarTiles {
  uint2* inbox;                  // [2][tiles][cap]
  uint32_t* count;               // [2][tiles]
  uint32_t* active;              // [2][tiles]: tiles to search in the step of that parity
  uint32_t* n_active;            // [2]
  uint2* park;                   // [tiles][cap]: a tile's entries beyond the window, until it reaches them
  uint32_t* park_count;          // [tiles]
  uint32_t* park_fmin;           // [tiles]: smallest f parked (0xFFFFFFFF: none)
  uint32_t* error;               // device error flags (the queue's window state)
  uint32_t tiles, tiles_x, cap;
};

struct AstarDev {
  const uint32_t* blocked;       // obstacle bitmap, 1 bit per cell
  uint32_t* g;                   // best known cost per cell
  uint32_t* closed;              // closed-set bitmap (cells expanded at least once)
  AstarState* st;
  const WindowState* ws;         // the open-list queue's window: gmin = smallest f of all lists
  uint32_t width, height, conn, tx, ty, target, lists, window;
};

__global__ void grid_generate_kernel(uint32_t* bits, uint64_t cells, uint32_t seed, uint32_t tag, uint32_t query,
                                     uint32_t threshold, uint64_t force_a, uint64_t force_b);
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v);
__global__ void grid_to_bytes_kernel(const uint32_t* bits, uint8_t* bytes, uint64_t cells);

class AstarModel {
 public:
  static gpq_status create(QueueBase* queue, const gpq_astar_config& c, AstarModel** out);
  ~AstarModel();
  gpq_status init();
  gpq_status set_grid(const uint8_t* bytes);
  gpq_status grid_bytes(uint8_t* out);
  gpq_status step();
  gpq_status run(gpq_drive_mode mode, uint32_t batch, gpq_astar_report* rep);
  gpq_status path(uint32_t* cells, uint32_t cap, uint32_t* len);

  gpq_astar_config cfg{};
  QueueBase* q = nullptr;

 private:
  gpq_status generate();
  gpq_status restart();
  gpq_status poll(AstarState* out);
  AstarDev view() const;
  DeviceArena arena;
  uint32_t* blocked = nullptr;
  uint32_t* g = nullptr;
  uint32_t* closed = nullptr;
  AstarState* st = nullptr;
  AstarTiles tiles_{};           // tile search only
  gpq_status restart_tiles();
  AstarState host_st{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_batch = 0;
};

gpq_status query_synthetic(uint32_t seed, uint32_t q, gpq_query_spec* out);
gpq_status batch_autotune(const gpq_query_spec* specs, uint32_t count, uint32_t sample, const gpq_batch_tuning& base,
                          gpq_batch_tuning* out, double* best_seconds);

class AstarBatch {
 public:
  static gpq_status create(const gpq_query_spec* specs, uint32_t count, const gpq_batch_tuning& t, AstarBatch** out);
  ~AstarBatch();
  gpq_status run(gpq_drive_mode mode, double* seconds);
  gpq_status results(int64_t* costs, uint64_t* expanded);
  gpq_status timing(double* start_us, double* finish_us);
  gpq_status grid_bytes(uint32_t q, uint8_t* out);
  // An optimal path of query q (source to target): q searched again alone in slot 0.
  gpq_status path(uint32_t q, uint32_t* cells, uint32_t cap, uint32_t* len);

  // Searches the queries d_ids[0..n) (device array) in one persistent launch on stream s,
  // workgroups in slots [slot_base, slot_base + slots), work counter d_counter.
  gpq_status launch(const uint32_t* d_ids, uint32_t n, uint32_t slot_base, uint32_t slots, uint32_t* d_counter,
                    hipStream_t s);
  // One query (device id d_id, spec spec_index) alone, one launch + host sync per iteration.
  gpq_status plain_one(const uint32_t* d_id, uint32_t spec_index, uint32_t slot, uint32_t* d_done, uint32_t* h_done,
                       hipStream_t s);
  uint32_t resident_slots() const { return resident; }
  const uint32_t* device_ids() const { return d_ids_all; }

  std::vector<gpq_query_spec> specs;
  gpq_batch_tuning tuning{};
  hipStream_t stream = nullptr;

  struct Dev;                    // device-side layout (astar_batch.hip)

 private:
  DeviceArena arena;
  void* dev = nullptr;           // Dev*
  uint32_t* d_ids_all = nullptr;
  uint32_t* counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t resident = 0;
  std::vector<uint64_t> cell_off;   // first cell of each query in the pooled arrays
};

}  // namespace gpq

struct gpq_astar {
  gpq::AstarModel* impl;
};
struct gpq_astar_batch {
  gpq::AstarBatch* impl;
};

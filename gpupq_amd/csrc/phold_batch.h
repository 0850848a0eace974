// phold_batch.h -- many PHOLD instances of uneven size in one local array queue.
//
// Each instance owns a range of LPs (a whole number of 256-LP post blocks) of one shared
// LAQ; an instance table in HBM gives every LP its instance.  One step of the pool is three
// launches over the whole pool -- the model step, the LAQ merge, and a publication of every
// instance's own window -- so a hipGraph replay advances all instances together, each in
// its own window, and an instance that finished idles.  Instances are admitted into free LP
// ranges and retired while the pool runs (continuous batching).
#pragma once

#include <vector>

#include "phold.h"

namespace gpq {

class PholdPool {
 public:
  // capacity_lps: LPs the pool can hold at once (rounded up to whole 256-LP blocks).
  // far_per_lp > 0: two-tier LAQ, each instance with its own near horizon.
  static gpq_status create(uint32_t capacity_lps, uint32_t capacity_per_lp, uint32_t stage_per_lp,
                           uint32_t max_instances, PholdPool** out, uint32_t far_per_lp = 0);
  ~PholdPool();

  // Bytes of HBM one LP of the pool holds (ring, staging, queue words, model state).
  static uint64_t bytes_per_lp(uint32_t capacity_per_lp, uint32_t stage_per_lp, uint32_t far_per_lp = 0);

  // Admits an instance into a free LP range: false in *ok when none is large enough.
  gpq_status admit(const gpq_phold_instance& inst, uint32_t* slot, bool* ok);
  // Steps captured per graph replay.
  gpq_status prepare(uint32_t batch);
  // One replay of `batch` steps, then the instance windows copied to the host mirror and
  // `done_event` recorded.  Non-blocking.
  gpq_status launch_async();
  bool replay_done();
  // Instance slots whose window reports done (from the last completed replay's mirror).
  void finished(std::vector<uint32_t>* slots) const;
  bool any_active() const;
  // Reads the instance's per-LP counters and hashes, frees its range.
  gpq_status retire(uint32_t slot, uint64_t* trace_hash, uint64_t* events);
  // retire() for several instances with one host round trip; trace_hash[k] and events[k]
  // belong to slots[k].
  gpq_status retire_many(const std::vector<uint32_t>& slots, uint64_t* trace_hash, uint64_t* events);
  uint32_t error_flags_host() const;
  gpq_status errors(uint32_t* flags);

  // HBM the pool holds: its queue's arena and the pool's own tables and model state.
  uint64_t device_bytes() const;
  QueueBase* queue() { return q_; }
  uint32_t capacity() const { return capacity_; }
  uint32_t free_lps() const;

 private:
  LaqQueue* q_ = nullptr;
  DeviceArena arena_;
  PoolInst* d_inst_ = nullptr;
  WindowState* d_ws_ = nullptr;
  uint8_t* d_block_live_ = nullptr;     // per 256-LP block: 1 while an instance owns it
  uint32_t* d_block_win_ = nullptr;     // per block, the owning instance slot (valid while the block is live)
  bool far_ = false;
  uint32_t* d_count_ = nullptr;
  uint64_t* d_hash_ = nullptr;
  WindowState* h_ws_ = nullptr;          // pinned mirror of d_ws_
  uint32_t* h_count_ = nullptr;          // pinned staging of retired instances' counters, by LP
  uint64_t* h_hash_ = nullptr;           // and hashes
  hipEvent_t done_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  uint32_t batch_ = 0;
  uint32_t capacity_ = 0, max_inst_ = 0;
  std::vector<PoolInst> inst_;           // host copy of the table
  std::vector<uint8_t> block_used_;      // per 256-LP block
  bool pending_ = false;
};

// The batched entry point: every instance admitted at once into a pool sized for all of
// them; GPQ_DRIVE_GRAPH advances them together, GPQ_DRIVE_PLAIN runs each alone through its
// own LAQ with a host synchronization per step (the reference's des_run pattern).
//
// With calendar_windows set, both drives use calendar-mode LAQs and the merge-and-commit step
// (drain 2).  Then GPQ_DRIVE_GRAPH puts all instances in one calendar LAQ on one window
// sequence (window width: the smallest lookahead; the last window ends at the latest end
// time): an instance's committed events are its events below its end time whatever the
// windows, so its trace equals its run alone, and one kernel per step advances all of them.
class PholdBatch {
 public:
  static gpq_status create(const gpq_phold_instance* inst, uint32_t count, uint32_t capacity_per_lp,
                           uint32_t stage_per_lp, uint32_t far_per_lp, PholdBatch** out);
  gpq_status run(gpq_drive_mode mode, uint32_t batch, double* seconds);
  gpq_status results(uint64_t* hashes, uint64_t* events);

  std::vector<gpq_phold_instance> inst;
  uint32_t capacity_per_lp = 128, stage_per_lp = 32, far_per_lp = 0, calendar_windows = 0;
  std::vector<uint64_t> hash, events;

 private:
  gpq_status run_calendar_union(uint32_t batch, double* seconds);
};

}  // namespace gpq

struct gpq_phold_batch {
  gpq::PholdBatch* impl;
};

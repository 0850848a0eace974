// phold.h -- host class of the PHOLD model handle and the queue-kind dispatch.
#pragma once

#include "queues.h"

namespace gpq {

// Calls f(view) with the concrete device view of q's queue kind.
template <class F>
gpq_status with_view(QueueBase* q, F&& f) {
  switch (q->cfg.kind) {
    case GPQ_LAQ: return f(static_cast<LaqQueue*>(q)->view());
    case GPQ_LHQ: return f(static_cast<LhqQueue*>(q)->view());
    case GPQ_LSQ: return f(static_cast<LsqQueue*>(q)->view());
    case GPQ_GAQ: return f(static_cast<GaqQueue*>(q)->view());
    case GPQ_HQ: return f(static_cast<HqQueue*>(q)->view());
    default:
      set_error("queue kind not provided");
      return GPQ_ERR_UNSUPPORTED;
  }
}

// One PHOLD instance of several sharing a queue: its LP range and model parameters.
struct PoolInst {
  uint32_t base, lps, init_per_lp, seed;
  float end_time, lookahead;
  uint32_t p_self, active;
  uint32_t block_begin, block_end;      // post blocks [block_begin, block_end) cover the range
  uint32_t pad[2];
};

// Several instances in one calendar-mode LAQ, all on the queue's window sequence (the queue's
// lookahead is at most every instance's, its horizon the latest end time): the LPs of 256-LP
// block b belong to instance inst[block_inst[b]].  One calendar step over all of them, and
// initial event j of every LP (staged for queue_post_init, or parked at or past the
// instance's end time).
gpq_status launch_calendar_instances(LaqQueue* q, const PoolInst* inst, const uint32_t* block_inst, uint32_t* count,
                                     uint64_t* hash);
gpq_status launch_populate_instances(LaqQueue* q, const PoolInst* inst, const uint32_t* block_inst, uint32_t j);

// queue_insert (update = false) or queue_insert_or_update (true) of n device items.
gpq_status launch_insert_items(QueueBase* q, const Event* d_items, uint32_t n, bool update = false);
gpq_status launch_peek_items(QueueBase* q, int set_done, Event* d_out, int32_t* d_index, uint32_t* d_found,
                             uint32_t max);

class PholdModel {
 public:
  static gpq_status create(QueueBase* queue, const gpq_phold_config& c, PholdModel** out);
  ~PholdModel();
  gpq_status init();
  gpq_status launch_step();
  gpq_status step();
  gpq_status events(uint64_t* out);
  gpq_status lp_state(uint32_t* counts, uint64_t* hashes);
  gpq_status trace_hash(uint64_t* out);
  gpq_status run(gpq_drive_mode mode, uint32_t batch, gpq_des_report* rep);
  gpq_status prepare_graph(uint32_t batch);
  gpq_status launch_async(hipEvent_t done);
  gpq_status profile(uint32_t batch, gpq_kernel_profile* out);
  // drain mode 2: the merge-and-commit kernel (queue_post then only publishes the window),
  // and the closing fold of the last step's staged events (a no-op in the other modes)
  gpq_status launch_fused(uint32_t settle);
  gpq_status settle();
  bool is_calendar() const;     // the queue is a calendar-mode LAQ

  gpq_phold_config cfg{};
  QueueBase* q = nullptr;
  DeviceArena arena;
  uint32_t* count = nullptr;
  uint64_t* hash = nullptr;
  uint32_t* last_ts = nullptr;
  uint64_t* sum = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_batch = 0;
};

}  // namespace gpq

struct gpq_phold {
  gpq::PholdModel* impl;
};

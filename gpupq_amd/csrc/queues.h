// queues.h -- host classes of the queue kinds.  Each exposes view(): the POD
// struct passed by value to kernels, whose device methods are the queue.cuh verbs.
#pragma once

#include "gpq_host.h"
#include "../../include/gpupq_device/queue_laq.cuh"
#include "../../include/gpupq_device/queue_lhq.cuh"
#include "../../include/gpupq_device/queue_lsq.cuh"
#include "../../include/gpupq_device/queue_gaq.cuh"
#include "../../include/gpupq_device/queue_hq.cuh"
#include "sortprim.h"

namespace gpq {

QueueBase* make_queue(const gpq_queue_config& cfg);   // by cfg.kind; nullptr if unsupported

class LaqQueue final : public QueueBase {
 public:
  explicit LaqQueue(const gpq_queue_config& c) : QueueBase(c) {}
  gpq_status init() override;
  gpq_status post() override;
  gpq_status post_init() override;
  gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) override;
  gpq_status pending(uint64_t* n) override;
  gpq_status length(uint32_t lp, uint32_t* n) override;
  gpq_status clear_lp(uint32_t lp) override;
  LaqView view() const { return v_; }
  // Merge of the staged inserts plus the window publication (queue_post without a model
  // step that merges by itself).
  gpq_status merge_publish();
  // The merge kernel of post() alone: one partial minimum per 256-LP block in `partial`,
  // no window publication (callers that keep several windows publish their own).
  gpq_status merge_only();
  static constexpr uint32_t kBlockLps = 256;
  static constexpr uint32_t kCalBlockLps = 256;   // LPs per commit block of the calendar step
  // Post blocks whose flag is 0 are skipped (their LPs hold nothing); nullptr = all live.
  void set_block_live(const uint8_t* flags) { v_.block_live = flags; }
  // Allocates the second staging buffer (stage_alt / ins_alt, empty) that a model step which
  // merges and commits in one kernel needs; idempotent.  From then on post() only advances
  // the window by the lookahead (the step did the merge), post_init() still merges.
  gpq_status enable_alt_stage();
  // After a closing merge of such a step: both buffers are empty, the next step reads stage.
  gpq_status reset_stage_select();
  uint32_t group() const { return group_; }
  // Decrease-key mode: every LP emptied, every id unmarked (a new search on the same queue).
  gpq_status empty_all();
  float far_step() const { return far_step_; }
  // Calendar mode: whether it is on; and, once every initial event is in (post_init), the
  // ends of the first windows and the filing of the initial events in their buckets.
  bool calendar() const { return v_.wlog != nullptr; }
  gpq_status calendar_start();
  static constexpr uint32_t kGatherCap = 1u << 16;   // records one LP may hold for lp_contents
  // Two-tier mode, several windows: block b follows the near horizon of ws_array[win[b]].
  void set_block_windows(const uint32_t* win, const WindowState* ws_array) {
    v_.block_win = win;
    v_.block_ws = ws_array;
  }

 private:
  gpq_status init_calendar();
  gpq_status calendar_gather(uint32_t lp, std::vector<Event>* out);
  LaqView v_{};
  uint64_t* sum_ = nullptr;
  Event* gather_ = nullptr;    // calendar mode: one LP's records for lp_contents / length (+ a count word)
  float far_step_ = 0.0f;
  uint32_t group_ = 16;   // lanes per LP in the post kernel
};

#define GPQ_LOCAL_QUEUE_CLASS(Name, View)                                        \
  class Name final : public QueueBase {                                          \
   public:                                                                       \
    explicit Name(const gpq_queue_config& c) : QueueBase(c) {}                   \
    gpq_status init() override;                                                  \
    gpq_status post() override;                                                  \
    gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) override; \
    gpq_status pending(uint64_t* n) override;                                    \
    gpq_status length(uint32_t lp, uint32_t* n) override;                        \
    gpq_status clear_lp(uint32_t lp) override;                                   \
    View view() const { return v_; }                                             \
                                                                                 \
   private:                                                                      \
    View v_{};                                                                   \
    uint64_t* sum_ = nullptr;                                                    \
  };

GPQ_LOCAL_QUEUE_CLASS(LhqQueue, LhqView)
GPQ_LOCAL_QUEUE_CLASS(LsqQueue, LsqView)

class GaqQueue final : public QueueBase {
 public:
  explicit GaqQueue(const gpq_queue_config& c) : QueueBase(c) {}
  gpq_status init() override;
  gpq_status pre() override;
  gpq_status post() override;
  gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) override;
  gpq_status pending(uint64_t* n) override;
  gpq_status length(uint32_t lp, uint32_t* n) override;
  gpq_status clear_lp(uint32_t lp) override;
  GaqView view() const { return v_; }

 private:
  gpq_status snapshot(std::vector<Event>* out);
  GaqView v_{};
  Event* fel_[2] = {nullptr, nullptr};
  Event* rest_ = nullptr;
  Event* sort_[2] = {nullptr, nullptr};
  MergeTask* task_ = nullptr;
  uint32_t* first_ = nullptr;
  uint32_t* scratch_ = nullptr;
  uint32_t* splits_ = nullptr;  // merge-path splits of the merge's tile boundaries
  uint32_t split_stride_ = 0;
  uint32_t* seg_pre_ = nullptr; // segment offsets of the step's striped inserts
  uint32_t cap_ = 0;
  uint32_t tile_ = kSortTile;   // records per sort / merge block
};

class HqQueue final : public QueueBase {
 public:
  explicit HqQueue(const gpq_queue_config& c) : QueueBase(c) {}
  ~HqQueue() override;
  gpq_status init() override;
  gpq_status post() override;
  gpq_status post_init() override;
  gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) override;
  gpq_status pending(uint64_t* n) override;
  gpq_status length(uint32_t lp, uint32_t* n) override;
  gpq_status clear_lp(uint32_t lp) override;
  gpq_status node_contents(uint32_t v, Event* out, uint32_t cap, uint32_t* n);
  HqView view() const { return v_; }

 private:
  gpq_status step(bool select);
  gpq_status merge_root(uint32_t h, hipStream_t s);
  gpq_status step_heap(uint32_t h, bool select, hipStream_t s);
  gpq_status snapshot(std::vector<Event>* out, std::vector<uint32_t>* counts);
  HqView v_{};
  uint32_t r_ = 0, nodes_ = 0, levels_ = 0, slots_ = 1, depth_ = 0;
  uint32_t heaps_ = 1, lps_per_heap_ = 0;   // nodes_ is per heap
  uint64_t mcap_ = 0;                       // merged-buffer records per heap
  size_t scratch_n_ = 0, tasks_n_ = 0;      // per-heap sizes of scratch_ and tasks_
  std::vector<hipStream_t> hs_;             // one stream per heap (several heaps only)
  std::vector<hipEvent_t> ev_;              // joins of both phases, then the two forks
  Event* nodes_buf_ = nullptr;
  uint32_t* counts_ = nullptr;
  uint8_t* dirty_ = nullptr;
  Event* sort_[2] = {nullptr, nullptr};
  Event* merged_ = nullptr;
  Event* rest_ = nullptr;
  Event* carry_ = nullptr;
  Event* chunk_buf_ = nullptr;
  Event* tmp2_ = nullptr;
  Event* tmp1_ = nullptr;
  uint32_t* first_ = nullptr;
  uint32_t* scratch_ = nullptr;
  MergeTask* tasks_ = nullptr;
  uint32_t* splits_ = nullptr;  // merge-path splits, split_stride_ words per task slot
  uint32_t split_stride_ = 0;
  uint32_t* seg_pre_ = nullptr; // per heap, segment offsets of the step's striped inserts
  uint64_t* sum_ = nullptr;
  uint32_t tile_ = kSortTile;   // records per sort / merge block
  // the split buffer of task slot t (tasks_ + t)
  uint32_t* splits_of(const MergeTask* t) const { return splits_ + (size_t)(t - tasks_) * split_stride_; }
};

}  // namespace gpq

// queues.h -- host classes of the queue kinds.  Each exposes view(): the POD
// struct passed by value to kernels, whose device methods are the queue.cuh verbs.
#pragma once

#include "gpq_host.h"
#include "queue_laq.cuh"
#include "queue_lhq.cuh"
#include "queue_lsq.cuh"

namespace gpq {

QueueBase* make_queue(const gpq_queue_config& cfg);   // by cfg.kind; nullptr if unsupported

class LaqQueue final : public QueueBase {
 public:
  explicit LaqQueue(const gpq_queue_config& c) : QueueBase(c) {}
  gpq_status init() override;
  gpq_status post() override;
  gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) override;
  gpq_status pending(uint64_t* n) override;
  gpq_status length(uint32_t lp, uint32_t* n) override;
  gpq_status clear_lp(uint32_t lp) override;
  LaqView view() const { return v_; }

 private:
  LaqView v_{};
  uint64_t* sum_ = nullptr;
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

}  // namespace gpq

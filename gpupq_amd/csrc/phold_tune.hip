// phold_tune.hip -- the PHOLD layout autotuner: calendar or sorted ring, capacity per LP,
// staging slots, far store, near-horizon step and step kind, each candidate timed on a prefix of
// the run itself (same LPs, same seed, end time cut short) through the LAQ.
#include <memory>

#include "phold.h"

namespace gpq {
namespace {

struct Candidate {
  uint32_t cap, stage, far;
  float far_step_la;   // near-horizon step in lookaheads (0 = the queue's default, 4)
  uint32_t drain;
  uint32_t cal;        // calendar mode: windows in the calendar (0 = off); cap = a window's log
                       // records per LP, far = beyond-the-calendar / parked records per LP
};

// The calendar layouts with the merge-and-commit step first, then the two-tier layouts with the
// same step, then the post kernel + drain step, then the one-tier ring.
constexpr Candidate kCandidates[] = {
    {8, 32, 128, 0.0f, 2, 64}, {8, 32, 128, 0.0f, 2, 32},
    {64, 32, 128, 0.0f, 2, 0}, {64, 32, 128, 6.0f, 2, 0}, {64, 32, 256, 0.0f, 2, 0},
    {64, 16, 128, 0.0f, 2, 0}, {64, 32, 128, 0.0f, 1, 0}, {128, 32, 0, 0.0f, 1, 0},
};

constexpr double kMargin = 1.02;

// Events per second and trace hash of one candidate over the prefix; false when it cannot
// hold the workload.
bool time_candidate(const gpq_queue_config& qc, const gpq_phold_config& pc, double* rate, uint64_t* hash) {
  std::unique_ptr<QueueBase> q(make_queue(qc));
  if (!q || q->init_common() != GPQ_OK || q->init() != GPQ_OK) return false;
  PholdModel* raw = nullptr;
  if (PholdModel::create(q.get(), pc, &raw) != GPQ_OK) return false;
  std::unique_ptr<PholdModel> m(raw);
  if (m->init() != GPQ_OK) return false;
  gpq_des_report r{};
  if (m->run(GPQ_DRIVE_GRAPH, 64, &r) != GPQ_OK) return false;
  if (m->trace_hash(hash) != GPQ_OK) return false;
  *rate = r.events_per_second;
  return true;
}

}  // namespace
}  // namespace gpq

extern "C" gpq_status gpq_phold_autotune(gpq_phold_config* model, gpq_queue_config* queue, float sample_end_time,
                                         gpq_phold_tuning* results, uint32_t max_results, uint32_t* n_results) {
  using namespace gpq;
  GPQ_CHECK_ARG(model && queue, "phold autotune: null argument");
  GPQ_CHECK_ARG(model->num_lps >= 1 && model->num_lps == queue->num_lps, "phold autotune: num_lps must match");
  GPQ_CHECK_ARG(model->end_time > 0.0f && queue->lookahead > 0.0f, "phold autotune: end_time and lookahead > 0");
  GPQ_CHECK_ARG(!results || max_results >= 1, "phold autotune: max_results must be >= 1");
  float t = sample_end_time > 0.0f ? sample_end_time : 5.0f;
  if (model->end_time < t) t = model->end_time;
  double best = 0.0;
  const Candidate* pick = nullptr;
  uint32_t n = 0;
  bool have_ref = false;
  uint64_t ref_hash = 0;
  for (const Candidate& c : kCandidates) {
    double rate = 0.0;
    uint64_t hash = 0;
    uint32_t status = 1;
    if (c.cal ? model->init_per_lp <= c.far : model->init_per_lp <= c.cap) {
      gpq_queue_config qc = *queue;
      qc.kind = GPQ_LAQ;
      qc.capacity_per_lp = c.cap;
      qc.stage_per_lp = c.stage;
      qc.far_per_lp = c.far;
      qc.far_step = c.far_step_la * queue->lookahead;
      qc.calendar_windows = c.cal;
      qc.horizon = t;
      qc.collect_stats = 0;
      qc.update_ids = 0;           // PHOLD records carry no ids: no decrease-key mode
      qc.update_per_lp = 0;
      gpq_phold_config pc = *model;
      pc.end_time = t;
      pc.drain = c.drain;
      pc.checks = 0;
      if (time_candidate(qc, pc, &rate, &hash)) {
        // every layout commits the same trace: one that does not is never picked
        status = 0;
        if (!have_ref) {
          have_ref = true;
          ref_hash = hash;
        } else if (hash != ref_hash) {
          status = 2;
        }
      }
      if (status) rate = 0.0;
    }
    if (results && n < max_results)
      results[n] = gpq_phold_tuning{c.cap, c.stage, c.far, c.far_step_la * queue->lookahead, c.drain, status, rate,
                                    c.cal};
    ++n;
    // candidates are listed in order of preference: a later one must be faster by more than
    // the run-to-run spread of the timing (2%) to replace the pick
    if (rate > (pick ? best * kMargin : 0.0)) {
      best = rate;
      pick = &c;
    }
  }
  if (n_results) *n_results = n < max_results || !results ? n : max_results;
  if (!pick) {
    set_error("phold autotune: no candidate layout holds the workload");
    return GPQ_ERR_CAPACITY;
  }
  queue->kind = GPQ_LAQ;
  queue->capacity_per_lp = pick->cap;
  queue->stage_per_lp = pick->stage;
  queue->far_per_lp = pick->far;
  queue->far_step = pick->far_step_la * queue->lookahead;
  queue->calendar_windows = pick->cal;
  queue->update_ids = 0;
  queue->update_per_lp = 0;
  model->drain = pick->drain;
  return GPQ_OK;
}

// Sort-tile autotuner of the global queues: the record sort and merge kernels with tiles of
// 2048 (the default, timed first), 1024 and 4096 records, each on the PHOLD prefix through the
// given queue kind, the trace cross-checked.
extern "C" gpq_status gpq_queue_autotune_sort(const gpq_phold_config* model, gpq_queue_config* queue,
                                              float sample_end_time, gpq_sort_tuning* results, uint32_t max_results,
                                              uint32_t* n_results) {
  using namespace gpq;
  GPQ_CHECK_ARG(model && queue, "sort autotune: null argument");
  GPQ_CHECK_ARG(queue->kind == GPQ_GAQ || queue->kind == GPQ_HQ, "sort autotune: the queue kind must be GAQ or HQ");
  GPQ_CHECK_ARG(model->num_lps >= 1 && model->num_lps == queue->num_lps, "sort autotune: num_lps must match");
  GPQ_CHECK_ARG(model->end_time > 0.0f && queue->lookahead > 0.0f, "sort autotune: end_time and lookahead > 0");
  GPQ_CHECK_ARG(!results || max_results >= 1, "sort autotune: max_results must be >= 1");
  float t = sample_end_time > 0.0f ? sample_end_time : 2.0f;
  if (model->end_time < t) t = model->end_time;
  static const uint32_t kTiles[] = {2048u, 1024u, 4096u};
  double best = 0.0;
  uint32_t pick = 0, n = 0;
  bool have_ref = false;
  uint64_t ref_hash = 0;
  for (uint32_t tile : kTiles) {
    gpq_queue_config qc = *queue;
    qc.sort_tile = tile;
    qc.horizon = t;
    qc.collect_stats = 0;
    gpq_phold_config pc = *model;
    pc.end_time = t;
    pc.drain = 0;
    pc.checks = 0;
    double rate = 0.0;
    uint64_t hash = 0;
    uint32_t status = 1;
    if (time_candidate(qc, pc, &rate, &hash)) {
      status = 0;
      if (!have_ref) {
        have_ref = true;
        ref_hash = hash;
      } else if (hash != ref_hash) {
        status = 2;
      }
    }
    if (status) rate = 0.0;
    if (results && n < max_results) results[n] = gpq_sort_tuning{tile, status, rate};
    ++n;
    if (rate > (pick ? best * kMargin : 0.0)) {
      best = rate;
      pick = tile;
    }
  }
  if (n_results) *n_results = n < max_results || !results ? n : max_results;
  if (!pick) {
    set_error("sort autotune: no tile ran the workload");
    return GPQ_ERR_CAPACITY;
  }
  queue->sort_tile = pick;
  return GPQ_OK;
}

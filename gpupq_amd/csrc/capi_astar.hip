// capi_astar.hip -- the extern "C" boundary of the A* models (include/gpupq.h).
#include <string>

#include "phold_batch.h"
#include "serve.h"

using gpq::set_error;

#define GPQ_NONNULL(p)                                      \
  do {                                                      \
    if (!(p)) { set_error("null argument"); return GPQ_ERR_INVALID; } \
  } while (0)

extern "C" {

gpq_status gpq_astar_create(gpq_queue* open_lists, const gpq_astar_config* cfg, gpq_astar** out) {
  GPQ_NONNULL(open_lists);
  GPQ_NONNULL(open_lists->impl);
  GPQ_NONNULL(cfg);
  GPQ_NONNULL(out);
  gpq::AstarModel* m = nullptr;
  GPQ_TRY(gpq::AstarModel::create(open_lists->impl, *cfg, &m));
  *out = new gpq_astar{m};
  return GPQ_OK;
}

gpq_status gpq_astar_destroy(gpq_astar* a) {
  if (!a) return GPQ_OK;
  if (a->impl && a->impl->q) (void)hipStreamSynchronize(a->impl->q->stream);
  delete a->impl;
  delete a;
  return GPQ_OK;
}

gpq_status gpq_astar_init(gpq_astar* a) {
  GPQ_NONNULL(a);
  return a->impl->init();
}

gpq_status gpq_astar_set_grid(gpq_astar* a, const uint8_t* blocked) {
  GPQ_NONNULL(a);
  GPQ_NONNULL(blocked);
  return a->impl->set_grid(blocked);
}

gpq_status gpq_astar_handle_next(gpq_astar* a) {
  GPQ_NONNULL(a);
  return a->impl->step();
}

gpq_status gpq_astar_run(gpq_astar* a, gpq_drive_mode mode, uint32_t batch, gpq_astar_report* rep) {
  GPQ_NONNULL(a);
  if (mode != GPQ_DRIVE_PLAIN && mode != GPQ_DRIVE_GRAPH) { set_error("bad drive mode"); return GPQ_ERR_INVALID; }
  return a->impl->run(mode, batch, rep);
}

gpq_status gpq_astar_grid(gpq_astar* a, uint8_t* blocked) {
  GPQ_NONNULL(a);
  GPQ_NONNULL(blocked);
  return a->impl->grid_bytes(blocked);
}

gpq_status gpq_astar_path(gpq_astar* a, uint32_t* cells, uint32_t cap, uint32_t* len) {
  GPQ_NONNULL(a);
  GPQ_NONNULL(len);
  if (cap && !cells) { set_error("null cells"); return GPQ_ERR_INVALID; }
  return a->impl->path(cells, cap, len);
}

gpq_status gpq_query_synthetic(uint32_t seed, uint32_t q, gpq_query_spec* out) {
  GPQ_NONNULL(out);
  return gpq::query_synthetic(seed, q, out);
}

gpq_status gpq_astar_batch_create(const gpq_query_spec* specs, uint32_t count, const gpq_batch_tuning* tuning,
                                  gpq_astar_batch** out) {
  GPQ_NONNULL(specs);
  GPQ_NONNULL(out);
  gpq_batch_tuning t{};
  if (tuning) t = *tuning;
  gpq::AstarBatch* b = nullptr;
  GPQ_TRY(gpq::AstarBatch::create(specs, count, t, &b));
  *out = new gpq_astar_batch{b};
  return GPQ_OK;
}

gpq_status gpq_astar_batch_autotune(const gpq_query_spec* specs, uint32_t count, uint32_t sample,
                                    gpq_batch_tuning* tuning, double* best_seconds) {
  GPQ_NONNULL(specs);
  GPQ_NONNULL(tuning);
  gpq_batch_tuning base = *tuning;
  return gpq::batch_autotune(specs, count, sample, base, tuning, best_seconds);
}

gpq_status gpq_astar_batch_destroy(gpq_astar_batch* b) {
  if (!b) return GPQ_OK;
  delete b->impl;
  delete b;
  return GPQ_OK;
}

gpq_status gpq_astar_batch_run(gpq_astar_batch* b, gpq_drive_mode mode, double* seconds) {
  GPQ_NONNULL(b);
  if (mode != GPQ_DRIVE_PLAIN && mode != GPQ_DRIVE_GRAPH) { set_error("bad drive mode"); return GPQ_ERR_INVALID; }
  return b->impl->run(mode, seconds);
}

gpq_status gpq_astar_batch_results(gpq_astar_batch* b, int64_t* costs, uint64_t* expanded) {
  GPQ_NONNULL(b);
  return b->impl->results(costs, expanded);
}

gpq_status gpq_astar_batch_timing(gpq_astar_batch* b, double* start_us, double* finish_us) {
  GPQ_NONNULL(b);
  return b->impl->timing(start_us, finish_us);
}

gpq_status gpq_astar_batch_path(gpq_astar_batch* b, uint32_t q, uint32_t* cells, uint32_t cap, uint32_t* len) {
  GPQ_NONNULL(b);
  GPQ_NONNULL(len);
  return b->impl->path(q, cells, cap, len);
}

gpq_status gpq_astar_batch_grid(gpq_astar_batch* b, uint32_t q, uint8_t* blocked) {
  GPQ_NONNULL(b);
  GPQ_NONNULL(blocked);
  return b->impl->grid_bytes(q, blocked);
}

gpq_status gpq_phold_batch_create(const gpq_phold_instance* inst, uint32_t count, uint32_t capacity_per_lp,
                                  uint32_t stage_per_lp, uint32_t far_per_lp, gpq_phold_batch** out) {
  GPQ_NONNULL(inst);
  GPQ_NONNULL(out);
  gpq::PholdBatch* b = nullptr;
  GPQ_TRY(gpq::PholdBatch::create(inst, count, capacity_per_lp, stage_per_lp, far_per_lp, &b));
  *out = new gpq_phold_batch{b};
  return GPQ_OK;
}

gpq_status gpq_phold_batch_set_calendar(gpq_phold_batch* b, uint32_t calendar_windows) {
  GPQ_NONNULL(b);
  GPQ_CHECK_ARG(calendar_windows == 0 ||
                    ((calendar_windows & (calendar_windows - 1u)) == 0 && calendar_windows >= 8 &&
                     calendar_windows <= gpq::kCalMaxWindows),
                "phold batch: calendar_windows must be 0 or a power of two in [8, 256]");
  b->impl->calendar_windows = calendar_windows;
  return GPQ_OK;
}

gpq_status gpq_phold_batch_destroy(gpq_phold_batch* b) {
  if (!b) return GPQ_OK;
  delete b->impl;
  delete b;
  return GPQ_OK;
}

gpq_status gpq_phold_batch_run(gpq_phold_batch* b, gpq_drive_mode mode, uint32_t batch, double* seconds) {
  GPQ_NONNULL(b);
  if (mode != GPQ_DRIVE_PLAIN && mode != GPQ_DRIVE_GRAPH) { set_error("bad drive mode"); return GPQ_ERR_INVALID; }
  return b->impl->run(mode, batch, seconds);
}

gpq_status gpq_phold_batch_results(gpq_phold_batch* b, uint64_t* trace_hashes, uint64_t* events) {
  GPQ_NONNULL(b);
  return b->impl->results(trace_hashes, events);
}

gpq_status gpq_serve_trace(const gpq_serve_config* cfg, gpq_request* out) {
  GPQ_NONNULL(cfg);
  GPQ_NONNULL(out);
  return gpq::serve_trace(*cfg, out, nullptr);
}

gpq_status gpq_serve_run(const gpq_serve_config* cfg, gpq_drive_mode mode, gpq_serve_report* rep, gpq_request* out) {
  GPQ_NONNULL(cfg);
  if (mode != GPQ_DRIVE_PLAIN && mode != GPQ_DRIVE_GRAPH) { set_error("bad drive mode"); return GPQ_ERR_INVALID; }
  return gpq::serve_run(*cfg, mode, rep, out);
}

}  // extern "C"

// serve.h -- the serving runtime: a synthetic trace of mixed PHOLD / A* requests, served by
// continuous batching under an HBM budget, or one request at a time (plain).
#pragma once

#include "astar.h"
#include "phold_batch.h"

namespace gpq {

gpq_status serve_trace(const gpq_serve_config& c, gpq_request* out, uint32_t* phold_seeds);
gpq_status serve_run(const gpq_serve_config& c, gpq_drive_mode mode, gpq_serve_report* rep, gpq_request* out);

}  // namespace gpq

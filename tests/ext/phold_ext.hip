// phold_ext.hip -- a PHOLD model compiled OUTSIDE libgpupq, against the public device operator
// API (include/gpupq_device.h) only: the shape of a reference model (src/phold.cu:53-117,
// model.h:9-19) written for this library.  tests/test_device_api_gpu.py drives it through the
// host queue API and compares its trace with the CPU oracle.  It links nothing from libgpupq:
// the queue's view comes from gpq_queue_device_view and is passed to the kernels by value.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/gpupq_device.h"

using gpq::Event;

namespace {

struct Model {
  uint32_t num_lps, init_per_lp, seed, p_self;
  float lookahead;
  uint32_t* count;
  uint64_t* hash;
};

// phold_populate (src/phold.cu:63-79): initial event j of LP lp, at ts = Exp(1) of draw
// {j, lp, 1, 0}, owned and sourced by the LP itself.
template <class Q>
__global__ void ext_populate_kernel(Q q, Model m, uint32_t j) {
  const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= m.num_lps) return;
  const uint4 r = gpq::philox4x32_10(make_uint4(j, lp, 1u, 0u), make_uint2(m.seed, 0u));
  gpq::queue_insert(q, Event{gpq::f_to_bits(gpq::exp_from_u32(r.x)), lp, lp, j});
}

// handle_next_event (src/phold.cu:81-117): the slot's safe event is committed (counted and
// hashed), and its successor is scheduled at (ts + lookahead) + Exp(1), at the LP itself with
// probability p_self, else at a uniform LP.
template <class Q>
__global__ void ext_handle_next_kernel(Q q, Model m) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= gpq::queue_slots(q)) return;
  Event* item = nullptr;
  const int idx = gpq::queue_peek(q, &item, (int)slot);
  if (idx < 0) return;
  const Event e = *item;
  const uint32_t lp = e.dst;
  const uint32_t n = m.count[lp]++;
  m.hash[lp] = gpq::trace_step(m.hash[lp], e.ts, e.src, e.seq);
  const uint4 r = gpq::philox4x32_10(make_uint4(n, lp, 0u, 0u), make_uint2(m.seed, 0u));
  const uint32_t dst = r.y < m.p_self ? lp : (uint32_t)(((uint64_t)r.z * m.num_lps) >> 32);
  const float ts = (gpq::bits_to_f(e.ts) + m.lookahead) + gpq::exp_from_u32(r.x);
  gpq::queue_set_done(q, idx);
  gpq::queue_insert(q, Event{gpq::f_to_bits(ts), dst, lp, m.init_per_lp + n});
}

// Per-LP verbs of the local queues: length and emptiness of LPs lps[0..n), then clear of
// those with clear_flag set.
template <class Q>
__global__ void ext_lp_verbs_kernel(Q q, const uint32_t* lps, uint32_t n, int32_t* length, uint32_t* empty,
                                    const uint32_t* clear_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int lp = (int)lps[i];
  length[i] = gpq::queue_length(q, lp);
  empty[i] = gpq::queue_is_empty(q, lp) ? 1u : 0u;
  if (clear_flag[i]) gpq::queue_clear(q, lp);
}

template <class F>
int dispatch(int kind, const void* view, F&& f) {
  switch (kind) {
    case GPQ_LAQ: return f(*static_cast<const gpq::LaqView*>(view));
    case GPQ_LSQ: return f(*static_cast<const gpq::LsqView*>(view));
    case GPQ_LHQ: return f(*static_cast<const gpq::LhqView*>(view));
    case GPQ_HQ: return f(*static_cast<const gpq::HqView*>(view));
    case GPQ_GAQ: return f(*static_cast<const gpq::GaqView*>(view));
    default: return -1;
  }
}

template <class F>
int dispatch_local(int kind, const void* view, F&& f) {
  switch (kind) {
    case GPQ_LAQ: return f(*static_cast<const gpq::LaqView*>(view));
    case GPQ_LSQ: return f(*static_cast<const gpq::LsqView*>(view));
    case GPQ_LHQ: return f(*static_cast<const gpq::LhqView*>(view));
    default: return -1;
  }
}

}  // namespace

struct ext_phold {
  Model m;
};

extern "C" {

// Largest view struct of any kind, so callers can size the buffer they hand to
// gpq_queue_device_view, and the size of the view of each kind.
uint64_t ext_view_bytes(int kind) {
  switch (kind) {
    case GPQ_LAQ: return sizeof(gpq::LaqView);
    case GPQ_LSQ: return sizeof(gpq::LsqView);
    case GPQ_LHQ: return sizeof(gpq::LhqView);
    case GPQ_HQ: return sizeof(gpq::HqView);
    case GPQ_GAQ: return sizeof(gpq::GaqView);
    default: return 0;
  }
}

int ext_phold_create(uint32_t num_lps, uint32_t init_per_lp, uint32_t seed, float lookahead, uint32_t p_self,
                     ext_phold** out) {
  auto* p = new ext_phold{};
  p->m = Model{num_lps, init_per_lp, seed, p_self, lookahead, nullptr, nullptr};
  if (hipMalloc(&p->m.count, sizeof(uint32_t) * num_lps) != hipSuccess ||
      hipMalloc(&p->m.hash, sizeof(uint64_t) * num_lps) != hipSuccess ||
      hipMemset(p->m.count, 0, sizeof(uint32_t) * num_lps) != hipSuccess ||
      hipMemset(p->m.hash, 0, sizeof(uint64_t) * num_lps) != hipSuccess) {
    (void)hipFree(p->m.count);
    (void)hipFree(p->m.hash);
    delete p;
    return -1;
  }
  *out = p;
  return 0;
}

void ext_phold_destroy(ext_phold* p) {
  if (!p) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(p->m.count);
  (void)hipFree(p->m.hash);
  delete p;
}

int ext_phold_populate(ext_phold* p, int kind, const void* view, uint32_t j, void* stream) {
  const uint32_t grid = (p->m.num_lps + 255) / 256;
  return dispatch(kind, view, [&](const auto& q) {
    hipLaunchKernelGGL(ext_populate_kernel<std::decay_t<decltype(q)>>, dim3(grid), dim3(256), 0,
                       static_cast<hipStream_t>(stream), q, p->m, j);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  });
}

int ext_phold_handle_next(ext_phold* p, int kind, const void* view, void* stream) {
  const uint32_t grid = (p->m.num_lps + 255) / 256;
  return dispatch(kind, view, [&](const auto& q) {
    hipLaunchKernelGGL(ext_handle_next_kernel<std::decay_t<decltype(q)>>, dim3(grid), dim3(256), 0,
                       static_cast<hipStream_t>(stream), q, p->m);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  });
}

int ext_phold_state(ext_phold* p, uint32_t* counts, uint64_t* hashes) {
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  if (hipMemcpy(counts, p->m.count, sizeof(uint32_t) * p->m.num_lps, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  if (hipMemcpy(hashes, p->m.hash, sizeof(uint64_t) * p->m.num_lps, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  return 0;
}

// queue_length / queue_is_empty / queue_clear on n LPs (per-LP queues); synchronous.
int ext_lp_verbs(int kind, const void* view, const uint32_t* lps, uint32_t n, const uint32_t* clear_flag,
                 int32_t* length, uint32_t* empty, void* stream) {
  uint32_t *d_lps = nullptr, *d_empty = nullptr, *d_clear = nullptr;
  int32_t* d_len = nullptr;
  int rc = -2;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hipMalloc(&d_lps, 4 * n) == hipSuccess && hipMalloc(&d_empty, 4 * n) == hipSuccess &&
      hipMalloc(&d_clear, 4 * n) == hipSuccess && hipMalloc(&d_len, 4 * n) == hipSuccess &&
      hipMemcpyAsync(d_lps, lps, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(d_clear, clear_flag, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess) {
    rc = dispatch_local(kind, view, [&](const auto& q) {
      hipLaunchKernelGGL(ext_lp_verbs_kernel<std::decay_t<decltype(q)>>, dim3((n + 63) / 64), dim3(64), 0, s, q,
                         d_lps, n, d_len, d_empty, d_clear);
      return hipGetLastError() == hipSuccess ? 0 : -2;
    });
    if (rc == 0 && (hipMemcpyAsync(length, d_len, 4 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipMemcpyAsync(empty, d_empty, 4 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess))
      rc = -2;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_lps);
  (void)hipFree(d_empty);
  (void)hipFree(d_clear);
  (void)hipFree(d_len);
  return rc;
}

}  // extern "C"

// gpq_host.h -- host-side internals: error plumbing, device buffers, the queue
// base class every queue kind derives from, and the window state shared with
// the model kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gpupq.h"
#include "gpq_device.h"

namespace gpq {

void set_error(const std::string& msg);
const std::string& last_error();

#define GPQ_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::gpq::set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
      return GPQ_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define GPQ_TRY(expr)                                                                   \
  do {                                                                                  \
    gpq_status s_ = (expr);                                                             \
    if (s_ != GPQ_OK) return s_;                                                        \
  } while (0)

#define GPQ_CHECK_ARG(cond, msg)                                                        \
  do {                                                                                  \
    if (!(cond)) { ::gpq::set_error(msg); return GPQ_ERR_INVALID; }                     \
  } while (0)

// Device allocation tracked by its owner; freed in the owner's destructor.
class DeviceArena {
 public:
  ~DeviceArena() { release(); }
  template <class T>
  gpq_status alloc(T** p, size_t count) {
    void* raw = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    if (hipMalloc(&raw, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
      return GPQ_ERR_NOMEM;
    }
    ptrs_.push_back(raw);
    bytes_ += bytes;
    *p = static_cast<T*>(raw);
    return GPQ_OK;
  }
  size_t bytes() const { return bytes_; }
  void release() {
    for (void* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
    bytes_ = 0;
  }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_ = 0;
};

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline uint32_t log2u(uint64_t x) { uint32_t r = 0; while ((1ull << r) < x) ++r; return r; }

// Common part of every queue handle.  The window (gmin, safe-window end, done flag)
// lives in device memory and is published by the queue's post kernels.
class QueueBase {
 public:
  explicit QueueBase(const gpq_queue_config& c) : cfg(c) {}
  virtual ~QueueBase() {
    if (stream) (void)hipStreamDestroy(stream);
    if (host_ws) (void)hipHostFree(host_ws);
  }
  gpq_status init_common();

  virtual gpq_status init() = 0;
  virtual gpq_status pre() { return GPQ_OK; }
  virtual gpq_status post() = 0;
  virtual gpq_status post_init() { return post(); }
  virtual gpq_status lp_contents(uint32_t lp, Event* out, uint32_t cap, uint32_t* n) = 0;
  virtual gpq_status pending(uint64_t* n) = 0;
  virtual gpq_status length(uint32_t lp, uint32_t* n) = 0;
  virtual gpq_status clear_lp(uint32_t lp) = 0;
  // Number of queue_peek() slots the model iterates over.
  virtual uint32_t peek_slots() const { return cfg.num_lps; }

  // Folds partial[0 .. partial_slots) into the window state (one small kernel on stream).
  gpq_status publish_window();
  gpq_status window(WindowState* out);   // synchronous copy of the device window state
  gpq_status error_flags(uint32_t* flags);
  gpq_status stats_get(gpq_queue_stats* out);
  // *out = sum of a device u32 array (synchronous); tmp is one device u64.
  gpq_status sum_u32(const uint32_t* d_values, uint32_t count, uint64_t* tmp, uint64_t* out);

  gpq_queue_config cfg;
  hipStream_t stream = nullptr;
  DeviceArena arena;
  WindowState* ws = nullptr;          // device
  WindowState* host_ws = nullptr;     // pinned mirror for polling
  hipEvent_t profile_mark = nullptr;  // when set, recorded just before the window publication
  uint32_t* partial = nullptr;        // device, one slot per reduction block
  uint32_t partial_slots = 0;
  uint32_t threads = 256;
  // Work counters (device, kStatSlots u64) when cfg.collect_stats; nullptr otherwise.
  // Slots: 0 inserted, 1 displaced, 2 merged LPs, 3 scanned LPs, 4 aux bytes, 5 records
  // sorted, 6 records merged, 7 records scanned by the safe-event selection.
  unsigned long long* stats = nullptr;
  unsigned long long* acct(int slot) const { return stats ? stats + slot : nullptr; }
};

constexpr int kStatSlots = 8;
constexpr int kStatSorted = 5, kStatMerged = 6, kStatSelected = 7;

}  // namespace gpq

struct gpq_queue {
  gpq::QueueBase* impl;
};

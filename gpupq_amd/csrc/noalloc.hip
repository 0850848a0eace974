// noalloc.hip -- the record sort for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// gpq_sort_events_noalloc sorts caller-owned device records with the same kernels as
// gpq_selftest_sort (one bitonic tile sort, then merge-path passes), but its ping-pong
// partner buffer and its device-side record count live in `scratch`, a file-scope device
// array of kScratchBytes that every GPU the library runs on holds until the process ends.
// The array's address on the current device is looked up once per device with
// hipGetSymbolAddress and kept; host code only ever passes that address to kernels.
//
// Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (on
// different streams, or from different host threads) race on it and must be serialised by
// the caller.  The array keeps the last call's contents; the count word is rewritten by each
// call and the partner buffer is written before it is read, so nothing needs clearing.
#include <array>
#include <mutex>

#include "sortprim.h"

namespace gpq {
namespace {

constexpr uint32_t kNoAllocMaxRecords = 1u << 20;       // 16 MiB of partner buffer
constexpr size_t kScratchHeader = 256;                  // the record count, padded to 256 B
constexpr size_t kScratchBytes = kScratchHeader + sizeof(Event) * (size_t)kNoAllocMaxRecords;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::array<unsigned char*, kMaxDevices> g_scratch_ptr{};   // per device, from hipGetSymbolAddress

// The current device's address of `scratch`, looked up on the first call on that device.
gpq_status scratch_on_current_device(unsigned char** out) {
  int dev = -1;
  GPQ_HIP(hipGetDevice(&dev));
  GPQ_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GPQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GPQ_CHECK_ARG(p != nullptr, "noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return GPQ_OK;
}

__global__ void noalloc_set_count_kernel(uint32_t* count, uint32_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = n;
}

}  // namespace
}  // namespace gpq

extern "C" gpq_status gpq_sort_events_noalloc(const gpq_event* d_in, uint32_t n, gpq_event* d_out, void* stream) {
  using gpq::Event;
  if (n == 0) return GPQ_OK;
  if (!d_in || !d_out) {
    gpq::set_error("sort_events_noalloc: null record pointer");
    return GPQ_ERR_INVALID;
  }
  if (n > gpq::kNoAllocMaxRecords) {
    gpq::set_error("sort_events_noalloc: " + std::to_string(n) + " records exceed the scratch limit of " +
                   std::to_string(gpq::kNoAllocMaxRecords));
    return GPQ_ERR_CAPACITY;
  }
  const uintptr_t in0 = reinterpret_cast<uintptr_t>(d_in), out0 = reinterpret_cast<uintptr_t>(d_out);
  const uintptr_t bytes = sizeof(Event) * (uintptr_t)n;
  if (in0 != out0 && in0 < out0 + bytes && out0 < in0 + bytes) {
    gpq::set_error("sort_events_noalloc: input and output overlap without being the same array");
    return GPQ_ERR_INVALID;
  }
  if ((in0 | out0) % alignof(Event) != 0) {
    gpq::set_error("sort_events_noalloc: records must be 16-byte aligned");
    return GPQ_ERR_INVALID;
  }
  unsigned char* base = nullptr;
  GPQ_TRY(gpq::scratch_on_current_device(&base));
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* d_n = reinterpret_cast<uint32_t*>(base);
  Event* partner = reinterpret_cast<Event*>(base + gpq::kScratchHeader);
  hipLaunchKernelGGL(gpq::noalloc_set_count_kernel, dim3(1), dim3(64), 0, s, d_n, n);
  GPQ_HIP(hipGetLastError());
  // The tile sort writes its first buffer and each merge pass swaps the two; with the pass
  // count's parity the sorted run ends in the caller's output array.
  uint32_t passes = 0;
  for (uint64_t w = gpq::kSortTile; w < n; w *= 2) ++passes;
  Event* out = reinterpret_cast<Event*>(d_out);
  Event* first = (passes & 1u) ? partner : out;
  Event* second = (passes & 1u) ? out : partner;
  Event* sorted = nullptr;
  GPQ_TRY(gpq::sort_records(reinterpret_cast<const Event*>(d_in), first, second, d_n, n, s, &sorted));
  if (sorted != out) {
    gpq::set_error("sort_events_noalloc: sorted run left in the scratch buffer");
    return GPQ_ERR_STATE;
  }
  return GPQ_OK;
}

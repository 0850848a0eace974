// selftest.hip -- bulk evaluation of the device random-number path (Philox4x32-10
// and the Exp(1) sampler) so tests can compare it bit-for-bit with the CPU reference.
#include <cstring>
#include <vector>

#include "gpq_host.h"

namespace gpq {
namespace {
__global__ void exp_table_kernel(uint32_t first, uint32_t count, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = exp_from_u32((first + i) << 8);
}
__global__ void philox_table_kernel(uint4 c, uint2 k, uint32_t count, uint4* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = philox4x32_10(make_uint4(c.x + i, c.y, c.z, c.w), k);
}
}  // namespace
}  // namespace gpq

extern "C" gpq_status gpq_selftest_exp(uint32_t first, uint32_t count, float* out) {
  if (!out || count == 0 || count > (1u << 24)) { gpq::set_error("bad selftest_exp args"); return GPQ_ERR_INVALID; }
  float* d = nullptr;
  GPQ_HIP(hipMalloc(&d, sizeof(float) * count));
  hipLaunchKernelGGL(gpq::exp_table_kernel, dim3(gpq::ceil_div(count, 256)), dim3(256), 0, 0, first, count, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(float) * count, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  GPQ_HIP(e);
  return GPQ_OK;
}

extern "C" gpq_status gpq_selftest_philox(uint32_t ctr0, uint32_t ctr1, uint32_t ctr2, uint32_t ctr3, uint32_t key0,
                                          uint32_t key1, uint32_t count, uint32_t* out) {
  if (!out || count == 0 || count > (1u << 24)) { gpq::set_error("bad selftest_philox args"); return GPQ_ERR_INVALID; }
  uint4* d = nullptr;
  GPQ_HIP(hipMalloc(&d, sizeof(uint4) * count));
  hipLaunchKernelGGL(gpq::philox_table_kernel, dim3(gpq::ceil_div(count, 256)), dim3(256), 0, 0,
                     make_uint4(ctr0, ctr1, ctr2, ctr3), make_uint2(key0, key1), count, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(uint4) * count, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  GPQ_HIP(e);
  return GPQ_OK;
}

// ---- record primitives of the global queues, driven from host arrays ------------------
#include "sortprim.h"

namespace {
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" gpq_status gpq_selftest_sort(const gpq_event* in, uint32_t n, gpq_event* out) {
  using gpq::Event;
  if (!in || !out || n == 0 || n >= (1u << 30)) { gpq::set_error("bad selftest_sort args"); return GPQ_ERR_INVALID; }
  DevBuf a, b0, b1, dn;
  GPQ_HIP(hipMalloc(&a.p, sizeof(Event) * n));
  GPQ_HIP(hipMalloc(&b0.p, sizeof(Event) * n));
  GPQ_HIP(hipMalloc(&b1.p, sizeof(Event) * n));
  GPQ_HIP(hipMalloc(&dn.p, sizeof(uint32_t)));
  GPQ_HIP(hipMemcpy(a.p, in, sizeof(Event) * n, hipMemcpyHostToDevice));
  GPQ_HIP(hipMemcpy(dn.p, &n, sizeof(uint32_t), hipMemcpyHostToDevice));
  // A capacity above n, as the queues pass it: the merge passes whose width already covers n
  // are skipped or reduced to one copy, by the parity of their number.
  const uint32_t cap = n <= (1u << 28) ? 6u * n : n;
  Event* res = nullptr;
  GPQ_TRY(gpq::sort_records(static_cast<Event*>(a.p), static_cast<Event*>(b0.p), static_cast<Event*>(b1.p),
                            static_cast<uint32_t*>(dn.p), cap, 0, &res));
  GPQ_HIP(hipMemcpy(out, res, sizeof(Event) * n, hipMemcpyDeviceToHost));
  return GPQ_OK;
}

extern "C" gpq_status gpq_selftest_merge(const gpq_event* a, uint32_t na, const gpq_event* b, uint32_t nb,
                                         gpq_event* out) {
  using gpq::Event;
  const uint64_t total = (uint64_t)na + nb;
  if (!out || total == 0 || total >= (1u << 30) || (na && !a) || (nb && !b)) {
    gpq::set_error("bad selftest_merge args");
    return GPQ_ERR_INVALID;
  }
  DevBuf da, db, dout, dt;
  GPQ_HIP(hipMalloc(&da.p, sizeof(Event) * (na ? na : 1)));
  GPQ_HIP(hipMalloc(&db.p, sizeof(Event) * (nb ? nb : 1)));
  GPQ_HIP(hipMalloc(&dout.p, sizeof(Event) * total));
  GPQ_HIP(hipMalloc(&dt.p, sizeof(gpq::MergeTask)));
  if (na) GPQ_HIP(hipMemcpy(da.p, a, sizeof(Event) * na, hipMemcpyHostToDevice));
  if (nb) GPQ_HIP(hipMemcpy(db.p, b, sizeof(Event) * nb, hipMemcpyHostToDevice));
  // A is split in two parts to exercise the concatenated first input.
  gpq::MergeTask t{};
  const uint32_t h = na / 3;
  t.a1 = static_cast<Event*>(da.p);
  t.n_a1 = h;
  t.a2 = static_cast<Event*>(da.p) + h;
  t.n_a2 = na - h;
  t.b = static_cast<Event*>(db.p);
  t.n_b = nb;
  // output split over three segments of the same buffer to exercise the segment map
  const uint32_t s1 = (uint32_t)(total / 4), s2 = (uint32_t)(total / 2);
  t.out[0] = static_cast<Event*>(dout.p);
  t.out[1] = static_cast<Event*>(dout.p) + s1;
  t.out[2] = static_cast<Event*>(dout.p) + s2;
  t.end[0] = s1;
  t.end[1] = s2;
  t.end[2] = (uint32_t)total;
  GPQ_HIP(hipMemcpy(dt.p, &t, sizeof(t), hipMemcpyHostToDevice));
  GPQ_TRY(gpq::merge_tasks(static_cast<gpq::MergeTask*>(dt.p), 1, total, 0));
  GPQ_HIP(hipMemcpy(out, dout.p, sizeof(Event) * total, hipMemcpyDeviceToHost));
  // the same merge with its splits computed by the partition launch, for every tile size:
  // it must give the same records
  std::vector<Event> again(total);
  for (uint32_t tile : {1024u, 2048u, 4096u}) {
    const uint32_t stride = gpq::merge_split_words(total, tile);
    DevBuf dsplit;
    GPQ_HIP(hipMalloc(&dsplit.p, sizeof(uint32_t) * stride));
    GPQ_HIP(hipMemset(dout.p, 0, sizeof(Event) * total));
    GPQ_TRY(gpq::merge_tasks(static_cast<gpq::MergeTask*>(dt.p), 1, total, 0, tile, nullptr,
                             static_cast<uint32_t*>(dsplit.p), stride, 0));
    GPQ_HIP(hipMemcpy(again.data(), dout.p, sizeof(Event) * total, hipMemcpyDeviceToHost));
    if (std::memcmp(again.data(), out, sizeof(Event) * total) != 0) {
      gpq::set_error("merge with precomputed splits differs (tile " + std::to_string(tile) + ")");
      return GPQ_ERR_STATE;
    }
  }
  return GPQ_OK;
}

extern "C" gpq_status gpq_selftest_select(const gpq_event* sorted, uint32_t n, uint32_t num_lps, float wend,
                                          const gpq_event* bound, gpq_event* deq, uint32_t* ndeq, gpq_event* rest,
                                          uint32_t* nrest, uint32_t* prefix) {
  using gpq::Event;
  if (!sorted || !deq || !rest || !ndeq || !nrest || !prefix || n == 0 || num_lps == 0) {
    gpq::set_error("bad selftest_select args");
    return GPQ_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n; ++i)
    if (sorted[i].dst >= num_lps) { gpq::set_error("dst out of range"); return GPQ_ERR_INVALID; }
  DevBuf ds, dsp, dfirst, dscr, dvals, dbound, ddeq, drest;
  const uint32_t nblk = (n + 1023) / 1024;
  GPQ_HIP(hipMalloc(&ds.p, sizeof(Event) * n));
  GPQ_HIP(hipMalloc(&dsp.p, sizeof(Event*)));
  GPQ_HIP(hipMalloc(&dfirst.p, sizeof(uint32_t) * num_lps));
  GPQ_HIP(hipMalloc(&dscr.p, sizeof(uint32_t) * (2 * nblk + 2)));
  GPQ_HIP(hipMalloc(&dvals.p, sizeof(uint32_t) * 8));
  GPQ_HIP(hipMalloc(&dbound.p, sizeof(Event)));
  GPQ_HIP(hipMalloc(&ddeq.p, sizeof(Event) * n));
  GPQ_HIP(hipMalloc(&drest.p, sizeof(Event) * n));
  GPQ_HIP(hipMemcpy(ds.p, sorted, sizeof(Event) * n, hipMemcpyHostToDevice));
  GPQ_HIP(hipMemcpy(dsp.p, &ds.p, sizeof(Event*), hipMemcpyHostToDevice));
  GPQ_HIP(hipMemset(dfirst.p, 0xFF, sizeof(uint32_t) * num_lps));
  uint32_t vals[8] = {n, 0, 0, 0, 0, bound ? 1u : 0u, 0, 0};
  std::memcpy(&vals[1], &wend, sizeof(float));
  GPQ_HIP(hipMemcpy(dvals.p, vals, sizeof(vals), hipMemcpyHostToDevice));
  if (bound) GPQ_HIP(hipMemcpy(dbound.p, bound, sizeof(Event), hipMemcpyHostToDevice));
  uint32_t* v = static_cast<uint32_t*>(dvals.p);
  const gpq::SelectBufs b{static_cast<uint32_t*>(dfirst.p), static_cast<uint32_t*>(dscr.p), v + 2, v + 3, v + 4};
  GPQ_TRY(gpq::select_safe(static_cast<const Event* const*>(dsp.p), v, n, v + 1, static_cast<Event*>(dbound.p),
                           v + 5, b, static_cast<Event*>(ddeq.p), static_cast<Event*>(drest.p), 0));
  GPQ_HIP(hipMemcpy(vals, dvals.p, sizeof(vals), hipMemcpyDeviceToHost));
  *prefix = vals[2];
  *ndeq = vals[3];
  *nrest = vals[4];
  if (vals[3]) GPQ_HIP(hipMemcpy(deq, ddeq.p, sizeof(Event) * vals[3], hipMemcpyDeviceToHost));
  if (vals[4]) GPQ_HIP(hipMemcpy(rest, drest.p, sizeof(Event) * vals[4], hipMemcpyDeviceToHost));
  // the per-LP workspace must be back to all-ones
  std::vector<uint32_t> first(num_lps);
  GPQ_HIP(hipMemcpy(first.data(), dfirst.p, sizeof(uint32_t) * num_lps, hipMemcpyDeviceToHost));
  for (uint32_t x : first)
    if (x != 0xFFFFFFFFu) { gpq::set_error("selection left per-LP workspace dirty"); return GPQ_ERR_STATE; }
  return GPQ_OK;
}

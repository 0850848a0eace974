// selftest.hip -- bulk evaluation of the device random-number path (Philox4x32-10
// and the Exp(1) sampler) so tests can compare it bit-for-bit with the CPU reference.
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

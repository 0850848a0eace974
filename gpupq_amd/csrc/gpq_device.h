// gpq_device.h -- the public device vocabulary (include/gpupq_device/core.cuh) plus the
// helpers only the library's own kernels use: the block-wide window minimum and the
// kernel that publishes it.
#pragma once

#include "../../include/gpupq_device/core.cuh"

namespace gpq {

// Block-wide min of one value per thread, written to partial[blockIdx.x].  No fence:
// a device-scope fence on gfx950 writes back the XCD's L2 (buffer_wbl2), far too
// costly per block; the partials are folded by window_publish_kernel, launched
// after the producing kernel on the same stream.  blockDim.x: multiple of 64, <= 1024.
__device__ __forceinline__ void block_min_to_partial(uint32_t v, uint32_t* partial) {
  __shared__ uint32_t s_wmin[16];
  v = wave_min_u32(v);
  if (lane_id() == 0) s_wmin[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = s_wmin[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) m = s_wmin[i] < m ? s_wmin[i] : m;
    partial[blockIdx.x] = m;
  }
}

// One block: folds the per-block minima and publishes gmin, the window end
// min(gmin + lookahead, horizon) and the done flag.
__global__ void window_publish_kernel(const uint32_t* partial, uint32_t count, WindowState* ws, float lookahead,
                                      float horizon);

}  // namespace gpq

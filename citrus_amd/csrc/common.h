// Shared helpers of the citrus HIP library: per-thread error message and the failure helper; the first-collision
// direction cosine that the first-collision kernels and the sweep (applying the sampling correction) share.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace citrus {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// eta = Omega_UC . w in float32, with the fused multiply-adds written out so that every kernel that forms it (the
// first-collision source, sums, moments and correction, and the sweep applying the correction to stored samples)
// gets the same value bit for bit
__host__ __device__ __forceinline__ float direction_cosine(float ux, float uy, float uz, float ox, float oy, float oz) {
  return fmaf(uz, oz, fmaf(uy, oy, ux * ox));
}

}  // namespace citrus

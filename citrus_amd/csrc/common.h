// Shared helpers of the citrus HIP library: per-thread error message and the failure helper; the first-collision
// direction cosine that the first-collision kernels and the sweep (applying the sampling correction) share; the
// attenuation factor of the image kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace citrus {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// A dispatch counts the work-items of each grid dimension in 32 bits: a 1-D grid of 2^32 or more work-items is cut
// short without an error. Launchers check their grids with grid_fits, and elementwise kernels over sizes that can
// pass the limit loop with a stride over at most kStrideBlocks workgroups.
constexpr int64_t kMaxGridItems = 0xffffffffLL;
constexpr int64_t kStrideBlocks = 1 << 20;

inline bool grid_fits(int64_t blocks, int64_t threads) { return blocks >= 0 && blocks * threads <= kMaxGridItems; }

inline unsigned stride_grid(int64_t total, int threads) {
  const int64_t blocks = (total + threads - 1) / threads;
  return (unsigned)(blocks < kStrideBlocks ? (blocks > 0 ? blocks : 1) : kStrideBlocks);
}

// eta = Omega_UC . w in float32, with the fused multiply-adds written out so that every kernel that forms it (the
// first-collision source, sums, moments and correction, and the sweep applying the correction to stored samples)
// gets the same value bit for bit
__host__ __device__ __forceinline__ float direction_cosine(float ux, float uy, float uz, float ox, float oy, float oz) {
  return fmaf(uz, oz, fmaf(uy, oy, ux * ox));
}

// exp(-t) for t >= 0 (+inf gives 0) as the hardware base-2 exponential of -t log2(e): relative error about
// |t| 6e-8 + 1 ulp, against the several instructions of expf's range handling per call. The attenuation factors of
// the collapsed-cone image (optical depths, never negative) use it.
__device__ __forceinline__ float attenuation(float t) { return __builtin_amdgcn_exp2f(-t * 1.44269504088896341f); }

}  // namespace citrus

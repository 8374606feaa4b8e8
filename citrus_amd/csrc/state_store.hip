// Compact per-view accumulators for gfx950: the LD cell mean of every (view, group, moment) in float16, scaled by the
// (group, moment)'s largest magnitude over 2^14, and back (the reference's StateStore.put/get,
// citrus/citrus.py:623-737: scale = max|m| / 2^14 (1 where the maximum is 0), q = float16(m / scale), m' = q scale).
//
// Moments arrive in the sweep's group-major layout [V][G][N][nsh][4] (the cell mean is component 0, every 16 bytes);
// the stored form is [V][G][nsh][N] float16 (the reference's stored shape [G, nsh, 1, Rc, Cc, Sc]) plus [V][G][nsh]
// float32 scales. Packing is three passes: per-block maxima of every moment (one thread per cell, 36 running maxima
// in registers, a wave-then-block reduction through LDS), the maxima over blocks and the scales, then the scaled
// conversion (a thread per cell reads its cell's moments and writes each moment's float16 row coalesced). The maximum
// is exact whatever the order, so the packed state is deterministic.

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "citrus_hip.h"
#include "common.h"

// Scratch of the allocation-free pack (citrus_state_pack_static): the per-block maxima live in this file-scope device
// array instead of a caller workspace. Every GPU the library runs on holds its own copy until the process ends; host
// code never passes the array's name to a kernel (on the host it is a stand-in, not a device address) but the address
// hipGetSymbolAddress returns for the current device, looked up once per device.
static __device__ __align__(256) float scratch[CITRUS_STATE_PACK_SCRATCH_FLOATS];

namespace citrus {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxMoments = 64;
constexpr float kFloat16Target = 16384.0f;  // the per-(group, moment) maximum maps to 2^14 (< 65504)

int blocks_per_slab(int64_t N) { return (int)((N + kThreads - 1) / kThreads); }

template <int NSH>
__global__ __launch_bounds__(kThreads) void state_max_kernel(const float* __restrict__ phi, int64_t N,
                                                             float* __restrict__ partial) {
  __shared__ float red[kThreads / 64][NSH];
  const int64_t slab = blockIdx.y;  // (view, group)
  const int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  float m[NSH];
#pragma unroll
  for (int n = 0; n < NSH; ++n) m[n] = 0.0f;
  if (c < N) {
    const float* p = phi + (slab * N + c) * NSH * 4;
#pragma unroll
    for (int n = 0; n < NSH; ++n) m[n] = fabsf(p[4 * n]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < NSH; ++n) {
    float v = m[n];
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if (lane == 0) red[wave][n] = v;
  }
  __syncthreads();
  if (threadIdx.x < NSH) {
    float v = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) v = fmaxf(v, red[w][threadIdx.x]);
    partial[(slab * gridDim.x + blockIdx.x) * NSH + threadIdx.x] = v;
  }
}

__global__ void state_scale_kernel(const float* __restrict__ partial, int nblk, int nsh, int64_t slabs,
                                   float* __restrict__ scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (slab, moment)
  if (i >= slabs * nsh) return;
  const int64_t slab = i / nsh;
  const int n = (int)(i % nsh);
  float m = 0.0f;
  for (int b = 0; b < nblk; ++b) m = fmaxf(m, partial[(slab * nblk + b) * nsh + n]);
  scale[i] = m > 0.0f ? m / kFloat16Target : 1.0f;
}

template <int NSH>
__global__ __launch_bounds__(kThreads) void state_pack_kernel(const float* __restrict__ phi, int64_t N,
                                                              const float* __restrict__ scale,
                                                              __half* __restrict__ q) {
  const int64_t slab = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (c >= N) return;
  const float* p = phi + (slab * N + c) * NSH * 4;
  const float* s = scale + slab * NSH;
  __half* o = q + slab * NSH * N + c;
#pragma unroll
  for (int n = 0; n < NSH; ++n) o[(int64_t)n * N] = __float2half_rn(p[4 * n] / s[n]);
}

template <int NSH>
__global__ __launch_bounds__(kThreads) void state_unpack_kernel(const __half* __restrict__ q,
                                                                const float* __restrict__ scale, int64_t N,
                                                                float* __restrict__ phi) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int64_t slab = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (c >= N) return;
  const __half* iq = q + slab * NSH * N + c;
  const float* s = scale + slab * NSH;
  f4* o = reinterpret_cast<f4*>(phi + (slab * N + c) * NSH * 4);
#pragma unroll
  for (int n = 0; n < NSH; ++n) o[n] = f4{__half2float(iq[(int64_t)n * N]) * s[n], 0.0f, 0.0f, 0.0f};
}

template <template <int> class K, class... A>
hipError_t launch_nsh(int nsh, dim3 grid, hipStream_t s, A... args) {
  switch (nsh) {
    case 1: hipLaunchKernelGGL(K<1>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 4: hipLaunchKernelGGL(K<4>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 9: hipLaunchKernelGGL(K<9>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 16: hipLaunchKernelGGL(K<16>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 25: hipLaunchKernelGGL(K<25>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 36: hipLaunchKernelGGL(K<36>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 49: hipLaunchKernelGGL(K<49>::fn, grid, dim3(kThreads), 0, s, args...); break;
    case 64: hipLaunchKernelGGL(K<64>::fn, grid, dim3(kThreads), 0, s, args...); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NSH>
struct MaxK { static constexpr auto fn = state_max_kernel<NSH>; };
template <int NSH>
struct PackK { static constexpr auto fn = state_pack_kernel<NSH>; };
template <int NSH>
struct UnpackK { static constexpr auto fn = state_unpack_kernel<NSH>; };

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
float* scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(CITRUS_ERR_UNSUPPORTED, "device index outside 0..63");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CITRUS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = scratch_ptr[dev];
  return CITRUS_OK;
}

int check(int V, int G, int nsh, int64_t N) {
  if (V < 1 || G < 1 || N < 1) return fail(CITRUS_ERR_ARG, "V, G and N must be >= 1");
  if (nsh < 1 || nsh > kMaxMoments) return fail(CITRUS_ERR_ARG, "nsh must be (L+1)^2 <= 64");
  if ((int64_t)V * G > 65535) return fail(CITRUS_ERR_ARG, "at most 65535 (view, group) slabs per call");
  if (!grid_fits(blocks_per_slab(N), kThreads)) return fail(CITRUS_ERR_ARG, "too many cells");
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

extern "C" int64_t citrus_state_pack_workspace_floats(int V, int G, int nsh, int64_t N) {
  if (V < 1 || G < 1 || nsh < 1 || N < 1) return -1;
  return (int64_t)V * G * citrus::blocks_per_slab(N) * nsh;
}

namespace citrus {
namespace {

// The three passes of the pack on `workspace` (>= citrus_state_pack_workspace_floats floats, fully written by the
// first pass before the second reads it, so it needs no clearing).
int pack(const float* phi, int V, int G, int nsh, int64_t N, float* scale, void* q, float* workspace,
         void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int nblk = blocks_per_slab(N);
  const int64_t slabs = (int64_t)V * G;
  const dim3 grid(nblk, (unsigned)slabs);
  hipError_t e = launch_nsh<MaxK>(nsh, grid, s, phi, N, workspace);
  if (e == hipSuccess) {
    const int64_t n = slabs * nsh;
    hipLaunchKernelGGL(state_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, workspace, nblk, nsh,
                       slabs, scale);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch_nsh<PackK>(nsh, grid, s, phi, N, (const float*)scale, static_cast<__half*>(q));
  if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("state pack: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_state_pack(const float* phi, int V, int G, int nsh, int64_t N, float* scale, void* q,
                                 float* workspace, int64_t workspace_floats, void* stream) {
  using namespace citrus;
  if (int rc = check(V, G, nsh, N)) return rc;
  if (!phi || !scale || !q || !workspace) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (workspace_floats < citrus_state_pack_workspace_floats(V, G, nsh, N))
    return fail(CITRUS_ERR_ARG, "workspace smaller than citrus_state_pack_workspace_floats");
  return pack(phi, V, G, nsh, N, scale, q, workspace, stream);
}

extern "C" int citrus_state_pack_static(const float* phi, int V, int G, int nsh, int64_t N, float* scale, void* q,
                                        void* stream) {
  using namespace citrus;
  if (int rc = check(V, G, nsh, N)) return rc;
  if (!phi || !scale || !q) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (citrus_state_pack_workspace_floats(V, G, nsh, N) > CITRUS_STATE_PACK_SCRATCH_FLOATS)
    return fail(CITRUS_ERR_ARG, "V * G * ceil(N / 256) * nsh exceeds CITRUS_STATE_PACK_SCRATCH_FLOATS (" +
                                    std::to_string((long long)CITRUS_STATE_PACK_SCRATCH_FLOATS) + " floats)");
  float* ws = nullptr;
  if (int rc = scratch_for_current_device(&ws)) return rc;
  return pack(phi, V, G, nsh, N, scale, q, ws, stream);
}

extern "C" int citrus_state_unpack(const void* q, const float* scale, int V, int G, int nsh, int64_t N, float* phi,
                                   void* stream) {
  using namespace citrus;
  if (int rc = check(V, G, nsh, N)) return rc;
  if (!phi || !scale || !q) return fail(CITRUS_ERR_ARG, "NULL buffer");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(blocks_per_slab(N), (unsigned)((int64_t)V * G));
  hipError_t e = launch_nsh<UnpackK>(nsh, grid, s, static_cast<const __half*>(q), scale, N, phi);
  if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("state unpack: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

// Footprint filter matrices of the collapsed-cone image (gfx950).
//
// box_matrix (citrus/scatter_solver/scatter_image.py:952-1007): B[c][j] = |[j, j+1] n W_c| / |W_c| with the window
// W_c = [c + 1/2 - w/2, c + 1/2 + w/2] n [0, N] of width w (clamped below at 1e-3) centred on pixel c, clipped to the
// panel and renormalised there. One thread per matrix entry; the blur itself is a batched GEMM of the images with
// these matrices, or, for the signal-only image, two banded kernels that sum each window directly
// (citrus_footprint_blur).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

// B[c][j] for a window of width w centred on pixel c of N
__device__ __forceinline__ float box_value(float w, int N, int c, int j) {
  const float half = 0.5f * fmaxf(w, 1e-3f);
  const float nf = (float)N;
  const float center = (float)c + 0.5f;
  const float lo = fminf(fmaxf(center - half, 0.0f), nf);
  const float hi = fminf(fmaxf(center + half, 0.0f), nf);
  const float jf = (float)j;
  return fmaxf(fminf(hi, jf + 1.0f) - fmaxf(lo, jf), 0.0f) / (hi - lo);
}

__global__ __launch_bounds__(256) void box_matrix_kernel(const float* __restrict__ w, int k, int N,
                                                         float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)k * N * N;
  if (idx >= total) return;
  const int j = (int)(idx % N);
  const int c = (int)((idx / N) % N);
  const int m = (int)(idx / ((int64_t)N * N));
  out[idx] = box_value(w[m], N, c, j);
}

// Emission-weighted mean panel height of each direction from the emission kernel's per-block sums
// partial [K][nb][2] = (sum e, sum e h): one workgroup per direction, a fixed-order tree over the blocks
// (scatter_image.py:1033-1060, the footprint's hbar_k); the box-centre height where the emission vanishes.
__global__ __launch_bounds__(256) void footprint_heights_kernel(const float* __restrict__ partial, int nb,
                                                                float center_height, float* __restrict__ heights) {
  __shared__ float s0[256], s1[256];
  const int k = blockIdx.x, t = threadIdx.x;
  float a = 0.0f, b = 0.0f;
  for (int i = t; i < nb; i += 256) {
    a += partial[((int64_t)k * nb + i) * 2];
    b += partial[((int64_t)k * nb + i) * 2 + 1];
  }
  s0[t] = a;
  s1[t] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s0[t] += s0[t + w];
      s1[t] += s1[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const float den = s0[0], num = s1[0];
    heights[k] = den > 0.0f ? num / fmaxf(den, 1.17549435e-38f) : center_height;
  }
}

// Both footprint filters of a chunk from the mean heights: widths (h cell_width[k][a]) scale_a in pixels; the
// channel filter transposed, channel_t[k][j][c] = B_c[c][j] w[k][c] (the right operand of the channel blur, its
// columns scaled by the optional channel weights w), the row filter rows[k][c][j] = B_r[c][j].
__global__ __launch_bounds__(256) void footprint_filters_kernel(const float* __restrict__ heights,
                                                                const float* __restrict__ cell_width, float sc,
                                                                float sr, int K, int Nc, int Nr,
                                                                const float* __restrict__ channel_weight,
                                                                float* __restrict__ channel_t,
                                                                float* __restrict__ rows) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc = (int64_t)K * Nc * Nc, nr = (int64_t)K * Nr * Nr;
  if (idx < nc) {
    const int c = (int)(idx % Nc);            // column of the transposed matrix: the window's pixel
    const int j = (int)((idx / Nc) % Nc);     // row: the pixel covered
    const int m = (int)(idx / ((int64_t)Nc * Nc));
    const float b = box_value(heights[m] * cell_width[2 * m] * sc, Nc, c, j);
    channel_t[idx] = channel_weight ? b * channel_weight[(int64_t)m * Nc + c] : b;
  } else if (idx < nc + nr) {
    const int64_t i = idx - nc;
    const int j = (int)(i % Nr);
    const int c = (int)((i / Nr) % Nr);
    const int m = (int)(i / ((int64_t)Nr * Nr));
    rows[i] = box_value(heights[m] * cell_width[2 * m + 1] * sr, Nr, c, j);
  }
}

// The window of box_value for output pixel c: [lo, hi] and 1 / (hi - lo); pixel j is covered by
// max(min(hi, j + 1) - max(lo, j), 0) of it.
struct BoxWindow {
  float lo, hi, inv;
  int j0, j1;  // pixels [j0, j1) with a nonzero share
};

__device__ __forceinline__ BoxWindow box_window(float w, int N, int c) {
  const float half = 0.5f * fmaxf(w, 1e-3f);
  const float nf = (float)N;
  const float center = (float)c + 0.5f;
  BoxWindow b;
  b.lo = fminf(fmaxf(center - half, 0.0f), nf);
  b.hi = fminf(fmaxf(center + half, 0.0f), nf);
  b.inv = 1.0f / (b.hi - b.lo);
  b.j0 = max((int)floorf(b.lo), 0);
  b.j1 = min((int)ceilf(b.hi), N);
  return b;
}

__device__ __forceinline__ float box_share(const BoxWindow& b, int j) {
  return fmaxf(fminf(b.hi, (float)j + 1.0f) - fmaxf(b.lo, (float)j), 0.0f);
}

// sum_j share_j v[j * stride] over the window's pixels, four independent partial sums (loads in flight together)
__device__ __forceinline__ float window_sum(const BoxWindow& b, const float* __restrict__ v, int64_t stride) {
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  int j = b.j0;
  for (; j + 4 <= b.j1; j += 4) {
    const float x0 = v[(int64_t)j * stride], x1 = v[(int64_t)(j + 1) * stride];
    const float x2 = v[(int64_t)(j + 2) * stride], x3 = v[(int64_t)(j + 3) * stride];
    a0 = fmaf(box_share(b, j), x0, a0);
    a1 = fmaf(box_share(b, j + 1), x1, a1);
    a2 = fmaf(box_share(b, j + 2), x2, a2);
    a3 = fmaf(box_share(b, j + 3), x3, a3);
  }
  for (; j < b.j1; ++j) a0 = fmaf(box_share(b, j), v[(int64_t)j * stride], a0);
  return (a0 + a1) + (a2 + a3);
}

// Channel footprint of every image row, banded: x[k][i][c] = chw[k][c] (sum_j share_j img[k][i][j]) / |W_c|, W_c the
// box of width heights[k] cell_width[k][0] sc around channel c (box_value's window). 16 x 16 tiles of (row, channel),
// k = blockIdx.z; a tile's rows are read through L1.
__global__ __launch_bounds__(256) void blur_channels_kernel(const float* __restrict__ img,
                                                            const float* __restrict__ heights,
                                                            const float* __restrict__ cell_width, float sc, int Nr,
                                                            int Nc, const float* __restrict__ channel_weight,
                                                            float* __restrict__ x) {
  const int c = blockIdx.x * 16 + (int)(threadIdx.x % 16), i = blockIdx.y * 16 + (int)(threadIdx.x / 16);
  const int k = blockIdx.z;
  if (c >= Nc || i >= Nr) return;
  const BoxWindow b = box_window(heights[k] * cell_width[2 * k] * sc, Nc, c);
  const float acc = window_sum(b, img + ((int64_t)k * Nr + i) * Nc, 1);
  const float wgt = channel_weight ? channel_weight[(int64_t)k * Nc + c] : 1.0f;
  x[((int64_t)k * Nr + i) * Nc + c] = acc * b.inv * wgt;
}

// Row footprint of the channel-blurred images and their sum over the chunk's directions, banded: acc[r][c] +=
// sum_k (sum_i share_i x[k][i][c]) / |W_r|, W_r the box of width heights[k] cell_width[k][1] sr around row r. A
// workgroup takes 64 pixels (4 rows x 16 channels) and splits the directions in four quarters over its threads; the
// quarters' sums are added in order through LDS, so each pixel is written once, with no atomics.
constexpr int kBlurSplit = 4;
__global__ __launch_bounds__(256) void blur_rows_accumulate_kernel(const float* __restrict__ x,
                                                                   const float* __restrict__ heights,
                                                                   const float* __restrict__ cell_width, float sr,
                                                                   int Nr, int Nc, int K, float* __restrict__ acc) {
  __shared__ float part[kBlurSplit][64];
  const int px = (int)(threadIdx.x % 64), part_id = (int)(threadIdx.x / 64);
  const int c = blockIdx.x * 16 + px % 16, r = blockIdx.y * 4 + px / 16;
  const bool on = c < Nc && r < Nr;
  const int kq = (K + kBlurSplit - 1) / kBlurSplit;
  const int k0 = part_id * kq, k1 = min(K, k0 + kq);
  float sum = 0.0f;
  if (on) {
    for (int k = k0; k < k1; ++k) {
      const BoxWindow b = box_window(heights[k] * cell_width[2 * k + 1] * sr, Nr, r);
      sum = fmaf(window_sum(b, x + (int64_t)k * Nr * Nc + c, Nc), b.inv, sum);
    }
  }
  part[part_id][px] = sum;
  __syncthreads();
  if (part_id == 0 && on) {
    float total = part[0][px];
#pragma unroll
    for (int q = 1; q < kBlurSplit; ++q) total += part[q][px];
    acc[(int64_t)r * Nc + c] += total;
  }
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_footprint_blur(const float* partial, int K, int nblocks, float center_height,
                                     const float* cell_width, float scale_channel, float scale_row, int Nc, int Nr,
                                     const float* channel_weight, const float* image, float* heights, float* work,
                                     float* acc, void* stream) {
  using namespace citrus;
  if (!partial || !cell_width || !image || !heights || !work || !acc) return fail(CITRUS_ERR_ARG, "NULL input or output");
  if (K < 0 || K > 65535 || nblocks < 1 || Nc < 1 || Nr < 1)
    return fail(CITRUS_ERR_ARG, "0 <= K <= 65535, nblocks >= 1, Nc >= 1 and Nr >= 1 required");
  if (K == 0) return CITRUS_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(footprint_heights_kernel, dim3((unsigned)K), dim3(256), 0, s, partial, nblocks, center_height,
                     heights);
  const dim3 tiles((unsigned)((Nc + 15) / 16), (unsigned)((Nr + 15) / 16));
  hipLaunchKernelGGL(blur_channels_kernel, dim3(tiles.x, tiles.y, (unsigned)K), dim3(256), 0, s, image, heights,
                     cell_width, scale_channel, Nr, Nc, channel_weight, work);
  hipLaunchKernelGGL(blur_rows_accumulate_kernel, dim3(tiles.x, (unsigned)((Nr + 3) / 4)), dim3(256), 0, s, work,
                     heights, cell_width, scale_row, Nr, Nc, K, acc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("footprint blur: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_box_matrix(const float* widths, int k, int N, float* out, void* stream) {
  using namespace citrus;
  if (!widths || !out) return fail(CITRUS_ERR_ARG, "NULL widths or output");
  if (k < 0 || N < 1) return fail(CITRUS_ERR_ARG, "k >= 0 and N >= 1 required");
  const int64_t total = (int64_t)k * N * N;
  if (total == 0) return CITRUS_OK;
  hipLaunchKernelGGL(box_matrix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), widths, k, N, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("box matrix: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_footprint_filters(const float* partial, int K, int nblocks, float center_height,
                                        const float* cell_width, float scale_channel, float scale_row, int Nc, int Nr,
                                        const float* channel_weight, float* heights,
                                        float* channel_filter_t, float* row_filter, void* stream) {
  using namespace citrus;
  if (!partial || !cell_width || !heights || !channel_filter_t || !row_filter)
    return fail(CITRUS_ERR_ARG, "NULL input or output");
  if (K < 0 || K > 65535 || nblocks < 1 || Nc < 1 || Nr < 1)
    return fail(CITRUS_ERR_ARG, "0 <= K <= 65535, nblocks >= 1, Nc >= 1 and Nr >= 1 required");
  if (K == 0) return CITRUS_OK;
  const int64_t total = (int64_t)K * Nc * Nc + (int64_t)K * Nr * Nr;
  if (!grid_fits((total + 255) / 256, 256)) return fail(CITRUS_ERR_ARG, "filters too large for one launch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(footprint_heights_kernel, dim3((unsigned)K), dim3(256), 0, s, partial, nblocks, center_height,
                     heights);
  hipLaunchKernelGGL(footprint_filters_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, heights,
                     cell_width, scale_channel, scale_row, K, Nc, Nr, channel_weight, channel_filter_t,
                     row_filter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("footprint filters: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

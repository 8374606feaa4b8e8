// Kernels of the Phi_{>2} surrogate network (gfx950): 3D convolution forward, input gradient and weight gradient with
// bfloat16 weights and float32 accumulation, leaky-ReLU backward, 2x average pooling and nearest upsampling with their
// adjoints, the Adam step and bfloat16 rounding. The network has no counterpart in the reference.
//
// Activations are [B][C][D][H][W] float32; the channels of one sample are contiguous, and a batch stride (elements
// between samples) lets a tensor be a channel slice of a larger one (the U-Net's concatenations are written in place).
// Weights are [Co][Ci][k][k][k] bfloat16 (k = 3 or 1, zero padding (k-1)/2, stride 1).
//
// Convolution: a workgroup computes a 4 x 8 x 8 voxel tile for kCob output channels. Per input channel it stages the
// input halo tile and the bfloat16 weights of the (output-channel block, input channel) pair, widened to float32 as
// [tap][output channel], in LDS; each thread accumulates kCob outputs of its voxel in float32, reading four weights
// per 16-byte LDS load. The input gradient is the same kernel reading the kernel transposed and flipped. The weight
// gradient stages the input halo and the output-gradient tile, has each thread sum the 27 taps of one output channel
// over a sixteenth of the tile's voxels, and reduces the slices and then the voxel chunks in fixed order (no atomics).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kTz = 4, kTy = 8, kTx = 8;      // voxel tile of a workgroup
constexpr int kThreads = kTz * kTy * kTx;     // 256
constexpr int kCob = 16;                      // output channels per workgroup
constexpr float kSlope = 0.01f;               // leaky-ReLU slope

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {  // round to nearest even
  const uint32_t b = __float_as_uint(f);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

struct Vol {
  int D, H, W;
  __host__ __device__ int64_t plane() const { return (int64_t)D * H * W; }
};

// out[b][co] = bias[co] + sum_ci sum_taps w * in (TRANSPOSED: out channel c, in channel i read w[i][c][flipped tap])
template <int K, bool TRANSPOSED, bool ACT>
__global__ __launch_bounds__(kThreads) void conv_kernel(const float* __restrict__ in, int64_t in_bs, int Cin,
                                                        Vol v, const uint16_t* __restrict__ w, int w_ci,
                                                        const float* __restrict__ bias, float* __restrict__ out,
                                                        int64_t out_bs, int Cout) {
  constexpr int P = (K - 1) / 2;
  constexpr int HZ = kTz + K - 1, HY = kTy + K - 1, HX = kTx + K - 1;
  constexpr int K3 = K * K * K;
  __shared__ float tile[HZ * HY * HX];
  __shared__ f4 wl[K3 * (kCob / 4)];  // weights of one input channel, [tap][output channel], widened to float32
  float* wlf = reinterpret_cast<float*>(wl);
  const int ntx = (v.W + kTx - 1) / kTx, nty = (v.H + kTy - 1) / kTy;
  const int t = blockIdx.x;
  const int bx = (t % ntx) * kTx, by = ((t / ntx) % nty) * kTy, bz = (t / (ntx * nty)) * kTz;
  const int co0 = blockIdx.y * kCob;
  const int b = blockIdx.z;
  const int tx = threadIdx.x % kTx, ty = (threadIdx.x / kTx) % kTy, tz = threadIdx.x / (kTx * kTy);
  const int64_t plane = v.plane();
  const float* inb = in + (int64_t)b * in_bs;
  f4 acc[kCob / 4];
#pragma unroll
  for (int c = 0; c < kCob / 4; ++c) acc[c] = f4{0.f, 0.f, 0.f, 0.f};
  for (int ci = 0; ci < Cin; ++ci) {
    __syncthreads();
    const float* src = inb + (int64_t)ci * plane;
    for (int i = threadIdx.x; i < HZ * HY * HX; i += kThreads) {
      const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
      const int z = bz + hz - P, y = by + hy - P, x = bx + hx - P;
      const bool ok = z >= 0 && z < v.D && y >= 0 && y < v.H && x >= 0 && x < v.W;
      tile[i] = ok ? src[((int64_t)z * v.H + y) * v.W + x] : 0.0f;
    }
    for (int i = threadIdx.x; i < K3 * kCob; i += kThreads) {
      const int tap = i / kCob, c = i % kCob;
      const int co = co0 + c;
      float wv = 0.0f;
      if (co < Cout) {
        const int64_t wi = TRANSPOSED ? ((int64_t)ci * w_ci + co) * K3 + (K3 - 1 - tap)
                                      : ((int64_t)co * w_ci + ci) * K3 + tap;
        wv = bf16_to_f32(w[wi]);
      }
      wlf[i] = wv;
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < K3; ++tap) {
      const int dz = tap / (K * K), dy = (tap / K) % K, dx = tap % K;
      const float xv = tile[((tz + dz) * HY + ty + dy) * HX + tx + dx];
#pragma unroll
      for (int c = 0; c < kCob / 4; ++c) acc[c] += wl[tap * (kCob / 4) + c] * xv;
    }
  }
  const int z = bz + tz, y = by + ty, x = bx + tx;
  if (z >= v.D || y >= v.H || x >= v.W) return;
  float* ob = out + (int64_t)b * out_bs + ((int64_t)z * v.H + y) * v.W + x;
#pragma unroll
  for (int c = 0; c < kCob; ++c) {
    const int co = co0 + c;
    if (co < Cout) {
      float r = acc[c / 4][c % 4] + (bias ? bias[co] : 0.0f);
      if (ACT) r = r > 0.0f ? r : kSlope * r;
      ob[(int64_t)co * plane] = r;
    }
  }
}

// partial[chunk][co][ci][tap] = sum over the chunk's tiles (and samples) of g[co] * in[ci] shifted by the tap
template <int K>
__global__ __launch_bounds__(kThreads) void wgrad_kernel(const float* __restrict__ in, int64_t in_bs, int Cin,
                                                         const float* __restrict__ g, int64_t g_bs, int Cout, Vol v,
                                                         int B, int tiles_per_chunk, float* __restrict__ partial) {
  constexpr int P = (K - 1) / 2;
  constexpr int HZ = kTz + K - 1, HY = kTy + K - 1, HX = kTx + K - 1;
  constexpr int K3 = K * K * K;
  __shared__ float xt[HZ * HY * HX];
  __shared__ float gt[kCob * kThreads];
  __shared__ float red[16 * kCob * K3];
  const int chunk = blockIdx.x, ci = blockIdx.y, co0 = blockIdx.z * kCob;
  const int ntx = (v.W + kTx - 1) / kTx, nty = (v.H + kTy - 1) / kTy, ntz = (v.D + kTz - 1) / kTz;
  const int ntiles = ntx * nty * ntz;
  const int64_t plane = v.plane();
  const int c = threadIdx.x % kCob;       // output channel of this thread
  const int slice = threadIdx.x / kCob;   // sixteenth of the tile: voxels 16 slice .. 16 slice + 15
  float acc[K3];
#pragma unroll
  for (int k = 0; k < K3; ++k) acc[k] = 0.0f;
  const int total = ntiles * B;
  for (int it = 0; it < tiles_per_chunk; ++it) {
    const int job = chunk * tiles_per_chunk + it;
    if (job >= total) break;  // uniform
    const int t = job % ntiles, b = job / ntiles;
    const int bx = (t % ntx) * kTx, by = ((t / ntx) % nty) * kTy, bz = (t / (ntx * nty)) * kTz;
    __syncthreads();
    const float* src = in + (int64_t)b * in_bs + (int64_t)ci * plane;
    for (int i = threadIdx.x; i < HZ * HY * HX; i += kThreads) {
      const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
      const int z = bz + hz - P, y = by + hy - P, x = bx + hx - P;
      const bool ok = z >= 0 && z < v.D && y >= 0 && y < v.H && x >= 0 && x < v.W;
      xt[i] = ok ? src[((int64_t)z * v.H + y) * v.W + x] : 0.0f;
    }
    {
      const int vx = threadIdx.x % kTx, vy = (threadIdx.x / kTx) % kTy, vz = threadIdx.x / (kTx * kTy);
      const int z = bz + vz, y = by + vy, x = bx + vx;
      const bool ok = z < v.D && y < v.H && x < v.W;
      const float* gb = g + (int64_t)b * g_bs + ((int64_t)z * v.H + y) * v.W + x;
#pragma unroll
      for (int cc = 0; cc < kCob; ++cc)
        gt[cc * kThreads + threadIdx.x] = (ok && co0 + cc < Cout) ? gb[(int64_t)(co0 + cc) * plane] : 0.0f;
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < 16; ++j) {
      const int vox = slice * 16 + j;
      const int vx = vox % kTx, vy = (vox / kTx) % kTy, vz = vox / (kTx * kTy);
      const float gv = gt[c * kThreads + vox];
#pragma unroll
      for (int tap = 0; tap < K3; ++tap) {
        const int dz = tap / (K * K), dy = (tap / K) % K, dx = tap % K;
        acc[tap] = fmaf(gv, xt[((vz + dz) * HY + vy + dy) * HX + vx + dx], acc[tap]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int tap = 0; tap < K3; ++tap) red[(slice * kCob + c) * K3 + tap] = acc[tap];
  __syncthreads();
  for (int i = threadIdx.x; i < kCob * K3; i += kThreads) {
    const int cc = i / K3, tap = i % K3;
    float s = 0.0f;
    for (int sl = 0; sl < 16; ++sl) s += red[(sl * kCob + cc) * K3 + tap];
    if (co0 + cc < Cout) partial[(((int64_t)chunk * Cout + co0 + cc) * Cin + ci) * K3 + tap] = s;
  }
}

__global__ void sum_chunks_kernel(const float* __restrict__ partial, int chunks, int64_t n, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += partial[(int64_t)c * n + i];
  out[i] = s;
}

// db[co] = sum over samples and voxels of g[b][co]: one workgroup per channel, fixed-order tree
__global__ __launch_bounds__(256) void bias_grad_kernel(const float* __restrict__ g, int64_t g_bs, int B,
                                                        int64_t plane, float* __restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) {
    const float* gb = g + (int64_t)b * g_bs + (int64_t)co * plane;
    for (int64_t i = threadIdx.x; i < plane; i += 256) s += gb[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

__global__ void leaky_backward_kernel(float* __restrict__ g, int64_t g_bs, const float* __restrict__ y, int64_t y_bs,
                                      int64_t n, int B) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t b = i / n, e = i % n;
  const float yv = y[b * y_bs + e];
  g[b * g_bs + e] *= yv > 0.0f ? 1.0f : kSlope;
}

// y[c][z][y][x] = mean of the 2x2x2 block of x (input dims v)
__global__ void pool_kernel(const float* __restrict__ x, int64_t x_bs, Vol v, int C, int B, float* __restrict__ y,
                            int64_t y_bs) {
  const int D2 = v.D / 2, H2 = v.H / 2, W2 = v.W / 2;
  const int64_t n = (int64_t)C * D2 * H2 * W2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t b = i / n, e = i % n;
  const int xo = (int)(e % W2), yo = (int)((e / W2) % H2), zo = (int)((e / ((int64_t)W2 * H2)) % D2);
  const int c = (int)(e / ((int64_t)W2 * H2 * D2));
  const float* s = x + b * x_bs + (int64_t)c * v.plane();
  float acc = 0.0f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) acc += s[((int64_t)(2 * zo + dz) * v.H + 2 * yo + dy) * v.W + 2 * xo + dx];
  y[b * y_bs + e] = acc * 0.125f;
}

// dx (+)= g / 8 broadcast over each 2x2x2 block (input dims v of the pooling)
__global__ void pool_backward_kernel(const float* __restrict__ g, int64_t g_bs, Vol v, int C, int B,
                                     float* __restrict__ dx, int64_t dx_bs, int accumulate) {
  const int64_t n = (int64_t)C * v.plane();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t b = i / n, e = i % n;
  const int x = (int)(e % v.W), y = (int)((e / v.W) % v.H), z = (int)((e / ((int64_t)v.W * v.H)) % v.D);
  const int c = (int)(e / v.plane());
  const int D2 = v.D / 2, H2 = v.H / 2, W2 = v.W / 2;
  const float gv = g[b * g_bs + (((int64_t)c * D2 + z / 2) * H2 + y / 2) * W2 + x / 2] * 0.125f;
  float* o = dx + b * dx_bs + e;
  *o = accumulate ? *o + gv : gv;
}

// y (output dims 2v) = nearest upsampling of x (dims v)
__global__ void upsample_kernel(const float* __restrict__ x, int64_t x_bs, Vol v, int C, int B, float* __restrict__ y,
                                int64_t y_bs) {
  const int D2 = 2 * v.D, H2 = 2 * v.H, W2 = 2 * v.W;
  const int64_t n = (int64_t)C * D2 * H2 * W2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t b = i / n, e = i % n;
  const int xo = (int)(e % W2), yo = (int)((e / W2) % H2), zo = (int)((e / ((int64_t)W2 * H2)) % D2);
  const int c = (int)(e / ((int64_t)W2 * H2 * D2));
  y[b * y_bs + e] = x[b * x_bs + (((int64_t)c * v.D + zo / 2) * v.H + yo / 2) * v.W + xo / 2];
}

// dx (dims v) = sum of g (dims 2v) over each 2x2x2 block
__global__ void upsample_backward_kernel(const float* __restrict__ g, int64_t g_bs, Vol v, int C, int B,
                                         float* __restrict__ dx, int64_t dx_bs) {
  const int64_t n = (int64_t)C * v.plane();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t b = i / n, e = i % n;
  const int x = (int)(e % v.W), y = (int)((e / v.W) % v.H), z = (int)((e / ((int64_t)v.W * v.H)) % v.D);
  const int c = (int)(e / v.plane());
  const int H2 = 2 * v.H, W2 = 2 * v.W;
  const float* s = g + b * g_bs + (int64_t)c * 8 * v.plane();
  float acc = 0.0f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx2 = 0; dx2 < 2; ++dx2) acc += s[((int64_t)(2 * z + dz) * H2 + 2 * y + dy) * W2 + 2 * x + dx2];
  dx[b * dx_bs + e] = acc;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ p16, int64_t n, float lr, float beta1,
                            float beta2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (p16) p16[i] = f32_to_bf16(pi);
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f32_to_bf16(x[i]);
}

int finish(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CITRUS_OK;
}

int check_vol(int B, int C, int D, int H, int W) {
  if (B < 1 || C < 1 || D < 1 || H < 1 || W < 1) return fail(CITRUS_ERR_ARG, "empty tensor");
  if ((int64_t)C * D * H * W * B > ((int64_t)1 << 40)) return fail(CITRUS_ERR_ARG, "tensor too large");
  return CITRUS_OK;
}

dim3 elementwise_grid(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <bool TRANSPOSED, bool ACT>
int launch_conv(int k, const float* in, int64_t in_bs, int Cin, Vol v, const uint16_t* w, int w_ci,
                const float* bias, float* out, int64_t out_bs, int Cout, int B, hipStream_t s) {
  const int64_t tiles = (int64_t)((v.W + kTx - 1) / kTx) * ((v.H + kTy - 1) / kTy) * ((v.D + kTz - 1) / kTz);
  if (!grid_fits(tiles, kThreads) || (Cout + kCob - 1) / kCob > 65535) return fail(CITRUS_ERR_ARG, "grid too large");
  const dim3 grid((unsigned)tiles, (unsigned)((Cout + kCob - 1) / kCob), (unsigned)B);
  if (k == 3)
    hipLaunchKernelGGL((conv_kernel<3, TRANSPOSED, ACT>), grid, dim3(kThreads), 0, s, in, in_bs, Cin, v, w, w_ci,
                       bias, out, out_bs, Cout);
  else
    hipLaunchKernelGGL((conv_kernel<1, TRANSPOSED, ACT>), grid, dim3(kThreads), 0, s, in, in_bs, Cin, v, w, w_ci,
                       bias, out, out_bs, Cout);
  return finish("surrogate convolution");
}

}  // namespace
}  // namespace citrus

using namespace citrus;

extern "C" int citrus_nn_conv3d_forward(const float* x, int64_t x_bstride, int B, int Ci, int D, int H, int W,
                                        const uint16_t* w, const float* bias, int Co, int k, int leaky, float* y,
                                        int64_t y_bstride, void* stream) {
  if (int rc = check_vol(B, Ci, D, H, W)) return rc;
  if (Co < 1 || (k != 1 && k != 3)) return fail(CITRUS_ERR_ARG, "Co >= 1 and k in {1, 3} required");
  if (!x || !w || !y) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if ((int64_t)B > 65535) return fail(CITRUS_ERR_ARG, "at most 65535 samples per launch");
  const Vol v{D, H, W};
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (leaky) return launch_conv<false, true>(k, x, x_bstride, Ci, v, w, Ci, bias, y, y_bstride, Co, B, s);
  return launch_conv<false, false>(k, x, x_bstride, Ci, v, w, Ci, bias, y, y_bstride, Co, B, s);
}

extern "C" int citrus_nn_conv3d_backward_data(const float* g, int64_t g_bstride, int B, int Co, int D, int H, int W,
                                              const uint16_t* w, int Ci, int k, float* dx, int64_t dx_bstride,
                                              void* stream) {
  if (int rc = check_vol(B, Co, D, H, W)) return rc;
  if (Ci < 1 || (k != 1 && k != 3)) return fail(CITRUS_ERR_ARG, "Ci >= 1 and k in {1, 3} required");
  if (!g || !w || !dx) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if ((int64_t)B > 65535) return fail(CITRUS_ERR_ARG, "at most 65535 samples per launch");
  const Vol v{D, H, W};
  // the layer's output gradient (Co channels) in, its input gradient (Ci channels) out, kernel read as w[co][ci][-tap]
  return launch_conv<true, false>(k, g, g_bstride, Co, v, w, Ci, nullptr, dx, dx_bstride, Ci, B,
                                  reinterpret_cast<hipStream_t>(stream));
}

extern "C" int64_t citrus_nn_wgrad_workspace_floats(int B, int Ci, int Co, int D, int H, int W, int k) {
  if (B < 1 || Ci < 1 || Co < 1 || D < 1 || H < 1 || W < 1 || (k != 1 && k != 3)) return -1;
  const int64_t tiles = (int64_t)((W + kTx - 1) / kTx) * ((H + kTy - 1) / kTy) * ((D + kTz - 1) / kTz) * B;
  const int64_t chunks = (tiles + 15) / 16;
  return chunks * Co * Ci * k * k * k;
}

extern "C" int citrus_nn_conv3d_backward_weight(const float* x, int64_t x_bstride, const float* g, int64_t g_bstride,
                                                int B, int Ci, int Co, int D, int H, int W, int k, float* dw,
                                                float* db, float* workspace, void* stream) {
  if (int rc = check_vol(B, Ci, D, H, W)) return rc;
  if (Co < 1 || (k != 1 && k != 3)) return fail(CITRUS_ERR_ARG, "Co >= 1 and k in {1, 3} required");
  if (!x || !g || !dw || !workspace) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (Ci > 65535 || (Co + kCob - 1) / kCob > 65535) return fail(CITRUS_ERR_ARG, "too many channels");
  const Vol v{D, H, W};
  const int tiles_per_chunk = 16;
  const int64_t tiles = (int64_t)((W + kTx - 1) / kTx) * ((H + kTy - 1) / kTy) * ((D + kTz - 1) / kTz) * B;
  const int64_t chunks = (tiles + tiles_per_chunk - 1) / tiles_per_chunk;
  if (!grid_fits(chunks, kThreads)) return fail(CITRUS_ERR_ARG, "too many tiles");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)chunks, (unsigned)Ci, (unsigned)((Co + kCob - 1) / kCob));
  if (k == 3)
    hipLaunchKernelGGL(wgrad_kernel<3>, grid, dim3(kThreads), 0, s, x, x_bstride, Ci, g, g_bstride, Co, v, B,
                       tiles_per_chunk, workspace);
  else
    hipLaunchKernelGGL(wgrad_kernel<1>, grid, dim3(kThreads), 0, s, x, x_bstride, Ci, g, g_bstride, Co, v, B,
                       tiles_per_chunk, workspace);
  const int64_t n = (int64_t)Co * Ci * k * k * k;
  hipLaunchKernelGGL(sum_chunks_kernel, elementwise_grid(n), dim3(256), 0, s, workspace, (int)chunks, n, dw);
  if (db) hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)Co), dim3(256), 0, s, g, g_bstride, B, v.plane(), db);
  return finish("surrogate weight gradient");
}

extern "C" int citrus_nn_leaky_backward(float* g, int64_t g_bstride, const float* y, int64_t y_bstride,
                                        int64_t n_per_sample, int B, void* stream) {
  if (!g || !y || n_per_sample < 1 || B < 1) return fail(CITRUS_ERR_ARG, "NULL buffer or empty tensor");
  hipLaunchKernelGGL(leaky_backward_kernel, elementwise_grid(n_per_sample * B), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), g, g_bstride, y, y_bstride, n_per_sample, B);
  return finish("surrogate activation gradient");
}

extern "C" int citrus_nn_avg_pool2(const float* x, int64_t x_bstride, int B, int C, int D, int H, int W, float* y,
                                   int64_t y_bstride, void* stream) {
  if (int rc = check_vol(B, C, D, H, W)) return rc;
  if (D % 2 || H % 2 || W % 2) return fail(CITRUS_ERR_ARG, "pooling needs even sizes");
  if (!x || !y) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const Vol v{D, H, W};
  hipLaunchKernelGGL(pool_kernel, elementwise_grid((int64_t)B * C * v.plane() / 8), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), x, x_bstride, v, C, B, y, y_bstride);
  return finish("surrogate pooling");
}

extern "C" int citrus_nn_avg_pool2_backward(const float* g, int64_t g_bstride, int B, int C, int D, int H, int W,
                                            float* dx, int64_t dx_bstride, int accumulate, void* stream) {
  if (int rc = check_vol(B, C, D, H, W)) return rc;
  if (D % 2 || H % 2 || W % 2) return fail(CITRUS_ERR_ARG, "pooling needs even sizes");
  if (!g || !dx) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const Vol v{D, H, W};
  hipLaunchKernelGGL(pool_backward_kernel, elementwise_grid((int64_t)B * C * v.plane()), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), g, g_bstride, v, C, B, dx, dx_bstride, accumulate);
  return finish("surrogate pooling gradient");
}

extern "C" int citrus_nn_upsample2(const float* x, int64_t x_bstride, int B, int C, int D, int H, int W, float* y,
                                   int64_t y_bstride, void* stream) {
  if (int rc = check_vol(B, C, D, H, W)) return rc;
  if (!x || !y) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const Vol v{D, H, W};
  hipLaunchKernelGGL(upsample_kernel, elementwise_grid((int64_t)B * C * v.plane() * 8), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), x, x_bstride, v, C, B, y, y_bstride);
  return finish("surrogate upsampling");
}

extern "C" int citrus_nn_upsample2_backward(const float* g, int64_t g_bstride, int B, int C, int D, int H, int W,
                                            float* dx, int64_t dx_bstride, void* stream) {
  if (int rc = check_vol(B, C, D, H, W)) return rc;
  if (!g || !dx) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const Vol v{D, H, W};
  hipLaunchKernelGGL(upsample_backward_kernel, elementwise_grid((int64_t)B * C * v.plane()), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), g, g_bstride, v, C, B, dx, dx_bstride);
  return finish("surrogate upsampling gradient");
}

extern "C" int citrus_nn_adam(float* p, const float* g, float* m, float* v, uint16_t* p_bf16, int64_t n, float lr,
                              float beta1, float beta2, float eps, int step, void* stream) {
  if (!p || !g || !m || !v || n < 1 || step < 1) return fail(CITRUS_ERR_ARG, "NULL buffer, n < 1 or step < 1");
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, elementwise_grid(n), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p, g, m,
                     v, p_bf16, n, lr, beta1, beta2, eps, c1, c2);
  return finish("surrogate Adam step");
}

extern "C" int citrus_nn_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
  if (!x || !y || n < 1) return fail(CITRUS_ERR_ARG, "NULL buffer or n < 1");
  hipLaunchKernelGGL(to_bf16_kernel, elementwise_grid(n), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), x, y,
                     n);
  return finish("surrogate bf16 rounding");
}

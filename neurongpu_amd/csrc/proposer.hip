// FP32 training kernels of the learned RK5 step-size proposer.
//
// One training step = one minibatch:
//   grad_kernel  one sample per thread: forward and backward pass of the
//                6-32-32-1 ReLU MLP, per-parameter gradients summed in LDS
//                (ds_add_f32) and then added to a global gradient buffer;
//   update_kernel  SGD or Adam on the 1313 parameters, clears the gradient
//                buffer, refreshes the BF16 inference copy.
#include <hip/hip_runtime.h>

#include "ngpu_device.h"
#include "proposer.h"

namespace ngpu {

namespace {

constexpr int kTB = 256;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(kTB) void grad_kernel(PropTrainArgs A) {
  __shared__ float s_w[kPropNParam];
  __shared__ float s_g[kPropNParam];
  __shared__ float s_loss;
  for (int k = threadIdx.x; k < kPropNParam; k += kTB) {
    s_w[k] = A.w[k];
    s_g[k] = 0.0f;
  }
  if (threadIdx.x == 0) s_loss = 0.0f;
  __syncthreads();
  int b = blockIdx.x * kTB + threadIdx.x;
  if (b < A.batch && A.n > 0) {
    unsigned long long idx;
    if (A.full_batch) {
      idx = (unsigned long long)b % A.n;
    } else {
      uint32_t u0 = dev_philox_w0((uint32_t)b, A.iter, 0x70726F70u, 0u,
                                  A.seed, 0x74726169u);
      uint32_t u1 = dev_philox_w0((uint32_t)b, A.iter, 0x70726F70u, 1u,
                                  A.seed, 0x74726169u);
      idx = (((unsigned long long)u1 << 32) | u0) % A.n;
    }
    const float *s = A.samples + idx * kPropSample;
    float x[kPropIn];
#pragma unroll
    for (int i = 0; i < kPropIn; i++) x[i] = s[i];
    float t = s[kPropIn];
    float h1[kPropHid], h2[kPropHid];
#pragma unroll
    for (int j = 0; j < kPropHid; j++) {
      float a = s_w[kPropB1 + j];
#pragma unroll
      for (int i = 0; i < kPropIn; i++)
        a = __builtin_fmaf(s_w[kPropW1 + j * kPropIn + i], x[i], a);
      h1[j] = a > 0.0f ? a : 0.0f;
    }
    float y = s_w[kPropB3];
#pragma unroll
    for (int j = 0; j < kPropHid; j++) {
      float a = s_w[kPropB2 + j];
#pragma unroll
      for (int i = 0; i < kPropHid; i++)
        a = __builtin_fmaf(s_w[kPropW2 + j * kPropHid + i], h1[i], a);
      h2[j] = a > 0.0f ? a : 0.0f;
      y = __builtin_fmaf(s_w[kPropW3 + j], h2[j], y);
    }
    float dy = y - t;  // d(0.5 (y - t)^2)/dy
    atomicAdd(&s_loss, 0.5f * dy * dy);
    atomicAdd(&s_g[kPropB3], dy);
    float dh1[kPropHid];
#pragma unroll
    for (int i = 0; i < kPropHid; i++) dh1[i] = 0.0f;
#pragma unroll
    for (int j = 0; j < kPropHid; j++) {
      atomicAdd(&s_g[kPropW3 + j], dy * h2[j]);
      float d2 = h2[j] > 0.0f ? dy * s_w[kPropW3 + j] : 0.0f;
      if (d2 != 0.0f) {
        atomicAdd(&s_g[kPropB2 + j], d2);
#pragma unroll
        for (int i = 0; i < kPropHid; i++) {
          if (h1[i] != 0.0f) atomicAdd(&s_g[kPropW2 + j * kPropHid + i], d2 * h1[i]);
          dh1[i] = __builtin_fmaf(s_w[kPropW2 + j * kPropHid + i], d2, dh1[i]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kPropHid; j++) {
      float d1 = h1[j] > 0.0f ? dh1[j] : 0.0f;
      if (d1 != 0.0f) {
        atomicAdd(&s_g[kPropB1 + j], d1);
#pragma unroll
        for (int i = 0; i < kPropIn; i++)
          atomicAdd(&s_g[kPropW1 + j * kPropIn + i], d1 * x[i]);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kPropNParam; k += kTB)
    if (s_g[k] != 0.0f) atomicAdd(A.grad + k, s_g[k]);
  if (threadIdx.x == 0) atomicAdd(A.loss, s_loss);
}

__global__ void update_kernel(PropTrainArgs A, uint32_t step) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kPropNParam) return;
  float g = A.grad[k] / (float)A.batch;
  float w = A.w[k];
  if (A.adam) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float m = b1 * A.m[k] + (1.0f - b1) * g;
    float v = b2 * A.v[k] + (1.0f - b2) * g * g;
    A.m[k] = m;
    A.v[k] = v;
    float mh = m / (1.0f - powf(b1, (float)step));
    float vh = v / (1.0f - powf(b2, (float)step));
    w -= A.lr * mh / (sqrtf(vh) + eps);
  } else {
    w -= A.lr * g;
  }
  A.w[k] = w;
  A.w_bf16[k] = to_bf16(w);
  A.grad[k] = 0.0f;
}

__global__ void pack_kernel(const float *w, unsigned char *pk) {
  int l = threadIdx.x;  // 64 threads
  if (l >= 64) return;
  uint16_t *w1p = reinterpret_cast<uint16_t *>(pk + kPropPkW1);
  uint16_t *w2p = reinterpret_cast<uint16_t *>(pk + kPropPkW2);
  float *b1p = reinterpret_cast<float *>(pk + kPropPkB1);
  float *b2p = reinterpret_cast<float *>(pk + kPropPkB2);
  float *w3p = reinterpret_cast<float *>(pk + kPropPkW3);
  int row = l % 32, half = l / 32;
  for (int j = 0; j < 4; j++) {
    int k = 4 * half + j;
    w1p[l * 4 + j] = k < kPropIn ? to_bf16(w[kPropW1 + row * kPropIn + k])
                                 : (uint16_t)0;
    for (int c = 0; c < 4; c++)
      w2p[(c * 64 + l) * 4 + j] =
          to_bf16(w[kPropW2 + row * kPropHid + 8 * c + 4 * half + j]);
  }
  for (int r = 0; r < 16; r++) {
    int unit = 8 * (r / 4) + 4 * half + (r % 4);
    b1p[l * 16 + r] = w[kPropB1 + unit];
    b2p[l * 16 + r] = w[kPropB2 + unit];
    w3p[l * 16 + r] = w[kPropW3 + unit];
  }
  if (l == 0) *reinterpret_cast<float *>(pk + kPropPkB3) = w[kPropB3];
}

__global__ void to_bf16_kernel(const float *w, uint16_t *o) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kPropNParam) o[k] = to_bf16(w[k]);
}

}  // namespace

const char *LaunchPropTrainStep(const PropTrainArgs &a, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (a.batch <= 0) return "batch must be positive";
  int grid = (a.batch + kTB - 1) / kTB;
  hipLaunchKernelGGL(grad_kernel, dim3(grid), dim3(kTB), 0, s, a);
  hipLaunchKernelGGL(update_kernel, dim3((kPropNParam + 255) / 256), dim3(256),
                     0, s, a, a.iter + 1u);
  return nullptr;
}

void LaunchPropPack(const float *w, unsigned char *packed, void *stream) {
  hipLaunchKernelGGL(pack_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     w, packed);
}

void LaunchPropToBf16(const float *w, uint16_t *w_bf16, void *stream) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((kPropNParam + 255) / 256),
                     dim3(256), 0, (hipStream_t)stream, w, w_bf16);
}

}  // namespace ngpu

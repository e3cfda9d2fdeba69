// Loss and optimiser kernels of batched optimisation loops.
//
// MSE_Intensity (xlumina/loss_functions.py:137-149) for a batch of detector
// fields with its field cotangent, and the Adam / AdamW update (optax.adam /
// optax.adamw, experiments/four_f_optimizer.py:112).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xl {

constexpr int kMseThreads = 256;
constexpr int kMseBlocks = 256;   // partial sums per batch member

struct MseArgs {
  const float2* field;
  int64_t fbs;
  const float* target;
  float2* grad;
  int64_t gbs;
  float* partial;      // batch x kMseBlocks
  float* loss;
  int64_t npix;
  float gscale;        // scale * 4 / npix
  float lscale;        // scale / npix
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// grid (kMseBlocks, batch): partial[b][blk] = sum over the block's pixels of
// (|E|^2 - T)^2, accumulated in a fixed order.
__global__ __launch_bounds__(kMseThreads) void mse_kernel(MseArgs a) {
  const int b = blockIdx.y;
  const float2* e = a.field + (int64_t)b * a.fbs;
  float2* g = a.grad ? a.grad + (int64_t)b * a.gbs : nullptr;
  float acc = 0.f;
  for (int64_t p = (int64_t)blockIdx.x * kMseThreads + threadIdx.x; p < a.npix;
       p += (int64_t)gridDim.x * kMseThreads) {
    const float2 v = e[p];
    const float d = v.x * v.x + v.y * v.y - a.target[p];
    acc += d * d;
    if (g) g[p] = make_float2(a.gscale * d * v.x, a.gscale * d * v.y);
  }
  __shared__ float red[kMseThreads / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kMseThreads / 64; ++w) s += red[w];
    a.partial[(int64_t)b * gridDim.x + blockIdx.x] = s;
  }
}

// One block per batch member: loss[b] = lscale * sum of its partials.
__global__ __launch_bounds__(kMseThreads) void mse_reduce_kernel(MseArgs a, int nblk) {
  const int b = blockIdx.x;
  float acc = 0.f;
  for (int i = threadIdx.x; i < nblk; i += kMseThreads) acc += a.partial[(int64_t)b * nblk + i];
  __shared__ float red[kMseThreads / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kMseThreads / 64; ++w) s += red[w];
    a.loss[b] = a.lscale * s;
  }
}

struct AdamArgs {
  float* p;
  const float* g;
  float* m;
  float* v;
  int64_t n;
  const int32_t* count;
  float lr, b1, b2, eps, wd;
  int vec;             // all four arrays 16-byte aligned: float4 accesses
};

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
  const int t = *a.count + 1;
  const float c1 = (float)(1.0 / (1.0 - pow((double)a.b1, (double)t)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)a.b2, (double)t)));
  const int64_t n4 = a.vec ? a.n >> 2 : 0;
  float4* p4 = (float4*)a.p;
  const float4* g4 = (const float4*)a.g;
  float4* m4 = (float4*)a.m;
  float4* v4 = (float4*)a.v;
  auto upd = [&](float& p, float g, float& m, float& v) {
    m = a.b1 * m + (1.f - a.b1) * g;
    v = a.b2 * v + (1.f - a.b2) * g * g;
    const float mh = m * c1, vh = v * c2;
    p -= a.lr * (mh / (sqrtf(vh) + a.eps) + a.wd * p);
  };
#pragma unroll 2
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256) {
    float4 p = p4[i], m = m4[i], v = v4[i];
    const float4 g = g4[i];
    upd(p.x, g.x, m.x, v.x);
    upd(p.y, g.y, m.y, v.y);
    upd(p.z, g.z, m.z, v.z);
    upd(p.w, g.w, m.w, v.w);
    p4[i] = p;
    m4[i] = m;
    v4[i] = v;
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256)
    upd(a.p[i], a.g[i], a.m[i], a.v[i]);
}

// Advances the step counter after every block of adam_kernel has read it.
__global__ void adam_tick_kernel(int32_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = count[0] + 1;
}

}  // namespace xl

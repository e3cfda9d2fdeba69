// Two-port beam splitters on (Ex, Ey) (xlumina/optical_elements.py:334-513):
// c = ca a + cb b, d = da a + db b per pixel, and the complex sums that give
// the cotangents of the four coefficients.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "opt.h"

namespace xl {

struct BsArgs {
  const float2* a;
  const float2* b;
  int64_t acs, bcs;
  float2* c;
  float2* d;
  int64_t ccs, dcs;
  float2 ca, cb, da, db;
  int64_t npix;
  int zero_ez;
  const float2* coef;  // optional device copy of (ca, cb, da, db)
};

__device__ __forceinline__ float2 bs_mac(float2 k, float2 v, float2 acc) {
  return make_float2(acc.x + k.x * v.x - k.y * v.y, acc.y + k.x * v.y + k.y * v.x);
}

// Grid-stride over the 2 npix samples of (Ex, Ey); with zero_ez also over the
// npix samples of component 2 of each output.
__global__ __launch_bounds__(256) void bs_kernel(BsArgs a) {
  const int64_t n2 = 2 * a.npix, total = a.zero_ez ? 3 * a.npix : n2;
  const float2 zero = make_float2(0.f, 0.f);
  const float2 ca = a.coef ? a.coef[0] : a.ca, cb = a.coef ? a.coef[1] : a.cb;
  const float2 da = a.coef ? a.coef[2] : a.da, db = a.coef ? a.coef[3] : a.db;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int comp = (int)(e / a.npix);
    const int64_t p = e - (int64_t)comp * a.npix;
    if (comp == 2) {
      if (a.c) a.c[2 * a.ccs + p] = zero;
      if (a.d) a.d[2 * a.dcs + p] = zero;
      continue;
    }
    const float2 va = a.a[comp * a.acs + p];
    const float2 vb = a.b ? a.b[comp * a.bcs + p] : zero;
    if (a.c) a.c[comp * a.ccs + p] = bs_mac(cb, vb, bs_mac(ca, va, zero));
    if (a.d) a.d[comp * a.dcs + p] = bs_mac(db, vb, bs_mac(da, va, zero));
  }
}

constexpr int kBsDotBlocks = 256;

struct BsDotsArgs {
  const float2* gc;
  const float2* gd;
  int64_t gccs, gdcs;
  const float2* a;
  const float2* b;
  int64_t acs, bcs;
  float* partial;      // kBsDotBlocks x 8
  float* dots;
  int64_t npix;
};

// conj(u) v accumulated into acc
__device__ __forceinline__ void cdot_acc(float2 u, float2 v, float& re, float& im) {
  re += u.x * v.x + u.y * v.y;
  im += u.x * v.y - u.y * v.x;
}

// grid kBsDotBlocks: partial[blk] = the block's sums of conj(gc) a, conj(gc) b,
// conj(gd) a, conj(gd) b over both components.
__global__ __launch_bounds__(256) void bs_dots_kernel(BsDotsArgs a) {
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float2 zero = make_float2(0.f, 0.f);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < 2 * a.npix; e += (int64_t)gridDim.x * 256) {
    const int comp = e >= a.npix ? 1 : 0;
    const int64_t p = e - comp * a.npix;
    const float2 va = a.a[comp * a.acs + p];
    const float2 vb = a.b ? a.b[comp * a.bcs + p] : zero;
    if (a.gc) {
      const float2 g = a.gc[comp * a.gccs + p];
      cdot_acc(g, va, s[0], s[1]);
      cdot_acc(g, vb, s[2], s[3]);
    }
    if (a.gd) {
      const float2 g = a.gd[comp * a.gdcs + p];
      cdot_acc(g, va, s[4], s[5]);
      cdot_acc(g, vb, s[6], s[7]);
    }
  }
  __shared__ float red[8][4];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float w = wave_sum(s[k]);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    a.partial[blockIdx.x * 8 + k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
  }
}

// One block of 256 threads: dots[k] = sum over blocks of partial[.][k].  Each
// thread adds a fixed stride of the partials, then a fixed-order tree in LDS:
// the sums are deterministic and the loads are issued in parallel.
__global__ __launch_bounds__(256) void bs_dots_reduce_kernel(BsDotsArgs a, int nblk) {
  __shared__ float red[8][256];
  float acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.f;
  for (int i = threadIdx.x; i < nblk; i += 256) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += a.partial[i * 8 + k];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 8; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) a.dots[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace xl

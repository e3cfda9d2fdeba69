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

// Bias corrections 1 / (1 - b^t) of step t = count + 1, in float64.
__device__ __forceinline__ void adam_factors(const AdamArgs& a, float& c1, float& c2) {
  const int t = *a.count + 1;
  c1 = (float)(1.0 / (1.0 - pow((double)a.b1, (double)t)));
  c2 = (float)(1.0 / (1.0 - pow((double)a.b2, (double)t)));
}

// One parameter's Adam / AdamW step.
__device__ __forceinline__ void adam_update(const AdamArgs& a, float c1, float c2, float& p, float g, float& m,
                                            float& v) {
  m = a.b1 * m + (1.f - a.b1) * g;
  v = a.b2 * v + (1.f - a.b2) * g * g;
  const float mh = m * c1, vh = v * c2;
  p -= a.lr * (mh / (sqrtf(vh) + a.eps) + a.wd * p);
}

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
  float c1, c2;
  adam_factors(a, c1, c2);
  const int64_t n4 = a.vec ? a.n >> 2 : 0;
  float4* p4 = (float4*)a.p;
  const float4* g4 = (const float4*)a.g;
  float4* m4 = (float4*)a.m;
  float4* v4 = (float4*)a.v;
#pragma unroll 2
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256) {
    float4 p = p4[i], m = m4[i], v = v4[i];
    const float4 g = g4[i];
    adam_update(a, c1, c2, p.x, g.x, m.x, v.x);
    adam_update(a, c1, c2, p.y, g.y, m.y, v.y);
    adam_update(a, c1, c2, p.z, g.z, m.z, v.z);
    adam_update(a, c1, c2, p.w, g.w, m.w, v.w);
    p4[i] = p;
    m4[i] = m;
    v4[i] = v;
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256)
    adam_update(a, c1, c2, a.p[i], a.g[i], a.m[i], a.v[i]);
}

// Advances the step counter after every block of adam_kernel has read it.
__global__ void adam_tick_kernel(int32_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = count[0] + 1;
}

// ---------------------------------------------------------------- field losses
// MSE_Intensity / MSE_Amplitude / MSE_Phase (xlumina/loss_functions.py:107-149)
// against a complex target field, batched over fields or components.
enum { LOSS_INTENSITY = 0, LOSS_AMPLITUDE = 1, LOSS_PHASE = 2 };

struct FieldLossArgs {
  const float2* field;
  int64_t fbs;
  const float2* target;
  int64_t tbs;
  float2* grad;
  int64_t gbs;
  float* partial;      // batch x kMseBlocks
  float* loss;
  int64_t npix;
  float gscale;        // 4 s (intensity) or 2 s, s = scale / npix
  float lscale;        // s
};

// d of one pixel and the direction of its field cotangent (grad = gscale d dir)
template <int MODE>
__device__ __forceinline__ float loss_term(float2 v, float2 t, float2& dir) {
  if constexpr (MODE == LOSS_INTENSITY) {
    dir = v;
    return v.x * v.x + v.y * v.y - (t.x * t.x + t.y * t.y);
  } else if constexpr (MODE == LOSS_AMPLITUDE) {
    const float av = sqrtf(v.x * v.x + v.y * v.y);
    const float inv = av > 0.f ? 1.f / av : 0.f;
    dir = make_float2(v.x * inv, v.y * inv);
    return av - sqrtf(t.x * t.x + t.y * t.y);
  } else {
    const float a2 = v.x * v.x + v.y * v.y;
    const float inv = a2 > 0.f ? 1.f / a2 : 0.f;
    dir = make_float2(-v.y * inv, v.x * inv);   // i E / |E|^2 = d angle / d(Re + i Im)
    return atan2f(v.y, v.x) - atan2f(t.y, t.x);
  }
}

// grid (kMseBlocks, batch): partial[b][blk] = sum over the block's pixels of d^2.
template <int MODE>
__global__ __launch_bounds__(kMseThreads) void field_loss_kernel(FieldLossArgs a) {
  const int b = blockIdx.y;
  const float2* e = a.field + (int64_t)b * a.fbs;
  const float2* tg = a.target + (int64_t)b * a.tbs;
  float2* g = a.grad ? a.grad + (int64_t)b * a.gbs : nullptr;
  float acc = 0.f;
  for (int64_t p = (int64_t)blockIdx.x * kMseThreads + threadIdx.x; p < a.npix;
       p += (int64_t)gridDim.x * kMseThreads) {
    float2 dir;
    const float d = loss_term<MODE>(e[p], tg[p], dir);
    acc += d * d;
    if (g) g[p] = make_float2(a.gscale * d * dir.x, a.gscale * d * dir.y);
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

__global__ __launch_bounds__(kMseThreads) void field_loss_reduce_kernel(FieldLossArgs a, int nblk) {
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

// ---------------------------------------------------------------- small-area loss
// small_area_hybrid (xlumina/loss_functions.py:24-40): per detector b with
// S = sum D, I = D / (S + eps), m = [I > 0.7 max I]:
// loss = scale * sum m / (sum m I + eps).
constexpr float kSaFrac = 0.7f;   // epsilon of the reference (fraction of the peak)
constexpr float kSaEps = 1e-8f;

struct SmallAreaArgs {
  const float* in;
  int64_t ibs;
  float* grad;
  int64_t gbs;
  float* partial;      // batch x kMseBlocks x 2
  float* stats;        // batch x 4: S, max D, count, sum m I
  float* loss;
  int64_t npix;
  float scale;
};

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Block sums of two per-thread values (PHASE 0: sum D and max D; PHASE 1:
// mask count and sum m I) into partial[b][blk][0..1].
template <int PHASE>
__global__ __launch_bounds__(kMseThreads) void small_area_kernel(SmallAreaArgs a) {
  const int b = blockIdx.y;
  const float* d = a.in + (int64_t)b * a.ibs;
  float s0 = 0.f, s1 = PHASE == 0 ? -__builtin_huge_valf() : 0.f;
  float den = 1.f, imax = 0.f;
  if constexpr (PHASE == 1) {
    den = a.stats[4 * b] + kSaEps;
    imax = a.stats[4 * b + 1] / den;
  }
  for (int64_t p = (int64_t)blockIdx.x * kMseThreads + threadIdx.x; p < a.npix;
       p += (int64_t)gridDim.x * kMseThreads) {
    const float v = d[p];
    if constexpr (PHASE == 0) {
      s0 += v;
      s1 = fmaxf(s1, v);
    } else {
      const float i = v / den;
      if (i > kSaFrac * imax) {
        s0 += 1.f;
        s1 += i;
      }
    }
  }
  __shared__ float red[2][kMseThreads / 64];
  s0 = wave_sum(s0);
  s1 = PHASE == 0 ? wave_max(s1) : wave_sum(s1);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s0;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float t0 = 0.f, t1 = PHASE == 0 ? -__builtin_huge_valf() : 0.f;
    for (int w = 0; w < kMseThreads / 64; ++w) {
      t0 += red[0][w];
      t1 = PHASE == 0 ? fmaxf(t1, red[1][w]) : t1 + red[1][w];
    }
    float* o = a.partial + ((int64_t)b * gridDim.x + blockIdx.x) * 2;
    o[0] = t0;
    o[1] = t1;
  }
}

// One block per detector: reduces the partials of PHASE into stats; after
// phase 1 also writes the loss.  Counts are summed as integers (each block's
// count is exact in float32).
template <int PHASE>
__global__ __launch_bounds__(kMseThreads) void small_area_reduce_kernel(SmallAreaArgs a, int nblk) {
  const int b = blockIdx.x;
  const float* part = a.partial + (int64_t)b * nblk * 2;
  double s0 = 0.0;
  float s1 = PHASE == 0 ? -__builtin_huge_valf() : 0.f;
  for (int i = threadIdx.x; i < nblk; i += kMseThreads) {
    s0 += PHASE == 0 ? (double)part[2 * i] : (double)(int64_t)part[2 * i];
    s1 = PHASE == 0 ? fmaxf(s1, part[2 * i + 1]) : s1 + part[2 * i + 1];
  }
  __shared__ double r0[kMseThreads];
  __shared__ float r1[kMseThreads];
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int w = kMseThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] = PHASE == 0 ? fmaxf(r1[threadIdx.x], r1[threadIdx.x + w]) : r1[threadIdx.x] + r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float* st = a.stats + 4 * b;
    if constexpr (PHASE == 0) {
      st[0] = (float)r0[0];
      st[1] = r1[0];
    } else {
      st[2] = (float)r0[0];
      st[3] = r1[0];
      if (a.loss) a.loss[b] = a.scale * (float)(r0[0] / ((double)r1[0] + kSaEps));
    }
  }
}

// dloss/dD with the mask held constant.
__global__ __launch_bounds__(kMseThreads) void small_area_grad_kernel(SmallAreaArgs a) {
  const int b = blockIdx.y;
  const float* d = a.in + (int64_t)b * a.ibs;
  float* g = a.grad + (int64_t)b * a.gbs;
  const float* st = a.stats + 4 * b;
  const float den = st[0] + kSaEps, imax = st[1] / den;
  const float sm = st[3] + kSaEps;
  const float coef = -a.scale * st[2] / (sm * sm) / den;
  for (int64_t p = (int64_t)blockIdx.x * kMseThreads + threadIdx.x; p < a.npix;
       p += (int64_t)gridDim.x * kMseThreads) {
    const float i = d[p] / den;
    const float m = i > kSaFrac * imax ? 1.f : 0.f;
    g[p] = coef * (m - st[3]);
  }
}

// ---------------------------------------------------------------- focal second moment
struct MomentArgs {
  const float2* e;
  int64_t cs;
  int nc;
  const float* r2;
  int64_t npix;
  float* partial;      // kMseBlocks x 2
  float* stats;        // sum I, L
  float* loss;
  const float* gscale;
  float2* grad;
};

// grid kMseBlocks: per block sums of I and I r2 (fixed order).
__global__ __launch_bounds__(kMseThreads) void moment_kernel(MomentArgs a) {
  float s0 = 0.f, s1 = 0.f;
  for (int64_t p = (int64_t)blockIdx.x * kMseThreads + threadIdx.x; p < a.npix;
       p += (int64_t)gridDim.x * kMseThreads) {
    float i = 0.f;
    for (int c = 0; c < a.nc; ++c) {
      const float2 v = a.e[c * a.cs + p];
      i += v.x * v.x + v.y * v.y;
    }
    s0 += i;
    s1 += i * a.r2[p];
  }
  __shared__ float red[2][kMseThreads / 64];
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s0;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float t0 = 0.f, t1 = 0.f;
    for (int w = 0; w < kMseThreads / 64; ++w) {
      t0 += red[0][w];
      t1 += red[1][w];
    }
    a.partial[2 * blockIdx.x] = t0;
    a.partial[2 * blockIdx.x + 1] = t1;
  }
}

// One block: stats = (S, T / S) from the partials, in a fixed order (float64).
__global__ __launch_bounds__(kMseThreads) void moment_reduce_kernel(MomentArgs a, int nblk) {
  __shared__ double r0[kMseThreads], r1[kMseThreads];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kMseThreads) {
    s0 += a.partial[2 * i];
    s1 += a.partial[2 * i + 1];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int w = kMseThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] += r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float L = (float)(r1[0] / r0[0]);
    a.stats[0] = (float)r0[0];
    a.stats[1] = L;
    if (a.loss) a.loss[0] = L;
  }
}

// grad_c = g 2 (r2 - L) E_c / S (grid-stride over pixels).
__global__ __launch_bounds__(256) void moment_grad_kernel(MomentArgs a) {
  const float g = a.gscale ? a.gscale[0] : 1.f;
  const float S = a.stats[0], L = a.stats[1];
  const float k = 2.f * g / S;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < a.npix; p += (int64_t)gridDim.x * 256) {
    const float w = k * (a.r2[p] - L);
    for (int c = 0; c < a.nc; ++c) {
      const float2 v = a.e[c * a.cs + p];
      a.grad[c * a.cs + p] = make_float2(w * v.x, w * v.y);
    }
  }
}


// --- distance gradient of a propagation --------------------------------------
// dL/dz = sum_c Re(conj(d_c) g_c) over the dz-propagated field d and the output
// cotangent g, plus (VRS) the z-dependence of the constructed Ez_in =
// (Ex' X + Ey' Y)/r, r^2 = X^2 + Y^2 + z^2: sum Re(conj(gez) Ex') X w +
// Re(conj(gez) Ey') Y w with w = -z / r^3 and gez the cotangent of Ez_in.
// One grid-stride pass with per-pixel products in float32 and float64 sums,
// block partials, then a fixed-order tree: the result does not depend on the
// launch.
constexpr int kDzBlocks = 2048;
constexpr int kDzThreads = 256;

struct DzDotArgs {
  const float2* d;
  const float2* g;
  int64_t cs;
  int nc;
  const float2* pre;     // Ex', Ey' (pre_cs apart), or null
  int64_t pre_cs;
  const float2* gez;
  const double* x;
  const double* y;
  int rows, cols;
  double z;
  const double* z_dev;
  double* partial;       // kDzBlocks
  double* out;
};

__device__ __forceinline__ double block_sum_f64(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kDzThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

// Rows are spread over the blocks (grid-stride), columns over the threads,
// so a pixel's row and column need no division.
__global__ __launch_bounds__(kDzThreads) void dz_dot_kernel(DzDotArgs a) {
  __shared__ double red[kDzThreads];
  const float z = (float)(a.z_dev ? a.z_dev[0] : a.z);
  double acc = 0.0;
  for (int i = blockIdx.x; i < a.rows; i += gridDim.x)
  for (int j = threadIdx.x; j < a.cols; j += kDzThreads) {
    const int64_t p = (int64_t)i * a.cols + j;
    float s = 0.f;
    for (int c = 0; c < a.nc; ++c) {
      const float2 dv = a.d[c * a.cs + p], gv = a.g[c * a.cs + p];
      s += dv.x * gv.x + dv.y * gv.y;
    }
    if (a.pre) {
      const float X = (float)a.x[j], Y = (float)a.y[i];
      const float ir = rsqrtf(X * X + Y * Y + z * z);
      const float w = -z * ir * ir * ir;
      const float2 gz = a.gez[p], px = a.pre[p], py = a.pre[a.pre_cs + p];
      s += (gz.x * px.x + gz.y * px.y) * (X * w) + (gz.x * py.x + gz.y * py.y) * (Y * w);
    }
    acc += (double)s;
  }
  const double t = block_sum_f64(acc, red);
  if (threadIdx.x == 0) a.partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kDzThreads) void dz_dot_reduce_kernel(DzDotArgs a, int nblk) {
  __shared__ double red[kDzThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kDzThreads) acc += a.partial[i];
  const double t = block_sum_f64(acc, red);
  if (threadIdx.x == 0) a.out[0] = t;
}

// Quality factor of a propagation (wave_optics.py:187-191) with z read from
// device memory: one thread, so a propagation to a distance held on the
// device reports it without a host round trip.
__global__ void quality_factor_kernel(const double* z, double rmax, double dr_real, double wl, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double zz = z[0];
    out[0] = (sqrt(wl * wl + rmax * rmax + 2.0 * wl * sqrt(rmax * rmax + zz * zz)) - rmax) / dr_real;
  }
}

}  // namespace xl

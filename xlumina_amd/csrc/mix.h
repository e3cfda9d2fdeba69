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
  const double* theta; // theta_mode > 0: splitter angle
  int theta_mode;      // 1 symmetric, 2 / 3 single input at port 0 / 1; + BS_THETA_RAW: theta holds p, t = 2 pi p - pi
  int adjoint;
};

constexpr int BS_THETA_RAW = 4;

// The splitter angle of theta_mode: as stored, or 2 pi p - pi of a raw p
// (two float64 roundings, as the reference's parameter map).
__device__ __forceinline__ double splitter_angle(int mode, double p) {
  return (mode & BS_THETA_RAW) ? __dsub_rn(__dmul_rn(p, 6.283185307179586), 3.141592653589793) : p;
}

// Lossy splitter of angle t: R = |sin t| - 0.01 |cos t|, T = 0.99 |cos t|
// (optical_elements.py:376-383) and its derivatives, in float64.
struct SplitterRT {
  double R, T, dR, dT;
  __device__ explicit SplitterRT(double t) {
    const double c = cos(t), s = sin(t);
    const double sc = (c > 0.0) - (c < 0.0), ss = (s > 0.0) - (s < 0.0);
    const double T0 = fabs(c), noise = T0 * 0.01;
    R = fabs(s) - noise;
    T = T0 - noise;
    dR = ss * c + 0.01 * sc * s;
    dT = -0.99 * sc * s;
  }
};

// (ca, cb, da, db) of the splitter modes: 1 (R, iT, iT, R), 2 (R, 0, iT, 0),
// 3 (iT, 0, R, 0) (optical_elements.py:334-460).
__device__ __forceinline__ void splitter_coef(int mode, double t, float2& ca, float2& cb, float2& da, float2& db) {
  const SplitterRT k(t);
  const float2 r = make_float2((float)k.R, 0.f), it = make_float2(0.f, (float)k.T), z = make_float2(0.f, 0.f);
  if (mode == 1) {
    ca = r; cb = it; da = it; db = r;
  } else if (mode == 2) {
    ca = r; cb = z; da = it; db = z;
  } else {
    ca = it; cb = z; da = r; db = z;
  }
}

__device__ __forceinline__ float2 bs_mac(float2 k, float2 v, float2 acc) {
  return make_float2(acc.x + k.x * v.x - k.y * v.y, acc.y + k.x * v.y + k.y * v.x);
}

// blockIdx.y = component (0, 1: Ex, Ey; 2 with zero_ez: the zeroed Ez of both
// outputs), grid-stride over its npix samples along x.
__global__ __launch_bounds__(256) void bs_kernel(BsArgs a) {
  const float2 zero = make_float2(0.f, 0.f);
  const int comp = blockIdx.y;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (comp == 2) {   // the zeroed Ez planes: no coefficients
    for (; p < a.npix; p += stride) {
      if (a.c) a.c[2 * a.ccs + p] = zero;
      if (a.d) a.d[2 * a.dcs + p] = zero;
    }
    return;   // whole block
  }
  const float2* ap = a.a + comp * a.acs;
  const float2* bp = a.b ? a.b + comp * a.bcs : nullptr;
  float2* cp = a.c ? a.c + comp * a.ccs : nullptr;
  float2* dp = a.d ? a.d + comp * a.dcs : nullptr;
  // this thread's first samples are loaded before the block forms its
  // coefficients, so the loads are in flight behind the float64 set-up
  float2 va = zero, vb = zero;
  if (p < a.npix) {
    va = ap[p];
    if (bp) vb = bp[p];
  }
  float2 ca = a.coef ? a.coef[0] : a.ca, cb = a.coef ? a.coef[1] : a.cb;
  float2 da = a.coef ? a.coef[2] : a.da, db = a.coef ? a.coef[3] : a.db;
  if (a.theta_mode) {   // formed once per block in float64, shared through LDS
    __shared__ float2 k4[4];
    if (threadIdx.x == 0) {
      splitter_coef(a.theta_mode & 3, splitter_angle(a.theta_mode, a.theta[0]), ca, cb, da, db);
      if (a.adjoint) {   // (conj ca, conj da, conj cb, conj db)
        const float2 t = cb;
        ca = make_float2(ca.x, -ca.y);
        cb = make_float2(da.x, -da.y);
        da = make_float2(t.x, -t.y);
        db = make_float2(db.x, -db.y);
      }
      k4[0] = ca; k4[1] = cb; k4[2] = da; k4[3] = db;
    }
    __syncthreads();
    ca = k4[0]; cb = k4[1]; da = k4[2]; db = k4[3];
  }
  for (; p < a.npix; p += stride) {
    if (cp) cp[p] = bs_mac(cb, vb, bs_mac(ca, va, zero));
    if (dp) dp[p] = bs_mac(db, vb, bs_mac(da, va, zero));
    const int64_t q = p + stride;
    if (q < a.npix) {
      va = ap[q];
      vb = bp ? bp[q] : zero;
    }
  }
}

constexpr int kBsDotBlocks = 2048;

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
  const double* theta;
  int theta_mode;
  double* dtheta;
  float2* ga;          // optional: the input cotangents (theta_mode > 0)
  float2* gb;
  int64_t gacs, gbcs;
  int g_planes;        // planes 2 .. g_planes-1 of ga / gb are zeroed
};

// conj(u) v accumulated into acc
__device__ __forceinline__ void cdot_acc(float2 u, float2 v, float& re, float& im) {
  re += u.x * v.x + u.y * v.y;
  im += u.x * v.y - u.y * v.x;
}

// grid kBsDotBlocks: partial[blk] = the block's sums of conj(gc) a, conj(gc) b,
// conj(gd) a, conj(gd) b over both components.  With ga / gb set, the same
// pass writes the input cotangents (the splitter's adjoint) from the gc / gd
// it has loaded.
__global__ __launch_bounds__(256) void bs_dots_kernel(BsDotsArgs a) {
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float2 zero = make_float2(0.f, 0.f);
  const bool fields = a.ga || a.gb;
  float2 ka = zero, kb = zero, kc = zero, kd = zero;   // conj(ca), conj(da), conj(cb), conj(db)
  if (fields) {   // formed once per block in float64, shared through LDS
    __shared__ float2 k4[4];
    if (threadIdx.x == 0) {
      float2 ca, cb, da, db;
      splitter_coef(a.theta_mode & 3, splitter_angle(a.theta_mode, a.theta[0]), ca, cb, da, db);
      k4[0] = make_float2(ca.x, -ca.y);
      k4[1] = make_float2(da.x, -da.y);
      k4[2] = make_float2(cb.x, -cb.y);
      k4[3] = make_float2(db.x, -db.y);
    }
    __syncthreads();
    ka = k4[0]; kb = k4[1]; kc = k4[2]; kd = k4[3];
  }
#pragma unroll 2
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < 2 * a.npix; e += (int64_t)gridDim.x * 256) {
    const int comp = e >= a.npix ? 1 : 0;
    const int64_t p = e - comp * a.npix;
    const float2 va = a.a[comp * a.acs + p];
    const float2 vb = a.b ? a.b[comp * a.bcs + p] : zero;
    const float2 gcv = a.gc ? a.gc[comp * a.gccs + p] : zero;
    const float2 gdv = a.gd ? a.gd[comp * a.gdcs + p] : zero;
    if (a.gc) {
      cdot_acc(gcv, va, s[0], s[1]);
      cdot_acc(gcv, vb, s[2], s[3]);
    }
    if (a.gd) {
      cdot_acc(gdv, va, s[4], s[5]);
      cdot_acc(gdv, vb, s[6], s[7]);
    }
    if (a.ga) a.ga[comp * a.gacs + p] = bs_mac(kb, gdv, bs_mac(ka, gcv, zero));
    if (a.gb) a.gb[comp * a.gbcs + p] = bs_mac(kd, gdv, bs_mac(kc, gcv, zero));
  }
  if (fields) {   // the planes the splitter does not act on
    const int64_t nz = (int64_t)(a.g_planes > 2 ? a.g_planes - 2 : 0) * a.npix;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nz; e += (int64_t)gridDim.x * 256) {
      const int c = 2 + (int)(e / a.npix);
      const int64_t p = e - (int64_t)(c - 2) * a.npix;
      if (a.ga) a.ga[c * a.gacs + p] = zero;
      if (a.gb) a.gb[c * a.gbcs + p] = zero;
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
  if (a.dots && threadIdx.x < 8) a.dots[threadIdx.x] = red[threadIdx.x][0];
  if (a.theta_mode && a.dtheta && threadIdx.x == 0) {
    // dL/dt = Re sum_k S_k dkappa_k/dt with S = (conj(gc) a, conj(gc) b, conj(gd) a, conj(gd) b)
    const SplitterRT k(splitter_angle(a.theta_mode, a.theta[0]));
    const double re0 = red[0][0], im0 = red[1][0], im1 = red[3][0];
    const double re2 = red[4][0], im2 = red[5][0], re3 = red[6][0];
    const int mode = a.theta_mode & 3;
    double g;
    if (mode == 1) g = k.dR * (re0 + re3) - k.dT * (im1 + im2);
    else if (mode == 2) g = k.dR * re0 - k.dT * im2;
    else g = -k.dT * im0 + k.dR * re2;
    a.dtheta[0] = (a.theta_mode & BS_THETA_RAW) ? 6.283185307179586 * g : g;   // dL/dp of a raw angle
  }
}

}  // namespace xl

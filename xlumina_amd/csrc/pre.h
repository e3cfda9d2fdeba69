// Standalone per-pixel pre-ops (Jones mask chain, Ez construction, high-NA
// lens map) and their adjoint.  The forward form is the same arithmetic the
// propagators fuse into their first pass; the adjoint is what the propagators'
// backward runs after the adjoint convolution.
//
// Cotangent convention (PyTorch): for complex y, g = dL/dRe(y) + i dL/dIm(y);
// a linear map y = A x sends g to A^H g; a real parameter p gets
// Re(sum conj(g) dy/dp).
#pragma once
#include "fft.h"
#include "masks.h"
#include "opt.h"
#include "rs.h"

namespace xl {

struct PreArgs {
  const float2* in;
  int64_t in_cs;
  int nin, ncomp;
  const float2* gout;
  float2* out;
  int64_t cs;
  float* gparam[2][2];     // float32 per pixel, or a double* in the float64 parameter forms
  int scal;                // bit 2i+j: gparam[i][j] is a scalar-form gradient, bit 4: ezdz is set (one launch
                           // row, blocks <= its buffer): per-block sums at element 1 + block, reduced into element 0
  double* ezdz;            // XL_EZ_VRS: sum of Re(conj(g_Ez) Ex') X w + Re(conj(g_Ez) Ey') Y w, w = -z / r^3
  int rows, cols, ez_mode, hna;
  int out_planes;          // backward: planes nin .. out_planes-1 of out receive zeros
  const double* x;
  const double* y;
  double z, hna_radius, hna_f, hna_stm2;
  const double* z_dev;   // if set, z is read from device memory
  MaskArgs mask;
  int64_t in_bs, gout_bs, out_bs, mask_bs, gparam_bs;   // batch strides (blockIdx.y = member)
};

// The arguments of batch member b (pointers offset by the batch strides).
__device__ __forceinline__ PreArgs pre_member(const PreArgs& a0, int b) {
  PreArgs a = a0;
  a.in += b * a0.in_bs;
  if (a.z_dev) a.z_dev += b;   // one distance per member (b = 0 without a batch)
  if (a.gout) a.gout += b * a0.gout_bs;
  if (a.out) a.out += b * a0.out_bs;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (a.mask.st[i].m0) a.mask.st[i].m0 += b * a0.mask_bs;
    if (a.mask.st[i].m1) a.mask.st[i].m1 += b * a0.mask_bs;
#pragma unroll
    for (int j = 0; j < 2; ++j)
      if (a.gparam[i][j]) a.gparam[i][j] += b * a0.gparam_bs;
  }
  return a;
}

// Richards-Wolf lens factors (optical_elements.py:528-590): the 3x2 real map
// (Ex', Ey') -> lens output, with Ez = (Ex' X + Ey' Y)/rho folded in.
__device__ __forceinline__ void hna_map(double X, double Y, double radius, double f, double stm2, float A[3][2]) {
  // rho = 0 (odd N) gives cp = sp = NaN through 0 * inf, as X / rho does in the reference
  const double rho2 = X * X + Y * Y;
  const double irho = rsqrt_f64(rho2);
  const double rho = rho2 * irho;
  const float cp = (float)(X * irho), sp = (float)(Y * irho);
  float st, ct;
  sincos_rad((float)(rho / f), &st, &ct);
  const double u2 = rho2 / (radius * radius);
  const float G = u2 < 1.0 ? (float)rsqrt_f64(fabs(1.0 - u2 * stm2)) : 0.f;
  const float w = sqrtf(fabsf(ct)) * G;
  const float R[3][3] = {{ct * cp * cp + sp * sp, ct * cp * sp - sp * cp, -cp * st},
                         {sp * ct * cp - cp * sp, ct * sp * sp + cp * cp, -sp * st},
                         {st * cp, st * sp, ct}};
  for (int i = 0; i < 3; ++i) {
    A[i][0] = w * (R[i][0] + R[i][2] * cp);
    A[i][1] = w * (R[i][1] + R[i][2] * sp);
  }
}

__device__ __forceinline__ float re_conj_mul(float2 g, float2 v) { return g.x * v.x + g.y * v.y; }  // Re(conj(g) v)

__global__ void pre_forward_kernel(PreArgs a0) {
  const PreArgs a = pre_member(a0, blockIdx.y);
  const int64_t npix = (int64_t)a.rows * a.cols;
  for (int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pix < npix;
       pix += (int64_t)gridDim.x * blockDim.x) {
    float2 e0 = a.in[pix];
    float2 e1 = a.nin > 1 ? a.in[a.in_cs + pix] : make_float2(0.f, 0.f);
    float2 e2 = a.nin > 2 ? a.in[2 * a.in_cs + pix] : make_float2(0.f, 0.f);
    apply_mask(a.mask, pix, e0, e1, e2);
    if (a.hna) {
      const int row = (int)(pix / a.cols), col = (int)(pix % a.cols);
      float A[3][2];
      hna_map(a.x[col], a.y[row], a.hna_radius, a.hna_f, a.hna_stm2, A);
      for (int i = 0; i < 3; ++i) a.out[i * a.cs + pix] = c_add(c_scale(e0, A[i][0]), c_scale(e1, A[i][1]));
      continue;
    }
    if (a.ez_mode != EZ_NONE) {
      const int row = (int)(pix / a.cols), col = (int)(pix % a.cols);
      e2 = build_ez(a.ez_mode, e0, e1, a.x[col], a.y[row], a.z_dev ? *a.z_dev : a.z);
    }
    a.out[pix] = e0;
    if (a.ncomp > 1) a.out[a.cs + pix] = e1;
    if (a.ncomp > 2) a.out[2 * a.cs + pix] = e2;
  }
}

// adjoint of one mask stage: e_in -> (stage) -> e_out; g holds the cotangent of
// e_out on entry and of e_in on exit; the cotangents of the parameter values
// are formed in d0 / d1 where want0 / want1 ask for them.
__device__ __forceinline__ void stage_backward(const MaskStage& m, int64_t pix, float2 e0, float2 e1, float2& g0,
                                               float2& g1, float2& g2, bool want0, bool want1, float& d0, float& d1) {
  switch (m.kind) {
    case MASK_SLM: {
      const float2 w = c_expi(mask_param(m, 0, pix));
      const float2 o = c_mul(e0, w);
      if (want0) d0 = re_conj_mul(g0, make_float2(-o.y, o.x));   // d/dphi = i o
      g0 = c_mul(g0, make_float2(w.x, -w.y));
      break;
    }
    case MASK_SSLM: {
      const float2 w0 = c_expi(mask_param(m, 0, pix)), w1 = c_expi(mask_param(m, 1, pix));
      const float2 o0 = c_mul(e0, w0), o1 = c_mul(e1, w1);
      if (want0) d0 = re_conj_mul(g0, make_float2(-o0.y, o0.x));
      if (want1) d1 = re_conj_mul(g1, make_float2(-o1.y, o1.x));
      g0 = c_mul(g0, make_float2(w0.x, -w0.y));
      g1 = c_mul(g1, make_float2(w1.x, -w1.y));
      break;
    }
    case MASK_LCD:
    case MASK_LCD_SCALAR: {
      const float eta = m.kind == MASK_LCD ? mask_param(m, 0, pix) : (float)m.s0;
      const float th = m.kind == MASK_LCD ? mask_param(m, 1, pix) : (float)m.s1;
      float sh, ch, st, ct;
      sincos_rad(0.5f * eta, &sh, &ch);
      sincos_rad(th, &st, &ct);
      const float c2 = ct * ct, s2 = st * st, sc = st * ct;
      const float2 j00 = make_float2(ch * (c2 + s2), -sh * c2 + sh * s2);
      const float2 j11 = make_float2(ch * (s2 + c2), -sh * s2 + sh * c2);
      const float2 j01 = make_float2(0.f, -2.f * sh * sc);
      // d/deta
      const float2 a00 = make_float2(-0.5f * sh * (c2 + s2), 0.5f * ch * (s2 - c2));
      const float2 a11 = make_float2(-0.5f * sh * (s2 + c2), 0.5f * ch * (c2 - s2));
      const float2 a01 = make_float2(0.f, -ch * sc);
      // d/dtheta: c^2 -> -2sc, s^2 -> 2sc, sc -> c^2 - s^2
      const float2 b00 = make_float2(ch * (-2.f * sc + 2.f * sc), 4.f * sh * sc);
      const float2 b11 = make_float2(ch * (2.f * sc - 2.f * sc), -4.f * sh * sc);
      const float2 b01 = make_float2(0.f, -2.f * sh * (c2 - s2));
      if (want0)
        d0 = re_conj_mul(g0, c_add(c_mul(a00, e0), c_mul(a01, e1))) +
                   re_conj_mul(g1, c_add(c_mul(a01, e0), c_mul(a11, e1)));
      if (want1)
        d1 = re_conj_mul(g0, c_add(c_mul(b00, e0), c_mul(b01, e1))) +
                   re_conj_mul(g1, c_add(c_mul(b01, e0), c_mul(b11, e1)));
      const float2 n0 = c_add(c_mul(conjf2(j00), g0), c_mul(conjf2(j01), g1));
      const float2 n1 = c_add(c_mul(conjf2(j01), g0), c_mul(conjf2(j11), g1));
      g0 = n0;
      g1 = n1;
      break;
    }
    case MASK_LP: {
      float s, c;
      sincos_rad(mask_param(m, 0, pix), &s, &c);
      if (want0) {
        const float d00 = -2.f * s * c, d01 = c * c - s * s, d11 = 2.f * s * c;
        d0 = re_conj_mul(g0, c_add(c_scale(e0, d00), c_scale(e1, d01))) +
                   re_conj_mul(g1, c_add(c_scale(e0, d01), c_scale(e1, d11)));
      }
      const float2 n0 = c_add(c_scale(g0, c * c), c_scale(g1, c * s));
      const float2 n1 = c_add(c_scale(g0, c * s), c_scale(g1, s * s));
      g0 = n0;
      g1 = n1;
      g2 = make_float2(0.f, 0.f);
      break;
    }
    case MASK_AMP: {   // e' = A e: dA = Re(conj(g) e), g <- A g
      const float a0 = mask_param(m, 0, pix);
      if (want0) d0 = re_conj_mul(g0, e0);
      g0 = c_scale(g0, a0);
      if (m.m1) {
        const float a1 = mask_param(m, 1, pix);
        if (want1) d1 = re_conj_mul(g1, e1);
        g1 = c_scale(g1, a1);
      }
      break;
    }
    default:
      break;
  }
}

// The pixel's row and column (32-bit arithmetic where the plane allows it).
__device__ __forceinline__ void pix_row_col(int64_t pix, int cols, bool small, int& row, int& col) {
  if (small) {
    const int ip = (int)pix;
    row = ip / cols;
    col = ip - row * cols;
  } else {
    row = (int)(pix / cols);
    col = (int)(pix - (int64_t)row * cols);
  }
}

// MODE (the host's choice from ez_mode / hna): 0 no Ez construction or lens
// map, 1 an Ez construction, 2 the high-NA lens map; each form compiles only
// its own branch, which keeps the kernel arguments in scalar registers.
template <int MODE>
__global__ __launch_bounds__(256) void pre_backward_kernel(PreArgs a0) {
  const PreArgs a = pre_member(a0, blockIdx.y);
  const int64_t npix = (int64_t)a.rows * a.cols;
  const bool small = npix < 0x7fffffff;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  double acc_ez = 0.0;
  for (int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pix < npix;
       pix += (int64_t)gridDim.x * blockDim.x) {
    // forward recompute of the stage inputs (the second stage's input only
    // when there is a second stage)
    float2 x0 = a.in[pix];
    float2 x1 = a.nin > 1 ? a.in[a.in_cs + pix] : make_float2(0.f, 0.f);
    float2 x2 = a.nin > 2 ? a.in[2 * a.in_cs + pix] : make_float2(0.f, 0.f);
    float2 m0 = x0, m1 = x1, m2 = x2;
    if (a.mask.st[1].kind != MASK_NONE) apply_stage(a.mask.st[0], pix, m0, m1, m2);
    // cotangent of the masked (Ex', Ey', Ez')
    float2 g0 = a.gout[pix];
    float2 g1 = a.ncomp > 1 ? a.gout[a.cs + pix] : make_float2(0.f, 0.f);
    float2 g2 = a.ncomp > 2 ? a.gout[2 * a.cs + pix] : make_float2(0.f, 0.f);
    if constexpr (MODE == 2) {
      int row, col;
      pix_row_col(pix, a.cols, small, row, col);
      float A[3][2];
      hna_map(a.x[col], a.y[row], a.hna_radius, a.hna_f, a.hna_stm2, A);
      const float2 h0 = c_add(c_add(c_scale(g0, A[0][0]), c_scale(g1, A[1][0])), c_scale(g2, A[2][0]));
      const float2 h1 = c_add(c_add(c_scale(g0, A[0][1]), c_scale(g1, A[1][1])), c_scale(g2, A[2][1]));
      g0 = h0; g1 = h1; g2 = make_float2(0.f, 0.f);
    } else if constexpr (MODE == 1) {
      int row, col;
      pix_row_col(pix, a.cols, small, row, col);
      const double X = a.x[col], Y = a.y[row], z = a.z_dev ? *a.z_dev : a.z;
      float cx, cy;
      if (a.ez_mode == EZ_HNA) {
        const double rho = sqrt(X * X + Y * Y);
        cx = (float)(X / rho); cy = (float)(Y / rho);
      } else {
        const double r = sqrt(X * X + Y * Y + z * z);
        double dx = X / r, dy = Y / r;
        if (a.ez_mode == EZ_VCZT) { dx *= z / r; dy *= z / r; }
        cx = (float)dx; cy = (float)dy;
        if (a.ezdz && a.ez_mode == EZ_VRS) {
          // d Ez_in / dz = -(Ex' X + Ey' Y) z / r^3 through Ez_in's cotangent (vectorized_optics.py:258-261),
          // with (Ex', Ey') the fully masked input (the last stage applied to m)
          float2 f0 = m0, f1 = m1, f2 = m2;
          apply_stage(a.mask.st[a.mask.st[1].kind != MASK_NONE ? 1 : 0], pix, f0, f1, f2);
          const double w = -z / (r * r * r);
          acc_ez += (double)(re_conj_mul(g2, f0) * (float)(X * w) + re_conj_mul(g2, f1) * (float)(Y * w));
        }
      }
      g0 = c_add(g0, c_scale(g2, cx));
      g1 = c_add(g1, c_scale(g2, cy));
      g2 = make_float2(0.f, 0.f);
    }
    float d[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    stage_backward(a.mask.st[1], pix, m0, m1, g0, g1, g2, a.gparam[1][0], a.gparam[1][1], d[1][0], d[1][1]);
    stage_backward(a.mask.st[0], pix, x0, x1, g0, g1, g2, a.gparam[0][0], a.gparam[0][1], d[0][0], d[0][1]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (!a.gparam[i][j]) continue;
        const int f = param_form(a.mask.st[i], j);
        if (!(f & PF_F64))
          a.gparam[i][j][pix] = d[i][j];
        else if (f & PF_SCALAR)
          acc[i][j] += (double)d[i][j];
        else
          reinterpret_cast<double*>(a.gparam[i][j])[pix] = (f & PF_UNIT) ? 6.283185307179586 * d[i][j] : d[i][j];
      }
    if (a.out) {
      a.out[pix] = g0;
      if (a.nin > 1) a.out[a.in_cs + pix] = g1;
      if (a.nin > 2) a.out[2 * a.in_cs + pix] = g2;
      for (int c = a.nin; c < a.out_planes; ++c) a.out[c * a.in_cs + pix] = make_float2(0.f, 0.f);
    }
  }
  // scalar-form parameters: this block's sum (times 2 pi for PF_UNIT) at
  // element 1 + block of the parameter's buffer; pgrad_reduce_kernel adds them
  if (a.scal) {
    __shared__ double red[5][4];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (!(a.scal >> k & 1)) continue;
      double v = k < 4 ? acc[k >> 1][k & 1] : acc_ez;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        if (!(a.scal >> k & 1)) continue;
        const double v = red[k][0] + red[k][1] + red[k][2] + red[k][3];
        if (k == 4) {
          a.ezdz[1 + blockIdx.x] = v;
        } else {
          const double sc = (param_form(a.mask.st[k >> 1], k & 1) & PF_UNIT) ? 6.283185307179586 : 1.0;
          reinterpret_cast<double*>(a.gparam[k >> 1][k & 1])[1 + blockIdx.x] = sc * v;
        }
      }
    }
  }
}

// buf[i][0] = sum of buf[i][1 .. n] for the scalar-form parameter gradients
// and the Ez_in distance term (block i; null pointers skip), fixed order;
// base4[0 .. nbase), if set, is added to the Ez_in term's sum.
__global__ __launch_bounds__(1024) void pgrad_reduce_kernel(double* b0, double* b1, double* b2, double* b3, double* b4,
                                                            const double* base4, int n, int nbase) {
  double* b = blockIdx.x == 0 ? b0 : blockIdx.x == 1 ? b1 : blockIdx.x == 2 ? b2 : blockIdx.x == 3 ? b3 : b4;
  if (!b) return;   // whole block
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += b[1 + i];
  if (blockIdx.x == 4 && base4)
    for (int i = threadIdx.x; i < nbase; i += 1024) s += base4[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < 16; ++w) v += red[w];
    b[0] = v;
  }
}

// Adjoint of a lone SLM on one component (the 4f optimiser's and the scalar
// requests' mask): e' = e w, w = e^{i phi}, so dL/dphi = Re(conj(g) i e') and
// the input cotangent is conj(w) g -- the arithmetic of stage_backward's
// MASK_SLM case.  A pure stream (20 B in, up to 12 B out per pixel): each
// thread issues the loads of kSlmUnroll pixels before any arithmetic, so that
// enough bytes are in flight to run at HBM rate.
constexpr int kSlmUnroll = 4;

__global__ __launch_bounds__(256) void pre_slm_backward_kernel(PreArgs a0) {
  const PreArgs a = pre_member(a0, blockIdx.y);
  const int64_t npix = (int64_t)a.rows * a.cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float* ph = a.mask.st[0].m0;
  float* gp = a.gparam[0][0];
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; base < npix; base += kSlmUnroll * stride) {
    float2 e[kSlmUnroll], g[kSlmUnroll];
    float q[kSlmUnroll];
#pragma unroll
    for (int u = 0; u < kSlmUnroll; ++u) {
      const int64_t pix = base + u * stride;
      if (pix < npix) {
        e[u] = a.in[pix];
        g[u] = a.gout[pix];
        q[u] = ph[pix];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlmUnroll; ++u) {
      const int64_t pix = base + u * stride;
      if (pix < npix) {
        const float2 w = c_expi(q[u]);
        const float2 o = c_mul(e[u], w);
        if (gp) gp[pix] = re_conj_mul(g[u], make_float2(-o.y, o.x));
        if (a.out) a.out[pix] = c_mul(g[u], make_float2(w.x, -w.y));
      }
    }
  }
}

// The same adjoint fused with the Adam / AdamW step of the phases it
// differentiates (the mask's m0, updated in place; m and v laid out like the
// phases, batch stride mask_bs): the phase cotangent stays in registers, so a
// training step neither writes nor re-reads a gradient array.  The step uses
// t = count + 1 and leaves count alone (adam_tick_kernel advances it once per
// training step).  gparam[0][0], when set, still receives the cotangent.
__global__ __launch_bounds__(256) void pre_slm_backward_adam_kernel(PreArgs a0, AdamArgs ad) {
  const PreArgs a = pre_member(a0, blockIdx.y);
  const int64_t npix = (int64_t)a.rows * a.cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float* ph = const_cast<float*>(a.mask.st[0].m0);
  float* mb = ad.m + blockIdx.y * a0.mask_bs;
  float* vb = ad.v + blockIdx.y * a0.mask_bs;
  float* gp = a.gparam[0][0];
  float c1, c2;
  adam_factors(ad, c1, c2);
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; base < npix; base += kSlmUnroll * stride) {
    float2 e[kSlmUnroll], g[kSlmUnroll];
    float q[kSlmUnroll], m[kSlmUnroll], v[kSlmUnroll];
#pragma unroll
    for (int u = 0; u < kSlmUnroll; ++u) {
      const int64_t pix = base + u * stride;
      if (pix < npix) {
        e[u] = a.in[pix];
        g[u] = a.gout[pix];
        q[u] = ph[pix];
        m[u] = mb[pix];
        v[u] = vb[pix];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlmUnroll; ++u) {
      const int64_t pix = base + u * stride;
      if (pix < npix) {
        const float2 w = c_expi(q[u]);
        const float2 o = c_mul(e[u], w);
        const float gq = re_conj_mul(g[u], make_float2(-o.y, o.x));
        if (gp) gp[pix] = gq;
        if (a.out) a.out[pix] = c_mul(g[u], make_float2(w.x, -w.y));
        float pq = q[u];
        adam_update(ad, c1, c2, pq, gq, m[u], v[u]);
        ph[pix] = pq;
        mb[pix] = m[u];
        vb[pix] = v[u];
      }
    }
  }
}

__global__ void transfer_function_kernel(float2* out, const double* x, const double* y, int nx, int ny, double z,
                                         double k) {
  const int64_t npix = (int64_t)nx * ny;
  for (int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pix < npix;
       pix += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(pix / nx), col = (int)(pix % nx);
    out[pix] = rs_kernel_value(x[col], y[row], z, k);
  }
}

}  // namespace xl

// Per-pixel Jones masks and Ez constructions, fused into the first pass of the
// propagators and also run standalone.
//
//   SLM   optical_elements.py:87-103   u * e^{i phi}
//   sSLM  optical_elements.py:141-152, :186-222   diag(e^{i alpha}, e^{i phi}), Ez kept
//   LCD   optical_elements.py:122-139, :169-180, :266-305   general retarder, delta = 0,
//         eta/theta per pixel (XL_MASK_LCD) or constant (XL_MASK_LCD_SCALAR), Ez kept
//   LP    optical_elements.py:109-120, :307-332   [[c^2, cs], [cs, s^2]], Ez -> 0
// Ez constructions (incoming Ez discarded):
//   VRS   vectorized_optics.py:258-261   (Ex X + Ey Y) / r,  r = sqrt(X^2+Y^2+z^2)
//   VCZT  vectorized_optics.py:341-344   (Ex X / r + Ey Y / r) z / r
//   HNA   optical_elements.py:532-538   (Ex X + Ey Y) / rho
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xl {

enum { MASK_NONE = 0, MASK_SLM = 1, MASK_SSLM = 2, MASK_LCD = 3, MASK_LCD_SCALAR = 4, MASK_LP = 5, MASK_AMP = 6 };
enum { EZ_NONE = 0, EZ_VRS = 1, EZ_VCZT = 2, EZ_HNA = 3 };

// Storage forms of a stage parameter (3 bits each, m0 in bits 0-2, m1 in
// bits 3-5 of MaskStage::pf): 0 float32 per pixel; PF_F64 float64 per pixel;
// PF_F64 | PF_SCALAR one float64 shared by every pixel; PF_UNIT (with PF_F64)
// the value used is 2 pi p - pi of the stored p, formed in float64 as the
// reference's parameter map is (optical_elements.py:1556-1559).
enum { PF_F64 = 1, PF_SCALAR = 2, PF_UNIT = 4 };

struct MaskStage {
  int kind;
  int pf;            // storage forms of m0 / m1
  const float* m0;   // phase / eta / alpha (a const double* in the float64 forms)
  const float* m1;   // phi / theta
  double s0, s1;     // constant eta, theta (MASK_LCD_SCALAR)
};

__device__ __forceinline__ int param_form(const MaskStage& m, int j) { return (m.pf >> (3 * j)) & 7; }

// The value of a float64-stored parameter p (two roundings, as p * 2pi - pi
// in float64 without contraction, then one to float32).
__device__ __forceinline__ float param_value(double p, int f) {
  if (f & PF_UNIT) p = __dsub_rn(__dmul_rn(p, 6.283185307179586), 3.141592653589793);
  return (float)p;
}

// Parameter j (0: m0, 1: m1) of stage m at pixel pix.
__device__ __forceinline__ float mask_param(const MaskStage& m, int j, int64_t pix) {
  const float* p = j ? m.m1 : m.m0;
  const int f = param_form(m, j);
  if (!(f & PF_F64)) return p[pix];
  return param_value(reinterpret_cast<const double*>(p)[(f & PF_SCALAR) ? 0 : pix], f);
}
// Up to two masks applied in order (e.g. sSLM then LCD).
struct MaskArgs {
  MaskStage st[2];
};

__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// sin and cos of an angle in radians through sincospif: the angle becomes
// half-turns in float32 (one rounding of x / pi), which avoids the
// large-argument reduction path of sincosf.  For |x| <= 2 pi the added phase
// error is below 1e-6 rad.
__device__ __forceinline__ void sincos_rad(float x, float* s, float* c) {
  sincospif(x * 0.31830988618379067f, s, c);
}
__device__ __forceinline__ float2 c_expi(float p) {
  float s, c;
  sincos_rad(p, &s, &c);
  return make_float2(c, s);
}

// Retarder Jones entries for (eta, theta), delta = 0 (optical_elements.py:134-139):
//   J00 = e^{-i eta/2} c^2 + e^{i eta/2} s^2, J01 = J10 = -2 i sin(eta/2) s c,
//   J11 = e^{-i eta/2} s^2 + e^{i eta/2} c^2.
__device__ __forceinline__ void lcd_jones(float eta, float theta, float2& j00, float2& j01, float2& j11) {
  float sh, ch, st, ct;
  sincos_rad(0.5f * eta, &sh, &ch);
  sincos_rad(theta, &st, &ct);
  const float c2 = ct * ct, s2 = st * st, sc = st * ct;
  j00 = make_float2(ch * (c2 + s2), -sh * c2 + sh * s2);
  j11 = make_float2(ch * (s2 + c2), -sh * s2 + sh * c2);
  j01 = make_float2(0.f, -2.f * sh * sc);
}

// Per-pixel parameters a stage reads (m0 / m1 arrays).
__device__ __forceinline__ bool stage_m0(const MaskStage& m) {
  return m.kind == MASK_SLM || m.kind == MASK_SSLM || m.kind == MASK_LCD || m.kind == MASK_LP || m.kind == MASK_AMP;
}
__device__ __forceinline__ bool stage_m1(const MaskStage& m) {
  return m.kind == MASK_SSLM || m.kind == MASK_LCD || (m.kind == MASK_AMP && m.m1);
}

// One stage; p0() / p1() give the stage's per-pixel parameters (m0[pix],
// m1[pix]) and are called only by the kinds that have them, so a direct
// caller loads inside the case and a caller that loaded ahead passes values.
template <class P0, class P1>
__device__ __forceinline__ void stage_apply(const MaskStage& m, P0 p0, P1 p1, float2& e0, float2& e1, float2& e2) {
  switch (m.kind) {
    case MASK_SLM:
      e0 = c_mul(e0, c_expi(p0()));
      break;
    case MASK_SSLM:
      e0 = c_mul(e0, c_expi(p0()));
      e1 = c_mul(e1, c_expi(p1()));
      break;
    case MASK_LCD:
    case MASK_LCD_SCALAR: {
      float2 j00, j01, j11;
      const float eta = m.kind == MASK_LCD ? p0() : (float)m.s0;
      const float th = m.kind == MASK_LCD ? p1() : (float)m.s1;
      lcd_jones(eta, th, j00, j01, j11);
      const float2 a = e0, b = e1;
      e0 = c_add(c_mul(j00, a), c_mul(j01, b));
      e1 = c_add(c_mul(j01, a), c_mul(j11, b));
      break;
    }
    case MASK_LP: {
      float s, c;
      sincos_rad(p0(), &s, &c);
      const float2 a = e0, b = e1;
      e0 = c_add(c_scale(a, c * c), c_scale(b, c * s));
      e1 = c_add(c_scale(a, c * s), c_scale(b, s * s));
      e2 = make_float2(0.f, 0.f);
      break;
    }
    case MASK_AMP:
      e0 = c_scale(e0, p0());
      if (m.m1) e1 = c_scale(e1, p1());
      break;
    default:
      break;
  }
}

__device__ __forceinline__ void apply_stage(const MaskStage& m, int64_t pix, float2& e0, float2& e1, float2& e2) {
  stage_apply(m, [&] { return mask_param(m, 0, pix); }, [&] { return mask_param(m, 1, pix); }, e0, e1, e2);
}

// A stage whose parameters were loaded ahead (v0 = m0[pix], v1 = m1[pix]).
__device__ __forceinline__ void apply_stage_vals(const MaskStage& m, float v0, float v1, float2& e0, float2& e1,
                                                 float2& e2) {
  stage_apply(m, [&] { return v0; }, [&] { return v1; }, e0, e1, e2);
}

__device__ __forceinline__ void apply_mask(const MaskArgs& m, int64_t pix, float2& e0, float2& e1, float2& e2) {
  apply_stage(m.st[0], pix, e0, e1, e2);
  apply_stage(m.st[1], pix, e0, e1, e2);
}

// The pre-op inputs of U pixels (field components and per-pixel mask
// parameters), loaded apart from their use so that the loads of all U pixels
// issue back to back and are in flight together, instead of one pixel's round
// trip at a time.  Each input kind is loaded for all U pixels under one
// uniform branch (an absent input issues no load and reads as zero).
struct PixRaw {
  float2 e[3];
  float v[2][2];   // [stage][m0, m1]
};

template <int U>
__device__ __forceinline__ void pix_load_n(const float2* in, int64_t cs, int nin, const MaskArgs& m,
                                           const int64_t (&pix)[U], PixRaw (&r)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    r[u].e[1] = r[u].e[2] = make_float2(0.f, 0.f);
    r[u].v[0][0] = r[u].v[0][1] = r[u].v[1][0] = r[u].v[1][1] = 0.f;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) r[u].e[0] = in[pix[u]];
  if (nin > 1) {
#pragma unroll
    for (int u = 0; u < U; ++u) r[u].e[1] = in[cs + pix[u]];
  }
  if (nin > 2) {
#pragma unroll
    for (int u = 0; u < U; ++u) r[u].e[2] = in[2 * cs + pix[u]];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (!(j ? stage_m1(m.st[i]) : stage_m0(m.st[i]))) continue;
      const float* p = j ? m.st[i].m1 : m.st[i].m0;
      const int f = param_form(m.st[i], j);
      if (!(f & PF_F64)) {
#pragma unroll
        for (int u = 0; u < U; ++u) r[u].v[i][j] = p[pix[u]];
      } else if (f & PF_SCALAR) {
        const float v = param_value(reinterpret_cast<const double*>(p)[0], f);
#pragma unroll
        for (int u = 0; u < U; ++u) r[u].v[i][j] = v;
      } else {
        double d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = reinterpret_cast<const double*>(p)[pix[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u].v[i][j] = param_value(d[u], f);
      }
    }
}

// The masked field of pixel pix, loaded and masked in one go (for loops that
// gain nothing from batching the loads).
__device__ __forceinline__ void load_masked(const float2* in, int64_t cs, int nin, const MaskArgs& m, int64_t pix,
                                            float2& e0, float2& e1, float2& e2) {
  e0 = in[pix];
  e1 = nin > 1 ? in[cs + pix] : make_float2(0.f, 0.f);
  e2 = nin > 2 ? in[2 * cs + pix] : make_float2(0.f, 0.f);
  apply_mask(m, pix, e0, e1, e2);
}

// The masked components of a loaded pixel (as apply_mask on the loaded field).
__device__ __forceinline__ void pix_masked(const MaskArgs& m, const PixRaw& r, float2& e0, float2& e1, float2& e2) {
  e0 = r.e[0];
  e1 = r.e[1];
  e2 = r.e[2];
  apply_stage_vals(m.st[0], r.v[0][0], r.v[0][1], e0, e1, e2);
  apply_stage_vals(m.st[1], r.v[1][0], r.v[1][1], e0, e1, e2);
}

// The direction cosines X/r, Y/r (and z/r) are formed in float32 with one
// reciprocal square root: they multiply float32 fields, and their relative
// error (a few float32 ulps) stays far below the field's own rounding budget.
__device__ __forceinline__ float2 build_ez(int mode, float2 ex, float2 ey, double X, double Y, double z) {
  if (mode == EZ_VRS || mode == EZ_VCZT) {
    const float xf = (float)X, yf = (float)Y, zf = (float)z;
    const float ir = rsqrtf(fmaf(xf, xf, fmaf(yf, yf, zf * zf)));
    float cx = xf * ir, cy = yf * ir;
    if (mode == EZ_VCZT) {
      const float cz = zf * ir;
      cx *= cz;
      cy *= cz;
    }
    return c_add(c_scale(ex, cx), c_scale(ey, cy));
  }
  // EZ_HNA: (Ex X + Ey Y) / rho; rho = 0 gives NaN as in the reference
  const double rho = sqrt(X * X + Y * Y);
  return c_add(c_scale(ex, (float)(X / rho)), c_scale(ey, (float)(Y / rho)));
}

}  // namespace xl

// Chirp z-transform (Bluestein) propagation: CZT / VCZT / high-NA VCZT and
// their adjoints, plain FP32 form.
//
// Reference: Bluestein_method / compute_fft (xlumina/wave_optics.py:385-460),
// CZT_jit (:333-357), CZT_for_high_NA_jit (:359-371), VCZT_jit
// (vectorized_optics.py:375-384), VCZT_objective_lens
// (optical_elements.py:600-638).  Per axis the reference computes
//   out_k = post_k * sum_j g[woff + k - j] * pre_j * x_j
// by a length-Q circular convolution (Q = 2^ceil(log2(m + M - 1))), with
//   pre_j  = A^{-j} W^{j^2/2},  g_s = W^{-(s-m+1)^2/2} (s < m+M; the reference's
//   table lacks s = m+M-1 when M >= m), post_k = W^{k^2/2} e^{-i2pi l_k(1/2-m/2)/Dm},
// woff = m.  The tables are generated on the device with float64 phases
// (czt_setup_kernel); the FFT of g is stored pre-divided by Q.  The adjoint axis
// uses pre' = conj(post), post' = conj(pre), g'[d mod Q] = conj(g[m-d]), woff = 0.
//
// The 2-D transform is separable (K_y E K_x^T), so the row pass runs along x on
// the row-major input (contiguous loads) and writes a 4x4-tiled intermediate
// ordered [x-output group][row group]; the column pass reads it back column-
// contiguously, runs the y axis and applies the output-plane factor.
#pragma once
#include "fft.h"
#include "masks.h"
#include "rs.h"
#include "pre.h"

namespace xl {

enum { PRE_NONE = 0, PRE_RS = 1, PRE_HNA = 2, PRE_CONJ_RS = 3, PRE_CONST = 4 };
enum { POST_NONE = 0, POST_RS = 1, POST_CONST = 2, POST_CONJ_RS = 3 };

struct CztAxis {
  int lin, lout, woff, q;
  const float2* pre;   // lin
  const float2* gf;    // q, FFT of the kernel divided by q
  const float2* post;  // lout
};

struct CztArgs {
  const float2* in;
  int64_t in_cs;
  int nin;
  float2* out;
  int64_t out_cs;
  float2* t;           // tiles: ncomp x round4(rin) x round4(ax.lout)
  int rin, ncomp;      // input rows (cols = ax.lin)
  int pre_mode, post_mode, ez_mode;
  const double* px;    // grid of the input plane (pre-op)
  const double* py;
  const double* qx;    // grid of the output plane (post-op)
  const double* qy;
  double z, k;
  float2 cte;          // constant factor (forward) or its conjugate (adjoint)
  double hna_radius, hna_f, hna_stm2;
  MaskArgs mask;
  CztAxis ax, ay;
};

struct CztSetupArgs {
  float2* pre;
  float2* gf;
  float2* post;
  int m, M, q, adjoint;
  double d1_dm;        // D1 / Dm
  double dl_dm;        // Delta / Dm,  Delta = (D2 - D1) / M
  double th_turns;     // angle(W) / (2 pi), principal branch
};

// One workgroup of q/16 threads builds one axis' tables.
template <int Q>
__global__ __launch_bounds__(Q / 16) void czt_setup_kernel(CztSetupArgs a) {
  constexpr int TPF = Q / 16;
  __shared__ float2 s[LdsLen<Q>::value];
  const int t = threadIdx.x;
  const double th = a.th_turns;
  const double shift = 0.5 - 0.5 * a.m;
  // forward tables
  auto pre_f = [&](int j) {   // A^{-j} W^{j^2/2}
    const double jj = (double)j;
    return expi_turns(-jj * a.d1_dm + th * 0.5 * jj * jj);
  };
  auto post_f = [&](int k) {  // W^{k^2/2} e^{-i 2pi l_k (1/2 - m/2)/Dm}
    const double kk = (double)k;
    return expi_turns(th * 0.5 * kk * kk - (a.d1_dm + kk * a.dl_dm) * shift);
  };
  auto g_f = [&](int sidx) -> float2 {   // W^{-(s-m+1)^2/2}, s < m+M, last tap dropped if M >= m
    if (sidx < 0 || sidx >= a.m + a.M) return make_float2(0.f, 0.f);
    if (a.M >= a.m && sidx == a.m + a.M - 1) return make_float2(0.f, 0.f);
    const double u = (double)(sidx - a.m + 1);
    return expi_turns(-th * 0.5 * u * u);
  };
  if (!a.adjoint) {
    for (int j = t; j < a.m; j += TPF) a.pre[j] = pre_f(j);
    for (int k = t; k < a.M; k += TPF) a.post[k] = post_f(k);
    for (int i = t; i < Q; i += TPF) s[lpad(i)] = g_f(i);
  } else {
    for (int k = t; k < a.M; k += TPF) a.pre[k] = conjf2(post_f(k));
    for (int j = t; j < a.m; j += TPF) a.post[j] = conjf2(pre_f(j));
    for (int i = t; i < Q; i += TPF) {
      const int d = i < a.m ? i : i - Q;   // lag d in [-(Q-m), m)
      s[lpad(i)] = (d > -a.M) ? conjf2(g_f(a.m - d)) : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  fft_lds<Q, TPF, false>(s, t);
  const float iq = 1.0f / (float)Q;
  for (int i = t; i < Q; i += TPF) a.gf[i] = s[lpad(i)] * iq;
}

// Pre-op for pixel (row, col) of component c: returns the value fed to the x axis.
__device__ __forceinline__ float2 czt_pre_value(const CztArgs& a, int c, int row, int col, float2 fpix) {
  const int64_t pix = (int64_t)row * a.ax.lin + col;
  float2 e0 = a.in[pix];
  float2 e1 = a.nin > 1 ? a.in[a.in_cs + pix] : make_float2(0.f, 0.f);
  float2 e2 = a.nin > 2 ? a.in[2 * a.in_cs + pix] : make_float2(0.f, 0.f);
  apply_mask(a.mask, pix, e0, e1, e2);
  if (a.pre_mode == PRE_HNA) {
    // optical_elements.py:528-594: theta = rho/f, phi = atan2(Y, X), Richards-Wolf map
    float A[3][2];
    hna_map(a.px[col], a.py[row], a.hna_radius, a.hna_f, a.hna_stm2, A);
    return c_add(c_scale(e0, A[c][0]), c_scale(e1, A[c][1]));
  }
  float2 v = c == 0 ? e0 : (c == 1 ? e1 : e2);
  if (c == 2 && a.ez_mode != EZ_NONE) v = build_ez(a.ez_mode, e0, e1, a.px[col], a.py[row], a.z);
  if (a.pre_mode == PRE_RS || a.pre_mode == PRE_CONJ_RS) v = c_mul(v, fpix);
  else if (a.pre_mode == PRE_CONST) v = c_mul(v, a.cte);
  return v;
}

template <int Q>
__global__ __launch_bounds__(Q / 4) void czt_rows_kernel(CztArgs a) {
  constexpr int TPF = Q / 16, LS = LdsLen<Q>::value;
  constexpr int PPT = Q / TPF;
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int nrg = (a.rin + 3) >> 2;
  const int rg = blockIdx.x;
  const int row = rg * 4 + f;
  const bool rowok = row < a.rin;
  const int lin = a.ax.lin, lout = a.ax.lout, woff = a.ax.woff;
  const int nog = (lout + 3) >> 2;
  // transfer function of the input plane, once per pixel (float64 phase)
  float2 fp[PPT];
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int col = t + i * TPF;
    fp[i] = make_float2(1.f, 0.f);
    if (rowok && col < lin && (a.pre_mode == PRE_RS || a.pre_mode == PRE_CONJ_RS)) {
      float2 h = rs_kernel_value(a.px[col], a.py[row], a.z, a.k);
      if (a.pre_mode == PRE_CONJ_RS) h = c_mul(conjf2(h), a.cte);
      fp[i] = h;
    }
  }
  for (int c = 0; c < a.ncomp; ++c) {
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int col = t + i * TPF;
      float2 v = make_float2(0.f, 0.f);
      if (rowok && col < lin) v = c_mul(czt_pre_value(a, c, row, col, fp[i]), a.ax.pre[col]);
      s[lpad(col)] = v;
    }
    __syncthreads();
    fft_lds<Q, TPF, false>(s, t);
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int idx = t + i * TPF;
      s[lpad(idx)] = cmul(s[lpad(idx)], a.ax.gf[idx]);
    }
    __syncthreads();
    fft_lds<Q, TPF, true>(s, t);
    // window * post -> tiles [out group][row group]
    float2* dst = a.t + (int64_t)c * nrg * 4 * (nog * 4);
    for (int e = tid; e < nog * 16; e += Q / 4) {
      const int og = e >> 4, rr = (e >> 2) & 3, cc = e & 3;
      const int k = og * 4 + cc;
      float2 v = make_float2(0.f, 0.f);
      if (k < lout) v = cmul(lds[rr * LS + lpad(woff + k)], a.ax.post[k]);
      dst[((int64_t)og * nrg + rg) * 16 + rr * 4 + cc] = v;
    }
    __syncthreads();
  }
}

template <int Q>
__global__ __launch_bounds__(Q / 4) void czt_cols_kernel(CztArgs a) {
  constexpr int TPF = Q / 16, LS = LdsLen<Q>::value;
  constexpr int PPT = Q / TPF;
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int nrg = (a.rin + 3) >> 2;
  const int og = blockIdx.x;          // output columns 4og..4og+3 of the x axis
  const int lout_x = a.ax.lout;
  const int nog = (lout_x + 3) >> 2;
  const int lin = a.ay.lin, lout = a.ay.lout, woff = a.ay.woff;
  for (int c = 0; c < a.ncomp; ++c) {
    const float2* src = a.t + (int64_t)c * nrg * 4 * (nog * 4) + (int64_t)og * nrg * 16;
    for (int e = tid; e < nrg * 16; e += Q / 4) {
      const int r = (e >> 4) * 4 + ((e >> 2) & 3), cc = e & 3;
      if (r < lin) lds[cc * LS + lpad(r)] = cmul(src[e], a.ay.pre[r]);
    }
    for (int e = tid; e < 4 * (Q - lin); e += Q / 4) {
      const int cc = e & 3, r = lin + (e >> 2);
      lds[cc * LS + lpad(r)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<Q, TPF, false>(s, t);
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int idx = t + i * TPF;
      s[lpad(idx)] = cmul(s[lpad(idx)], a.ay.gf[idx]);
    }
    __syncthreads();
    fft_lds<Q, TPF, true>(s, t);
    float2* oc = a.out + (int64_t)c * a.out_cs;
    for (int e = tid; e < lout * 4; e += Q / 4) {
      const int ky = e >> 2, cc = e & 3;
      const int kx = og * 4 + cc;
      if (kx < lout_x) {
        float2 v = cmul(lds[cc * LS + lpad(woff + ky)], a.ay.post[ky]);
        if (a.post_mode == POST_RS) v = c_mul(c_mul(v, rs_kernel_value(a.qx[kx], a.qy[ky], a.z, a.k)), a.cte);
        else if (a.post_mode == POST_CONJ_RS) v = c_mul(v, conjf2(rs_kernel_value(a.qx[kx], a.qy[ky], a.z, a.k)));
        else if (a.post_mode == POST_CONST) v = c_mul(v, a.cte);
        oc[(int64_t)ky * lout_x + kx] = v;
      }
    }
    __syncthreads();
  }
}

template <int Q>
inline void czt_rows_launch(const CztArgs& a, hipStream_t st) {
  KTimer t(K_CZT_ROWS, st);
  hipLaunchKernelGGL(czt_rows_kernel<Q>, dim3((a.rin + 3) / 4), dim3(Q / 4), 0, st, a);
}
template <int Q>
inline void czt_cols_launch(const CztArgs& a, hipStream_t st) {
  KTimer t(K_CZT_COLS, st);
  hipLaunchKernelGGL(czt_cols_kernel<Q>, dim3((a.ax.lout + 3) / 4), dim3(Q / 4), 0, st, a);
}
template <int Q>
inline void czt_setup_launch(const CztSetupArgs& a, hipStream_t st) {
  KTimer t(K_CZT_SETUP, st);
  hipLaunchKernelGGL(czt_setup_kernel<Q>, dim3(1), dim3(Q / 16), 0, st, a);
}

}  // namespace xl

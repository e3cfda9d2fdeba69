// Chirp z-transform (Bluestein) propagation: CZT / VCZT / high-NA VCZT and
// their adjoints, plain FP32 form.
//
// Reference: Bluestein_method / compute_fft (xlumina/wave_optics.py:385-460),
// CZT_jit (:333-357), CZT_for_high_NA_jit (:359-371), VCZT_jit
// (vectorized_optics.py:375-384), VCZT_objective_lens
// (optical_elements.py:600-638).  Per axis the reference computes
//   out_k = post_k * sum_j g[m + k - j] * pre_j * x_j,   k < M, j < m,
// by a length-np2 circular convolution, with
//   pre_j  = A^{-j} W^{j^2/2},  g_s = W^{-(s-m+1)^2/2} (s < m+M; the reference's
//   table lacks s = m+M-1 when M >= m), post_k = W^{k^2/2} e^{-i2pi l_k(1/2-m/2)/Dm}.
// That is a Toeplitz product out_k = post_k sum_j gamma(k - j) pre_j x_j with the
// lag kernel gamma(d) = g[m + d]; the adjoint axis has pre' = conj(post),
// post' = conj(pre), gamma'(d) = conj(g[m - d]).  Here the kernel is laid out
// on the circle of length Q >= m + M - 1 by lag (lag d at index d mod Q), so
// the outputs are the first M samples of the inverse transform whatever the
// direction.  When the input and the output each fit in half the (power-of-
// two) length, the first forward pass skips the zero upper half of its
// inputs and the last inverse pass forms only the lower half of its outputs.
// The tables are generated on the device with float64 phases
// (czt_setup_kernel); the FFT of the kernel is stored pre-divided by Q.
//
// The 2-D transform is separable (K_y E K_x^T), so the row pass runs along x on
// the row-major input (contiguous loads) and writes a 4x4-tiled intermediate
// ordered [x-output group][row group]; the column pass reads it back column-
// contiguously, runs the y axis and applies the output-plane factor.
#pragma once
#include "fft.h"
#include "fftio.h"
#include "masks.h"
#include "rs.h"
#include "pre.h"

namespace xl {

enum { PRE_NONE = 0, PRE_RS = 1, PRE_HNA = 2, PRE_CONJ_RS = 3, PRE_CONST = 4 };
enum { POST_NONE = 0, POST_RS = 1, POST_CONST = 2, POST_CONJ_RS = 3 };

struct CztAxis {
  int lin, lout, zl, q;   // zl: zero-half first / last passes (lin, lout <= q/2, q a power of two)
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
  float2* t;           // intermediate, per component rin4 x lout4 in 4x4 tiles (czt_tidx)
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
  int64_t in_bs, out_bs, t_bs, mask_bs;   // batch strides (blockIdx.y / .z selects the member)
  int rin4;            // rin rounded up to a multiple of 4
  int64_t t_cs;        // component stride of the intermediate (rin4 * lout4)
  // Optional tables of the input-plane factor (rin x ax.lin) and the
  // output-plane factor (ay.lout x ax.lout), shared by a batch: built once by
  // czt_ftab_kernel instead of per member in float64.
  const float2* ftab;
  const float2* f0tab;
};

// Offset of (row, k) in the tiled intermediate: 4 rows x 4 outputs per 128-byte
// tile, tiles ordered [k / 4][row / 4], so the column pass reads whole lines.
__device__ __forceinline__ int64_t czt_tidx(int rin4, int row, int k) {
  return ((int64_t)(k >> 2) * (rin4 >> 2) + (row >> 2)) * 16 + (row & 3) * 4 + (k & 3);
}

// The arguments of batch member b: inputs, outputs, intermediate and
// per-pixel mask parameters offset by their batch strides.
__device__ __forceinline__ CztArgs czt_member(const CztArgs& a0, int b) {
  CztArgs a = a0;
  a.in += b * a0.in_bs;
  a.out += b * a0.out_bs;
  a.t += b * a0.t_bs;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (a.mask.st[i].m0) a.mask.st[i].m0 += b * a0.mask_bs;
    if (a.mask.st[i].m1) a.mask.st[i].m1 += b * a0.mask_bs;
  }
  return a;
}

struct CztSetupArgs {
  float2* pre;
  float2* gf;
  float2* post;
  int m, M, q, adjoint;
  double d1_dm;        // D1 / Dm
  double dl_dm;        // Delta / Dm,  Delta = (D2 - D1) / M
  double th_turns;     // angle(W) / (2 pi), principal branch
};

// One workgroup of q/16 threads builds one axis' tables (block b takes args b).
struct CztSetupPair {
  CztSetupArgs ax[2];
};

// Tables of one axis on a block of NT threads: the chirps and the kernel
// (float64 phases, the bulk of the work) on all NT threads, the kernel's FFT
// on the first Q/16 (the others only join its barriers).
template <int Q, int NT>
__device__ __forceinline__ void czt_setup_body(const CztSetupArgs& a, float2* s, int t) {
  constexpr int TPF = Q / 16;
  static_assert(NT >= TPF, "the FFT needs Q/16 threads");
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
    for (int j = t; j < a.m; j += NT) a.pre[j] = pre_f(j);
    for (int k = t; k < a.M; k += NT) a.post[k] = post_f(k);
  } else {
    for (int k = t; k < a.M; k += NT) a.pre[k] = conjf2(post_f(k));
    for (int j = t; j < a.m; j += NT) a.post[j] = conjf2(pre_f(j));
  }
  // lag d = k - j in [-(lin-1), lout-1] at index d mod Q (Q >= lin + lout - 1)
  const int lin = a.adjoint ? a.M : a.m, lout = a.adjoint ? a.m : a.M;
  for (int i = t; i < Q; i += NT) {
    const int d = i < lout ? i : (i > Q - lin ? i - Q : INT_MIN);
    float2 v = make_float2(0.f, 0.f);
    if (d != INT_MIN) v = a.adjoint ? conjf2(g_f(a.m - d)) : g_f(a.m + d);
    s[lpad(i)] = v;
  }
  __syncthreads();
  fft_lds<Q, TPF, false, 0, (NT > TPF)>(s, t);
  const float iq = 1.0f / (float)Q;
  for (int i = t; i < Q; i += NT) a.gf[i] = s[lpad(i)] * iq;
}

// Threads per set-up block: the float64 table work spreads over 1024 threads
// (the block runs alone on its CU; one or two such blocks make up the launch).
constexpr int CZT_SETUP_THREADS = 1024;

template <int Q>
__global__ __launch_bounds__(CZT_SETUP_THREADS) void czt_setup_kernel(CztSetupPair pair) {
  __shared__ float2 s[LdsLen<Q>::value];
  czt_setup_body<Q, CZT_SETUP_THREADS>(pair.ax[blockIdx.x], s, threadIdx.x);
}

// Both axes in one launch when their lengths differ (block 0: QA, block 1: QB).
template <int QA, int QB>
__global__ __launch_bounds__(CZT_SETUP_THREADS) void czt_setup2_kernel(CztSetupPair pair) {
  constexpr int QM = QA > QB ? QA : QB;
  __shared__ float2 s[LdsLen<QM>::value];
  if (blockIdx.x == 0) czt_setup_body<QA, CZT_SETUP_THREADS>(pair.ax[0], s, threadIdx.x);
  else czt_setup_body<QB, CZT_SETUP_THREADS>(pair.ax[1], s, threadIdx.x);
}

// Pre-op families the row kernel is specialised for.
enum { PREK_SIMPLE = 0, PREK_F = 1, PREK_HNA = 2 };

// All pre-op components of a pixel's masked field (e0, e1, e2) at input-plane
// coordinates (X, Y): Ez or the lens map, and the plane factor (fpix for
// PREK_F, cte for PRE_CONST).
template <int K>
__device__ __forceinline__ void czt_pre_compute(const CztArgs& a, float2 e0, float2 e1, float2 e2, double X, double Y,
                                                float2 fpix, float2 (&v)[3]) {
  if constexpr (K == PREK_HNA) {
    // optical_elements.py:528-594: theta = rho/f, phi = atan2(Y, X), Richards-Wolf map
    float A[3][2];
    hna_map(X, Y, a.hna_radius, a.hna_f, a.hna_stm2, A);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = c_add(c_scale(e0, A[c][0]), c_scale(e1, A[c][1]));
  } else {
    if (a.ez_mode != EZ_NONE) e2 = build_ez(a.ez_mode, e0, e1, X, Y, a.z);
    v[0] = e0;
    v[1] = e1;
    v[2] = e2;
    float2 m = make_float2(1.f, 0.f);
    if constexpr (K == PREK_F) m = fpix;
    else if (a.pre_mode == PRE_CONST) m = a.cte;
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = c_mul(v[c], m);
  }
}

// Transfer function of the input plane at (X, Y) (or its conjugate times cte
// for the adjoint).
__device__ __forceinline__ float2 czt_f_at(const CztArgs& a, double X, double Y) {
  float2 h = rs_kernel_value(X, Y, a.z, a.k);
  if (a.pre_mode == PRE_CONJ_RS) h = c_mul(conjf2(h), a.cte);
  return h;
}
__device__ __forceinline__ float2 czt_f_compute(const CztArgs& a, int row, int col) {
  return czt_f_at(a, a.px[col], a.py[row]);
}

// Output-plane factor of element (ky, kx): F0 cte (forward), conj(F0)
// (adjoint), cte, or 1.
__device__ __forceinline__ float2 czt_post_factor(const CztArgs& a, int ky, int kx) {
  if (a.post_mode == POST_RS) return c_mul(rs_kernel_value(a.qx[kx], a.qy[ky], a.z, a.k), a.cte);
  if (a.post_mode == POST_CONJ_RS) return conjf2(rs_kernel_value(a.qx[kx], a.qy[ky], a.z, a.k));
  if (a.post_mode == POST_CONST) return a.cte;
  return make_float2(1.f, 0.f);
}

// Both factor tables of a batched RS-mode CZT (grid-stride).
__global__ void czt_ftab_kernel(CztArgs a, float2* ftab, float2* f0tab) {
  const int64_t n1 = (int64_t)a.rin * a.ax.lin, n2 = (int64_t)a.ay.lout * a.ax.lout;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n1 + n2; e += (int64_t)gridDim.x * blockDim.x) {
    if (e < n1) ftab[e] = czt_f_compute(a, (int)(e / a.ax.lin), (int)(e % a.ax.lin));
    else {
      const int64_t f = e - n1;
      f0tab[f] = czt_post_factor(a, (int)(f / a.ax.lout), (int)(f % a.ax.lout));
    }
  }
}

// One block per input row: component c runs its FFT on threads
// [c*TPF, (c+1)*TPF) and LDS buffer c; the pre-op of each pixel (masks, Ez,
// transfer function in float64 phase, lens map) is computed once by one
// thread for all components.  The chirp spectrum is applied in registers
// between the forward and inverse chains and only the output window is
// stored, into the tiled intermediate (consecutive rows run on one XCD, so
// the four rows of a tile meet in one L2).  ZL: the input and the output fit
// in half the length (zero-half first pass, lower-half last pass).
template <int Q, bool ZL> struct CztRowsShape {
  // At 1024 and 2048 the three buffers (26 / 52 KB) leave room for 6 / 3
  // blocks of 3 / 6 waves per CU, which a SIMD holds at 5 waves (<= 96 VGPRs);
  // the zero-half forms fit that without spilling.
  static constexpr int WAVES = (ZL && (Q == 1024 || Q == 2048)) ? 5 : 1;
  // Pixels staged per thread at a time, their loads grouped ahead of the
  // pre-ops (measured: 1 -> 3 took the 4f rows from 3.04 to 2.44 ms per step
  // and the VCZT rows from 122 to 111 us; 2 is close behind).
  static constexpr int U = 3;
};

template <int Q, int K, bool ZL>
__global__ __launch_bounds__(3 * Q / 16)
__attribute__((amdgpu_waves_per_eu(CztRowsShape<Q, ZL>::WAVES))) void czt_rows_kernel(CztArgs a0) {
  const CztArgs a = czt_member(a0, blockIdx.y);
  constexpr int TPF = Q / 16, LS = LdsLen<Q>::value;
  using HO = HandOff<Q, TPF>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int f = tid / TPF, t = tid % TPF;
  const int row = xcd_remap(blockIdx.x, gridDim.x);
  const int lin = a.ax.lin, lout = a.ax.lout;
  const int nc = a.ncomp;
  // Staging: the pre-op of each pixel once for all components, U pixels per
  // thread at a time with their loads issued ahead of the pre-ops; columns
  // lin..LIM-1 are the zero padding.
  constexpr int LIM = ZL ? Q / 2 : Q, U = CztRowsShape<Q, ZL>::U;
  // the input-plane grid is read only by the forms that use it
  const bool grid = K == PREK_HNA || (K == PREK_F && !a.ftab) || a.ez_mode != EZ_NONE;
  const bool ftab = K == PREK_F && a.ftab;
  const double Y = grid ? a.py[row] : 0.0;
  for (int c0 = tid; c0 < LIM; c0 += U * nthr) {
    PixRaw raw[U];
    int64_t pix[U];
    double X[U];
    float2 fr[U], pr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int col = c0 + u * nthr;
      const int cc = col < lin ? col : 0;
      pix[u] = (int64_t)row * lin + cc;
      pr[u] = a.ax.pre[cc];
      X[u] = 0.0;
      fr[u] = make_float2(1.f, 0.f);
    }
    pix_load_n<U>(a.in, a.in_cs, a.nin, a.mask, pix, raw);
    if (grid) {
#pragma unroll
      for (int u = 0; u < U; ++u) X[u] = a.px[pix[u] - (int64_t)row * lin];
    }
    if (ftab) {
#pragma unroll
      for (int u = 0; u < U; ++u) fr[u] = a.ftab[pix[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int col = c0 + u * nthr;
      if (col < LIM) {
        float2 v[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        if (col < lin) {
          float2 e0, e1, e2;
          pix_masked(a.mask, raw[u], e0, e1, e2);
          float2 fv = make_float2(1.f, 0.f);
          if constexpr (K == PREK_F) fv = ftab ? fr[u] : czt_f_at(a, X[u], Y);
          czt_pre_compute<K>(a, e0, e1, e2, X[u], Y, fv, v);
#pragma unroll
          for (int c = 0; c < 3; ++c) v[c] = c_mul(v[c], pr[u]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
          if (c < nc) lds[c * LS + lpad(col)] = v[c];
      }
    }
  }
  __syncthreads();
  float2* s = lds + f * LS;
  typename HO::Type u;
  NoIO none;
  fft_forward_to_regs<Q, TPF, SRC_LDS, ZL>(s, t, none, u);
#pragma unroll
  for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
    for (int q = 0; q < HO::R; ++q) u.v[b][q] = cmul(u.v[b][q], a.ax.gf[HO::index(t, b, q)]);
  float2* dst = a.t + (int64_t)f * a.t_cs;
  auto store = [&](int k, float2 v) {
    if (k < lout) dst[czt_tidx(a.rin4, row, k)] = cmul(v, a.ax.post[k]);
  };
  fft_inverse_from_regs<Q, TPF, DST_FN, ZL>(s, t, u, store);
}

// Per block CPB output columns of the x axis (one per thread group) of one
// component: the columns of the tiled intermediate feed the first pass, the
// chirp spectrum is applied in registers, and only the output window is
// stored with the output-plane factor.  A 128 B tile line holds four columns:
// the CPB groups of a block (and at CPB = 2 the neighbouring block, kept on
// the same XCD) share it.  Column-wise, each lane's 8 B access is a line of
// its own (64 lines per wave for the output rows); staged through LDS, a wave
// stores CPB-column row segments and loads whole tile lines.  ZL as in the
// row kernel.
template <int Q> struct CztColsShape {
  static constexpr int CPB = Q >= 3072 ? 2 : 4;   // columns per block (more blocks at 3072 and 4096)
  // At 3072 three blocks (52 KB of LDS, 6 waves each) fit a CU when a SIMD
  // holds 5 waves, i.e. at most 96 VGPRs; unbounded, the compiler takes 112.
  static constexpr int WAVES = Q == 3072 ? 5 : 1;
  // Power-of-two lengths stage the tile loads and the output rows through
  // LDS; the 3 * 2^k lengths read and write columns directly (staged, 3072
  // no longer fits the 5-wave bound and measured slower).
  static constexpr bool STAGE = (Q & (Q - 1)) == 0;
};

template <int Q, bool ZL>
__global__ __launch_bounds__(CztColsShape<Q>::CPB * FftThreads<Q>::value)
__attribute__((amdgpu_waves_per_eu(CztColsShape<Q>::WAVES))) void czt_cols_kernel(CztArgs a0) {
  const CztArgs a = czt_member(a0, blockIdx.z);
  constexpr int TPF = FftThreads<Q>::value, LS = LdsLen<Q>::value;
  constexpr int CPB = CztColsShape<Q>::CPB;
  using HO = HandOff<Q, TPF>;
  __shared__ float2 lds[CPB * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int og = xcd_remap(blockIdx.x, gridDim.x);
  const int c = blockIdx.y;
  const int lout_x = a.ax.lout;
  const int kx = og * CPB + f;
  const bool colok = kx < lout_x;
  const int lin = a.ay.lin, lout = a.ay.lout;
  const float2* src = a.t + (int64_t)c * a.t_cs;
  typename HO::Type u;
  NoIO none;
  constexpr int NT = CPB * TPF;
  const int kx0 = og * CPB;
  if constexpr (CztColsShape<Q>::STAGE) {
    // The block's CPB columns lie in one 4-column tile group: they are staged
    // into LDS with lane-contiguous loads (column fastest) and the first pass
    // reads them from there.
    // IT elements per thread, loaded U at a time before their products
    constexpr int IT = (ZL ? Q / 2 : Q) / TPF, U = IT < 8 ? IT : 8;
#pragma unroll
    for (int i0 = 0; i0 < IT; i0 += U) {
      float2 tv[U], pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = threadIdx.x + (i0 + u) * NT, r = e / CPB, g = e % CPB;
        const bool ok = r < lin && kx0 + g < lout_x;
        tv[u] = src[czt_tidx(a.rin4, ok ? r : 0, ok ? kx0 + g : kx0)];
        pv[u] = a.ay.pre[ok ? r : 0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = threadIdx.x + (i0 + u) * NT, r = e / CPB, g = e % CPB;
        const bool ok = r < lin && kx0 + g < lout_x;
        lds[g * LS + lpad(r)] = ok ? cmul(tv[u], pv[u]) : make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    fft_forward_to_regs<Q, TPF, SRC_LDS, ZL>(s, t, none, u);
  } else {
    auto load = [&](int r) -> float2 {
      return (colok && r < lin) ? cmul(src[czt_tidx(a.rin4, r, kx)], a.ay.pre[r]) : make_float2(0.f, 0.f);
    };
    // (twiddle bases loaded at each pass: prefetched, they cost the 3 * 2^k
    // forms their 5-wave bound)
    fft_forward_to_regs<Q, TPF, SRC_FN, ZL, false>(s, t, load, u);
  }
#pragma unroll
  for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
    for (int q = 0; q < HO::R; ++q) u.v[b][q] = cmul(u.v[b][q], a.ay.gf[HO::index(t, b, q)]);
  float2* oc = a.out + (int64_t)c * a.out_cs;
  if constexpr (CztColsShape<Q>::STAGE) {
    // Last pass into registers, the outputs transposed through LDS to
    // [ky][column] and stored row by row (CPB consecutive columns per row).
    using LP = LastPass<Q, TPF, false>;
    typename LP::Type o;
    chain_io<Q, TPF, true, false, 0, SRC_REG, DST_REG, false, ZL>(s, t, none, none, u, o);
    __syncthreads();   // the last pass of other waves may still read LDS
#pragma unroll
    for (int b = 0; b < LP::BPT; ++b) {
      if (!LP::valid(t, b)) continue;
#pragma unroll
      for (int q = 0; q < (ZL ? LP::R / 2 : LP::R); ++q) {
        const int ky = LP::index(t, b, q);
        if (ky < lout) lds[ky * CPB + f] = cmul(o.v[b][q], a.ay.post[ky]);
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < lout * CPB; e += NT) {
      const int ky = e / CPB, kx2 = kx0 + e % CPB;
      if (kx2 < lout_x) {
        float2 v = lds[e];
        if (a.f0tab) v = c_mul(v, a.f0tab[(int64_t)ky * lout_x + kx2]);
        else if (a.post_mode != POST_NONE) v = c_mul(v, czt_post_factor(a, ky, kx2));
        oc[(int64_t)ky * lout_x + kx2] = v;
      }
    }
  } else {
    auto store = [&](int ky, float2 v) {
      if (colok && ky < lout) {
        v = cmul(v, a.ay.post[ky]);
        if (a.f0tab) v = c_mul(v, a.f0tab[(int64_t)ky * lout_x + kx]);
        else if (a.post_mode != POST_NONE) v = c_mul(v, czt_post_factor(a, ky, kx));
        oc[(int64_t)ky * lout_x + kx] = v;
      }
    };
    fft_inverse_from_regs<Q, TPF, DST_FN, ZL, false>(s, t, u, store);
  }
}

// Zero-half forms exist for the power-of-two lengths.
template <int Q> struct CztZl { static constexpr bool value = (Q & (Q - 1)) == 0; };

template <int Q, bool ZL>
inline void czt_rows_launch_zl(const CztArgs& a, int batch, hipStream_t st) {
  const dim3 g(a.rin, batch), b(a.ncomp * Q / 16);
  const size_t lds = (size_t)a.ncomp * LdsLen<Q>::value * sizeof(float2);
  if (a.pre_mode == PRE_RS || a.pre_mode == PRE_CONJ_RS)
    hipLaunchKernelGGL((czt_rows_kernel<Q, PREK_F, ZL>), g, b, lds, st, a);
  else if (a.pre_mode == PRE_HNA)
    hipLaunchKernelGGL((czt_rows_kernel<Q, PREK_HNA, ZL>), g, b, lds, st, a);
  else
    hipLaunchKernelGGL((czt_rows_kernel<Q, PREK_SIMPLE, ZL>), g, b, lds, st, a);
}
template <int Q>
inline void czt_rows_launch(const CztArgs& a, int batch, hipStream_t st) {
  KTimer t(K_CZT_ROWS, st);
  if constexpr (CztZl<Q>::value)
    if (a.ax.zl) return czt_rows_launch_zl<Q, true>(a, batch, st);
  czt_rows_launch_zl<Q, false>(a, batch, st);
}
// Dynamic-LDS limit of the row kernel (three Q-point buffers), set once per
// device by the library's initialisation for every Bluestein length.
template <int Q, bool ZL> inline hipError_t czt_set_attribute_zl() {
  const int mx = 3 * LdsLen<Q>::value * (int)sizeof(float2);
  hipError_t e = hipSuccess;
  for (const void* f : {(const void*)czt_rows_kernel<Q, PREK_F, ZL>, (const void*)czt_rows_kernel<Q, PREK_HNA, ZL>,
                        (const void*)czt_rows_kernel<Q, PREK_SIMPLE, ZL>}) {
    const hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, mx);
    if (r != hipSuccess) e = r;
  }
  return e;
}
template <int Q> inline hipError_t czt_set_attribute() {
  hipError_t e = czt_set_attribute_zl<Q, false>();
  if constexpr (CztZl<Q>::value) {
    const hipError_t r = czt_set_attribute_zl<Q, true>();
    if (r != hipSuccess) e = r;
  }
  return e;
}
inline hipError_t czt_set_attributes() {
  hipError_t e = hipSuccess;
  for (hipError_t r : {czt_set_attribute<64>(), czt_set_attribute<128>(), czt_set_attribute<256>(),
                       czt_set_attribute<512>(), czt_set_attribute<1024>(), czt_set_attribute<2048>(),
                       czt_set_attribute<4096>(), czt_set_attribute<768>(), czt_set_attribute<1536>(),
                       czt_set_attribute<3072>()})
    if (r != hipSuccess) e = r;
  return e;
}

template <int Q>
inline void czt_cols_launch(const CztArgs& a, int batch, hipStream_t st) {
  KTimer t(K_CZT_COLS, st);
  constexpr int CPB = CztColsShape<Q>::CPB;
  const dim3 g((a.ax.lout + CPB - 1) / CPB, a.ncomp, batch), b(CPB * FftThreads<Q>::value);
  if constexpr (CztZl<Q>::value)
    if (a.ay.zl) {
      hipLaunchKernelGGL((czt_cols_kernel<Q, true>), g, b, 0, st, a);
      return;
    }
  hipLaunchKernelGGL((czt_cols_kernel<Q, false>), g, b, 0, st, a);
}
template <int Q>
inline void czt_setup_launch(const CztSetupPair& p, int nblocks, hipStream_t st) {
  KTimer t(K_CZT_SETUP, st);
  hipLaunchKernelGGL(czt_setup_kernel<Q>, dim3(nblocks), dim3(CZT_SETUP_THREADS), 0, st, p);
}
template <int QA, int QB>
inline void czt_setup2_launch(const CztSetupPair& p, hipStream_t st) {
  KTimer t(K_CZT_SETUP, st);
  hipLaunchKernelGGL((czt_setup2_kernel<QA, QB>), dim3(2), dim3(CZT_SETUP_THREADS), 0, st, p);
}

}  // namespace xl

// Chirp z-transform as a dense zoom-DFT on MFMA (split-BF16, FP32 accumulate).
//
// The 2-D CZT of the reference (CZT_jit, wave_optics.py:333-357; high-NA form
// optical_elements.py:600-638) is  out = post2d ( K_y G K_x^T )  with G the
// pre-op'd field and, per axis, K[k][j] = post_k g[m+k-j] pre_j -- the same
// float64 phases as the Bluestein tables (czt.h), so both forms compute the
// same operator.  Here it runs as two complex GEMMs of the form C = A B^T
// (A: M x K, B: N x K, both K-contiguous):
//   stage 1: C1 = K_x [G_0; G_1; G_2]^T   (Mx x 3*Nyp)   components batched in N
//   stage 2: C2 = K_y  B2^T, B2 row (c, kx) = C1[kx][c*Nyp : (c+1)*Nyp]
// Operands are stored as four bf16 planes (re_hi, re_lo, im_hi, im_lo) with
// x = hi + lo (hi = bf16(x), lo = bf16(x - hi)); each complex MAC is
// 4 real products x 3 split terms (hi*hi, hi*lo, lo*hi) = 12 MFMAs
// (v_mfma_f32_32x32x16_bf16), accumulated in FP32 as C_rr, C_ii, C_ri.
// The adjoint uses the conjugate-transposed kernels with the same GEMMs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft.h"
#include "masks.h"
#include "pre.h"
#include "rs.h"

namespace xl {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int GBM = 128, GBK = 32;   // block tile rows (complex), K step; the tile width is 32 * WN
constexpr int GLDK = GBK + 8;                   // LDS row stride in bf16 (80 B: 16 B aligned, 2-way banks)

__device__ __forceinline__ uint16_t f2bf(float x) {   // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffffu) ? 0x40 : 0);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// x -> (hi, lo) bf16 pair
__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));
}

// Four-plane split storage of a complex matrix.
struct Planes {
  uint16_t* p;
  int64_t ps;   // plane stride (elements)
  int64_t ld;   // row stride (elements)
};

__device__ __forceinline__ void store_split(const Planes& d, int64_t off, float2 v) {
  uint16_t h, l;
  split_bf16(v.x, h, l);
  d.p[off] = h;
  d.p[d.ps + off] = l;
  split_bf16(v.y, h, l);
  d.p[2 * d.ps + off] = h;
  d.p[3 * d.ps + off] = l;
}

// Two consecutive elements of a split matrix, packed for 4-byte stores (a
// wave then writes 256 contiguous bytes per plane while its loads of the
// source stay lane-contiguous).
struct Split2 {
  uint32_t w[4];   // [plane]: element 0 in the low half, element 1 in the high half
};
__device__ __forceinline__ void split2_put(Split2& s, int i, float2 v) {
  uint16_t h[4];
  split_bf16(v.x, h[0], h[1]);
  split_bf16(v.y, h[2], h[3]);
#pragma unroll
  for (int pl = 0; pl < 4; ++pl) {
    if (i) s.w[pl] |= (uint32_t)h[pl] << 16;
    else s.w[pl] = h[pl];
  }
}
// off: element offset of the first of the two (even).
__device__ __forceinline__ void store_split2(const Planes& d, int64_t off, const Split2& s) {
#pragma unroll
  for (int pl = 0; pl < 4; ++pl) *reinterpret_cast<uint32_t*>(d.p + pl * d.ps + off) = s.w[pl];
}

// ---------------------------------------------------------------------------
// kernel matrices: K[k][j] = post_k g[m+k-j] pre_j as float64 turns
// ---------------------------------------------------------------------------
struct KgenArgs {
  Planes dst;
  int m, M;            // forward axis lengths (input m, output M)
  int rows, cols;      // stored matrix (forward: M x colsp, adjoint: m x colsp)
  int valid_cols;      // forward: m, adjoint: M
  int adjoint;         // store conj(K)^T
  double d1_dm, dl_dm, th_turns;
};

__device__ __forceinline__ double czt_pre_turns(const KgenArgs& a, int j) {
  const double jj = (double)j;
  return -jj * a.d1_dm + a.th_turns * 0.5 * jj * jj;
}
__device__ __forceinline__ double czt_post_turns(const KgenArgs& a, int k) {
  const double kk = (double)k;
  return a.th_turns * 0.5 * kk * kk - (a.d1_dm + kk * a.dl_dm) * (0.5 - 0.5 * a.m);
}

__device__ __forceinline__ float2 kgen_value(const KgenArgs& a, int r, int c) {
  float2 v = make_float2(0.f, 0.f);
  if (c < a.valid_cols) {
    const int k = a.adjoint ? c : r;   // output index of the forward axis
    const int j = a.adjoint ? r : c;   // input index
    const int s = a.m + k - j;          // g index, in [1, m+M-1]
    if (!(a.M >= a.m && s == a.m + a.M - 1)) {
      const double u = (double)(s - a.m + 1);
      const double turns = czt_post_turns(a, k) - a.th_turns * 0.5 * u * u + czt_pre_turns(a, j);
      v = expi_turns(a.adjoint ? -turns : turns);
    }
  }
  return v;
}

// Two consecutive columns per thread (ld is a multiple of 32); blocks
// [0, nblocks) of the caller's grid.
__device__ __forceinline__ void kgen_body(const KgenArgs& a, int block, int nblocks) {
  const int64_t total2 = (int64_t)a.rows * a.dst.ld / 2;
  for (int64_t e2 = block * (int64_t)blockDim.x + threadIdx.x; e2 < total2; e2 += (int64_t)nblocks * blockDim.x) {
    const int64_t e = e2 * 2;
    const int r = (int)(e / a.dst.ld), c0 = (int)(e % a.dst.ld);
    Split2 sp;
    split2_put(sp, 0, kgen_value(a, r, c0));
    split2_put(sp, 1, kgen_value(a, r, c0 + 1));
    store_split2(a.dst, e, sp);
  }
}

// ---------------------------------------------------------------------------
// pre-op: field -> split planes of G, rows (c, y) with per-component row stride
// ---------------------------------------------------------------------------
struct GemmPreArgs {
  const float2* in;
  int64_t in_cs;
  int nin, ncomp;
  int rows, cols, rowsp;    // input plane and padded rows per component
  Planes dst;               // ld = padded cols
  int pre_mode;             // PRE_* of czt.h
  int ez_mode;
  const double* px;
  const double* py;
  double z, k;
  float2 cte;
  double hna_radius, hna_f, hna_stm2;
  MaskArgs mask;
};

// Pre-op values of a pixel's masked field at input-plane coordinates (X, Y)
// for every component.
__device__ __forceinline__ void gemm_pre_compute(const GemmPreArgs& a, float2 e0, float2 e1, float2 e2, double X,
                                                 double Y, float2 (&v)[3]) {
  if (a.pre_mode == 2 /* PRE_HNA */) {
    float A[3][2];
    hna_map(X, Y, a.hna_radius, a.hna_f, a.hna_stm2, A);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = c_add(c_scale(e0, A[c][0]), c_scale(e1, A[c][1]));
  } else {
    v[0] = e0;
    v[1] = e1;
    v[2] = (a.ez_mode != EZ_NONE) ? build_ez(a.ez_mode, e0, e1, X, Y, a.z) : e2;
    if (a.pre_mode == 1 /* PRE_RS */ || a.pre_mode == 3 /* PRE_CONJ_RS */) {
      float2 h = rs_kernel_value(X, Y, a.z, a.k);
      if (a.pre_mode == 3) h = c_mul(conjf2(h), a.cte);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = c_mul(v[c], h);
    } else if (a.pre_mode == 4 /* PRE_CONST */) {
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = c_mul(v[c], a.cte);
    }
  }
}

// Two consecutive columns per thread (ld is a multiple of 32), 4-byte plane
// stores; blocks [0, nblocks) of the caller's grid.  Both pixels' loads are
// grouped ahead of their pre-ops.
__device__ __forceinline__ void gemm_pre_body(const GemmPreArgs& a, int block, int nblocks) {
  const int64_t total2 = (int64_t)a.rowsp * a.dst.ld / 2;
  const bool grid = (a.pre_mode >= 1 && a.pre_mode <= 3) || a.ez_mode != EZ_NONE;   // RS, HNA, conj RS, or Ez
  for (int64_t e2 = block * (int64_t)blockDim.x + threadIdx.x; e2 < total2; e2 += (int64_t)nblocks * blockDim.x) {
    const int64_t e = e2 * 2;
    const int row = (int)(e / a.dst.ld), col0 = (int)(e % a.dst.ld);
    const bool rowok = row < a.rows;
    bool ok[2];
    int64_t pix[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ok[i] = rowok && col0 + i < a.cols;
      pix[i] = ok[i] ? (int64_t)row * a.cols + col0 + i : 0;
    }
    PixRaw raw[2];
    pix_load_n<2>(a.in, a.in_cs, a.nin, a.mask, pix, raw);
    double X[2] = {0.0, 0.0}, Y = 0.0;
    if (grid) {
      X[0] = a.px[ok[0] ? col0 : 0];
      X[1] = a.px[ok[1] ? col0 + 1 : 0];
      Y = a.py[rowok ? row : 0];
    }
    Split2 sp[3];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      float2 v[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      if (ok[i]) {
        float2 e0, e1, e2;
        pix_masked(a.mask, raw[i], e0, e1, e2);
        gemm_pre_compute(a, e0, e1, e2, X[i], Y, v);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) split2_put(sp[c], i, v[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (c < a.ncomp) store_split2(a.dst, (int64_t)c * a.rowsp * a.dst.ld + e, sp[c]);
  }
}

// Operand preparation in one launch: the two kernel matrices (float64 phase,
// compute-bound) and the pre-op'd field planes (memory-bound) on disjoint
// block ranges, so they overlap instead of running back to back.
__global__ __launch_bounds__(256) void gemm_prep_kernel(KgenArgs k0, KgenArgs k1, GemmPreArgs pre, int nb0, int nb1) {
  const int b = blockIdx.x;
  if (b < nb0) kgen_body(k0, b, nb0);
  else if (b < nb0 + nb1) kgen_body(k1, b - nb0, nb1);
  else gemm_pre_body(pre, b - nb0 - nb1, gridDim.x - nb0 - nb1);
}

// ---------------------------------------------------------------------------
// C = A B^T, complex, split-BF16 MFMA
// ---------------------------------------------------------------------------
struct GemmOperand {
  const uint16_t* p;
  int64_t ps;       // plane stride
  int64_t ld;       // row stride within a batch
  int rows;         // valid rows (others read as zero)
  int rb;           // rows per batch: row r lives at (r % rb) * ld + (r / rb) * boff
  int64_t boff;
};

struct GemmArgs {
  GemmOperand A, B;
  int M, N, K;       // K multiple of GBK (zero padded)
  int mode;          // 0: split planes C, 1: complex output with post-op
  Planes c;          // mode 0
  // mode 1: output element (m, n) -> out[n / nb][m][n % nb] * post
  float2* out;
  int64_t out_cs;
  int nb;            // output columns per component
  int post_mode;     // POST_* of czt.h
  const double* qx;
  const double* qy;
  double z, k;
  float2 cte;
  float2* partial;   // split-K (gridDim.z > 1): fp32 partial sums [split][m][n]
};

__device__ __forceinline__ int64_t gop_off(const GemmOperand& o, int r) {
  return (int64_t)(r % o.rb) * o.ld + (int64_t)(r / o.rb) * o.boff;
}

// Writes output element (m, n) of the product (mode 0 or 1).
__device__ __forceinline__ void gemm_store(const GemmArgs& a, int m, int n, float2 v) {
  if (a.mode == 0) {
    store_split(a.c, (int64_t)m * a.c.ld + n, v);
  } else {
    const int comp = n / a.nb, kx = n - comp * a.nb;
    float2 o = v;
    if (a.post_mode == 1 /* POST_RS */) o = c_mul(c_mul(o, rs_kernel_value(a.qx[kx], a.qy[m], a.z, a.k)), a.cte);
    else if (a.post_mode == 3 /* POST_CONJ_RS */) o = c_mul(o, conjf2(rs_kernel_value(a.qx[kx], a.qy[m], a.z, a.k)));
    else if (a.post_mode == 2 /* POST_CONST */) o = c_mul(o, a.cte);
    a.out[(int64_t)comp * a.out_cs + (int64_t)m * a.nb + kx] = o;
  }
}

// Block tile GBM x (32 WN) complex on 4 x WN waves of 32 x 32; blockIdx.z
// takes an even share of the K steps (split-K) and then writes fp32 partial
// sums for cgemm_reduce_kernel instead of the output.
template <int WN>
__global__ __launch_bounds__(256 * WN) void cgemm_kernel(GemmArgs a) {
  constexpr int BN = 32 * WN, NT = 256 * WN;
  constexpr int CH_A = 4 * GBM * (GBK / 8);   // chunks of 8 bf16 per K step: A planes
  constexpr int CH = CH_A + 4 * BN * (GBK / 8);
  constexpr int PER = (CH + NT - 1) / NT;
  // [buffer][plane][row][k] bf16, two buffers: step ks computes on one while
  // the registers prefetched for step ks + 1 are stored into the other, so one
  // barrier per step suffices.  The block runs alone on its CU either way (its
  // register use allows at most four waves per SIMD, i.e. one block of 4 x WN
  // waves), so the second buffer costs no occupancy.
  __shared__ __attribute__((aligned(16))) uint16_t smA[2][4][GBM][GLDK];
  __shared__ __attribute__((aligned(16))) uint16_t smB[2][4][BN][GLDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 3, wn = w >> 2;
  const int bm = blockIdx.y * GBM, bn = blockIdx.x * BN;
  const int ksteps = a.K / GBK;
  const int kbeg = (int)((int64_t)ksteps * blockIdx.z / gridDim.z);
  const int kend = (int)((int64_t)ksteps * (blockIdx.z + 1) / gridDim.z);
  uint4 pre[PER];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int ch = tid + i * NT;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (CH % NT == 0 || ch < CH) {
        const bool isA = ch < CH_A;
        const int c2 = isA ? ch : ch - CH_A;
        const int rows = isA ? GBM : BN;
        const int pl = c2 / (4 * rows), r = (c2 >> 2) % rows, kc = c2 & 3;
        const GemmOperand& o = isA ? a.A : a.B;
        const int gr = (isA ? bm : bn) + r;
        if (gr < o.rows) v = *reinterpret_cast<const uint4*>(o.p + pl * o.ps + gop_off(o, gr) + k0 + kc * 8);
      }
      pre[i] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int ch = tid + i * NT;
      if (CH % NT == 0 || ch < CH) {
        if (ch < CH_A) {
          const int pl = ch / (4 * GBM), r = (ch >> 2) % GBM, kc = ch & 3;
          *reinterpret_cast<uint4*>(&smA[buf][pl][r][kc * 8]) = pre[i];
        } else {
          const int c2 = ch - CH_A;
          const int pl = c2 / (4 * BN), r = (c2 >> 2) % BN, kc = c2 & 3;
          *reinterpret_cast<uint4*>(&smB[buf][pl][r][kc * 8]) = pre[i];
        }
      }
    }
  };
  f32x16 crr = {}, cii = {}, cri = {};
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 16) {
      const int ra = wm * 32 + (lane & 31), rb = wn * 32 + (lane & 31), kq = kk + 8 * (lane >> 5);
      const bf16x8 arh = *reinterpret_cast<const bf16x8*>(&smA[buf][0][ra][kq]);
      const bf16x8 arl = *reinterpret_cast<const bf16x8*>(&smA[buf][1][ra][kq]);
      const bf16x8 aih = *reinterpret_cast<const bf16x8*>(&smA[buf][2][ra][kq]);
      const bf16x8 ail = *reinterpret_cast<const bf16x8*>(&smA[buf][3][ra][kq]);
      const bf16x8 brh = *reinterpret_cast<const bf16x8*>(&smB[buf][0][rb][kq]);
      const bf16x8 brl = *reinterpret_cast<const bf16x8*>(&smB[buf][1][rb][kq]);
      const bf16x8 bih = *reinterpret_cast<const bf16x8*>(&smB[buf][2][rb][kq]);
      const bf16x8 bil = *reinterpret_cast<const bf16x8*>(&smB[buf][3][rb][kq]);
      crr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, brh, crr, 0, 0, 0);
      crr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brl, crr, 0, 0, 0);
      crr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, brh, crr, 0, 0, 0);
      cii = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, bih, cii, 0, 0, 0);
      cii = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, bil, cii, 0, 0, 0);
      cii = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, bih, cii, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arl, bih, cri, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bil, cri, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, brh, cri, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brl, cri, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arh, bih, cri, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, brh, cri, 0, 0, 0);
    }
  };
  if (kbeg < kend) {
    gload(kbeg * GBK);
    sstore(0);
    __syncthreads();
    if (kbeg + 1 < kend) gload((kbeg + 1) * GBK);
  }
  for (int ks = kbeg; ks < kend; ++ks) {
    const int cur = (ks - kbeg) & 1;
    compute(cur);
    // the other buffer was last read in step ks - 1, before the barrier that ended it
    if (ks + 1 < kend) {
      sstore(cur ^ 1);
      if (ks + 2 < kend) gload((ks + 2) * GBK);
    }
    __syncthreads();
  }
  // epilogue: lane owns column n = lane % 32 and rows 8*(r/4) + 4*(lane/32) + r%4
  const int n = bn + wn * 32 + (lane & 31);
  if (n >= a.N) return;
  const bool split = gridDim.z > 1;
  float2* part = a.partial + (int64_t)blockIdx.z * a.M * a.N;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = bm + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= a.M) continue;
    const float2 v = make_float2(crr[r] - cii[r], cri[r]);
    if (split) part[(int64_t)m * a.N + n] = v;
    else gemm_store(a, m, n, v);
  }
}

// Sums the split-K partials in a fixed order and writes the output.
__global__ void cgemm_reduce_kernel(GemmArgs a, int splits) {
  const int64_t total = (int64_t)a.M * a.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    float2 v = a.partial[e];
    for (int s = 1; s < splits; ++s) v = v + a.partial[s * total + e];
    gemm_store(a, (int)(e / a.N), (int)(e % a.N), v);
  }
}

}  // namespace xl

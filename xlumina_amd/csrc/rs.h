// Rayleigh-Sommerfeld convolution (scalar RS / vectorial VRS) on a P x P grid,
// P = next power of two >= 2N-1, with the pre-propagation Jones masks and the
// Ez construction fused into the first row pass.
//
// Reference operator (xlumina/wave_optics.py:281-289): zero-pad the N x N field
// into a (2N-1)^2 grid, multiply fft2(U) by fft2(H) of the sampled RS impulse
// response on the lag grid (wave_optics.py:265-279, :291-297), ifft2, scale by
// dx*dy, keep [N-1:, N-1:].  That is the linear convolution
//     out[q,p] = sum_ij U[j,i] h(x_p - x_i, y_q - y_j) dx dy.
// Here lag +t sits at index t and lag -t at index P-t (slots N..P-N are zero),
// which gives the same linear convolution on any P >= 2N-1.  The kernel array
// is even in both axes, so only rows 0..P/2-1 of its row transform are built
// and the column pass mirrors them.
//
// Passes (all intermediates in 4x4-tile layout, 128 B per tile, ordered
// [column group][row group], so row passes write and column passes read whole
// 128 B lines):
//   rs_hrow   : H rows 0..P/2-1 generated in float64 phase -> row FFT -> Hr
//   rs_urow   : masks + Ez on the input rows -> zero-padded row FFT -> T
//   rs_col    : per 4 columns: FFT(Hr col) -> Hhat (registers), then for each
//               component FFT(T col) * Hhat -> IFFT, keep rows 0..N-1 -> T
//   rs_final  : row IFFT of T, keep columns 0..N-1 -> output (+ |E|^2)
#pragma once
#include "fft.h"
#include "masks.h"
#include "timing.h"

namespace xl {

struct RsArgs {
  const float2* in;       // nin components of N x N, row-major [y][x]
  int64_t in_cs;          // component stride (elements)
  float2* out;            // ncomp components of N x N
  int64_t out_cs;
  float* intensity;       // optional N x N, sum_c |out_c|^2
  float2* t;              // tiles, ncomp x NR4*4 x P
  float2* hr;             // tiles, (P/2) x P
  const double* x;
  const double* y;
  double z, k;
  float scale;            // dx*dy/P^2
  int n, nin, ncomp, ez_mode, adjoint;
  MaskArgs mask;
};

// lag value of circular slot s on a P-point axis with N samples (0 => unused)
__device__ __forceinline__ bool lag_of(const double* g, int s, int n, int P, double* d) {
  if (s < n) { *d = g[s] - g[0]; return true; }
  if (s > P - n) { *d = g[P - s] - g[0]; return true; }
  return false;
}

// h(x, y; z) of wave_optics.py:291-297 with the phase k r reduced in float64.
__device__ __forceinline__ float2 rs_kernel_value(double X, double Y, double z, double k) {
  const double r2 = X * X + Y * Y + z * z;
  const double r = sqrt(r2);
  const double ir = 1.0 / r;
  const double turns = k * r * 0.15915494309189535;   // k r / (2 pi)
  const float2 e = expi_turns(z > 0 ? turns : -turns);
  const double f = 0.15915494309189535 * z * ir * ir;   // (1/2pi) z / r^2
  const float fr = (float)(f * ir), fi = (float)(-f * k);   // f * (1/r - i k)
  return make_float2(fr * e.x - fi * e.y, fr * e.y + fi * e.x);
}

// tile index of element (r, c) for an array with nrg row groups
__device__ __forceinline__ int64_t tile_idx(int r, int c, int nrg) {
  return ((int64_t)(c >> 2) * nrg + (r >> 2)) * 16 + (r & 3) * 4 + (c & 3);
}

// Write 4 LDS rows (length >= ncols) to tiles at row group rg; thread-linear so
// each group of 16 lanes writes one 128 B tile.
template <int P>
__device__ __forceinline__ void store_rows_to_tiles(float2* dst, const float2* lds, int rg, int nrg, int ncols,
                                                    int tid, int nthr) {
  constexpr int LS = LdsLen<P>::value;
  const int total = ncols * 4;
  for (int e = tid; e < total; e += nthr) {
    const int cgp = e >> 4, rr = (e >> 2) & 3, cc = e & 3;
    const int c = cgp * 4 + cc;
    dst[((int64_t)cgp * nrg + rg) * 16 + rr * 4 + cc] = lds[rr * LS + lpad(c)];
  }
}

template <int P>
__global__ __launch_bounds__(P / 4) void rs_hrow_kernel(RsArgs a) {
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int rg = blockIdx.x;             // rows 4rg..4rg+3 of 0..P/2-1
  const int r = rg * 4 + f;
  double dy;
  const bool rowok = lag_of(a.y, r, a.n, P, &dy);
  for (int c = t; c <= P / 2; c += TPF) {
    double dx;
    float2 v = make_float2(0.f, 0.f);
    if (rowok && lag_of(a.x, c, a.n, P, &dx)) v = rs_kernel_value(dx, dy, a.z, a.k);
    s[lpad(c)] = v;
    if (c > 0 && c < P / 2) s[lpad(P - c)] = v;
  }
  __syncthreads();
  fft_lds<P, TPF, false>(s, t);
  store_rows_to_tiles<P>(a.hr, lds, rg, P / 8, P, tid, P / 4);
}

template <int P>
__global__ __launch_bounds__(P / 4) void rs_urow_kernel(RsArgs a) {
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  constexpr int PPT = (P / 2 + TPF - 1) / TPF;   // pixels per thread (row length <= P/2)
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int nrg = (a.n + 3) >> 2;
  const int rg = blockIdx.x;
  const int row = rg * 4 + f;
  const bool rowok = row < a.n;
  float2 v[3][PPT];
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    const int col = t + i * TPF;
    float2 e0 = make_float2(0.f, 0.f), e1 = e0, e2 = e0;
    if (rowok && col < a.n) {
      const int64_t pix = (int64_t)row * a.n + col;
      e0 = a.in[pix];
      if (a.nin > 1) e1 = a.in[a.in_cs + pix];
      if (a.nin > 2) e2 = a.in[2 * a.in_cs + pix];
      apply_mask(a.mask, pix, e0, e1, e2);
      if (a.ez_mode != EZ_NONE) e2 = build_ez(a.ez_mode, e0, e1, a.x[col], a.y[row], a.z);
    }
    v[0][i] = e0; v[1][i] = e1; v[2][i] = e2;
  }
  for (int c = 0; c < a.ncomp; ++c) {
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int col = t + i * TPF;
      if (col < P / 2) {
        s[lpad(col)] = c == 0 ? v[0][i] : (c == 1 ? v[1][i] : v[2][i]);
        s[lpad(col + P / 2)] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    fft_lds<P, TPF, false>(s, t);
    store_rows_to_tiles<P>(a.t + (int64_t)c * nrg * 4 * P, lds, rg, nrg, P, tid, P / 4);
    __syncthreads();
  }
}

template <int P>
__global__ __launch_bounds__(P / 4) void rs_col_kernel(RsArgs a) {
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  constexpr int EPT = P / TPF;   // 16 elements per thread per column
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int cg = xcd_remap(blockIdx.x, gridDim.x);
  const int nrg = (a.n + 3) >> 2;
  // 1) Hr columns (rows 0..P/2-1 stored, mirrored to P/2+1..P-1, row P/2 zero)
  {
    const float2* src = a.hr + (int64_t)cg * (P / 8) * 16;
    for (int e = tid; e < (P / 2) * 4; e += P / 4) {
      const int r = (e >> 4) * 4 + ((e >> 2) & 3), cc = e & 3;
      const float2 v = src[e];
      lds[cc * LS + lpad(r)] = v;
      if (r > 0) lds[cc * LS + lpad(P - r)] = v;
    }
    if (tid < 4) lds[tid * LS + lpad(P / 2)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<P, TPF, false>(s, t);
  float2 hh[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    float2 h = s[lpad(t + i * TPF)] * a.scale;
    hh[i] = a.adjoint ? conjf2(h) : h;
  }
  __syncthreads();
  for (int c = 0; c < a.ncomp; ++c) {
    float2* tc = a.t + (int64_t)c * nrg * 4 * P + (int64_t)cg * nrg * 16;
    for (int e = tid; e < nrg * 16; e += P / 4) {
      const int r = (e >> 4) * 4 + ((e >> 2) & 3), cc = e & 3;
      lds[cc * LS + lpad(r)] = tc[e];
    }
    for (int e = tid; e < 4 * (P - nrg * 4); e += P / 4) {
      const int cc = e & 3, r = nrg * 4 + (e >> 2);
      lds[cc * LS + lpad(r)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<P, TPF, false>(s, t);
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int idx = lpad(t + i * TPF);
      s[idx] = cmul(s[idx], hh[i]);
    }
    __syncthreads();
    fft_lds<P, TPF, true>(s, t);
    for (int e = tid; e < nrg * 16; e += P / 4) {
      const int r = (e >> 4) * 4 + ((e >> 2) & 3), cc = e & 3;
      tc[e] = lds[cc * LS + lpad(r)];
    }
    __syncthreads();
  }
}

template <int P>
__global__ __launch_bounds__(P / 4) void rs_final_kernel(RsArgs a) {
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  constexpr int PPT = (P / 2 + TPF - 1) / TPF;
  __shared__ float2 lds[4 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int nrg = (a.n + 3) >> 2;
  const int rg = blockIdx.x;
  const int row = rg * 4 + f;
  float inten[PPT];
#pragma unroll
  for (int i = 0; i < PPT; ++i) inten[i] = 0.f;
  for (int c = 0; c < a.ncomp; ++c) {
    const float2* tc = a.t + (int64_t)c * nrg * 4 * P;
    for (int e = tid; e < 4 * P; e += P / 4) {
      const int cgp = e >> 4, rr = (e >> 2) & 3, cc = e & 3;
      lds[rr * LS + lpad(cgp * 4 + cc)] = tc[((int64_t)cgp * nrg + rg) * 16 + rr * 4 + cc];
    }
    __syncthreads();
    fft_lds<P, TPF, true>(s, t);
    if (row < a.n) {
      float2* oc = a.out + (int64_t)c * a.out_cs + (int64_t)row * a.n;
#pragma unroll
      for (int i = 0; i < PPT; ++i) {
        const int col = t + i * TPF;
        if (col < a.n) {
          const float2 v = s[lpad(col)];
          oc[col] = v;
          inten[i] += v.x * v.x + v.y * v.y;
        }
      }
    }
    __syncthreads();
  }
  if (a.intensity && row < a.n) {
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int col = t + i * TPF;
      if (col < a.n) a.intensity[(int64_t)row * a.n + col] = inten[i];
    }
  }
}

template <int P>
inline hipError_t rs_launch(const RsArgs& a, hipStream_t st) {
  const int nrg = (a.n + 3) >> 2;
  dim3 blk(P / 4);
  { KTimer t(K_RS_HROW, st); hipLaunchKernelGGL(rs_hrow_kernel<P>, dim3(P / 8), blk, 0, st, a); }
  { KTimer t(K_RS_UROW, st); hipLaunchKernelGGL(rs_urow_kernel<P>, dim3(nrg), blk, 0, st, a); }
  { KTimer t(K_RS_COL, st); hipLaunchKernelGGL(rs_col_kernel<P>, dim3(P / 4), blk, 0, st, a); }
  { KTimer t(K_RS_FINAL, st); hipLaunchKernelGGL(rs_final_kernel<P>, dim3(nrg), blk, 0, st, a); }
  return hipGetLastError();
}

}  // namespace xl

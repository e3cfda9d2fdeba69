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
//   rs_hcol   : per column: FFT(Hr col, mirrored) -> Hhat, in place (half spectrum)
//   rs_col    : per (column, component): FFT(T col) * Hhat -> IFFT, keep rows
//               0..N-1 -> T
//   rs_final  : row IFFT of T, keep columns 0..N-1 -> output (+ |E|^2)
#pragma once
#include "fft.h"
#include "fftio.h"
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
  float2* hr;             // tiles, (P/2) x (P/2 + 1) of a (P/2) x P array: Hr after the row pass,
                          // Hhat (ky < P/2) after rs_hcol; columns kx > P/2 mirror (Hhat is even in kx)
  float2* hmid;           // P/2 + 1: Hhat at ky = P/2
  const double* x;
  const double* y;
  double z, k;
  const double* z_dev;    // if set, z is read from device memory
  float scale;            // dx*dy/P^2
  int n, nin, ncomp, ez_mode, adjoint;
  int kernel_dz;          // 1: convolve with dH/dz instead of H, 2: with dH/dk
  int hstored;            // the spectrum lives in a caller buffer, unconjugated (adjoints conjugate on read)
  int hreuse;             // the spectrum is already there: skip rs_hrow / rs_hcol
  MaskArgs mask;
  // Batch of independent problems of one geometry (blockIdx.y = member): element strides
  // between the members' inputs, outputs, intensities, mask parameters, tiles and kernel
  // spectra; with z_dev set, member b reads z_dev[b].
  int batch;
  int64_t in_bs, out_bs, inten_bs, mask_bs, t_bs, hr_bs;
  // Gradients by Parseval (one problem).  spec_mode 1: rs_col keeps the input
  // column spectra in `spec`, unit-major in the hand-off register order
  // (unit u, slot j, thread t at u P + j TPF + t).  spec_mode 2: rs_col adds
  // Re(conj(G) dg[i] U) over its column into gpart[i][unit] for each
  // derivative spectrum dg[i] (Hhat layout) that is set.
  int spec_mode;
  int dg_built;           // spec_mode 2: the derivative spectra were built by the forward call
  int partials_only;      // spec_mode 2: leave the partial sums in gpart (no grad_out reduction)
  float2* spec;
  float2* dg[2];          // spec_mode 1: derivative spectra to build with the kernel spectrum
  double* gpart[2];
  double* grad_out;       // spec_mode 2: out[i] = sum of gpart[i]
  double* qf_out;         // optional: quality factor of member b at qf_out[b]
  double qf_rmax, qf_dr, qf_wl;
};

// Quality factor of a propagation to z (wave_optics.py:187-191).
__device__ __forceinline__ double rs_quality_factor(double z, double rmax, double dr_real, double wl) {
  return (sqrt(wl * wl + rmax * rmax + 2.0 * wl * sqrt(rmax * rmax + z * z)) - rmax) / dr_real;
}

// The arguments of batch member b (pointers offset by the batch strides).
__device__ __forceinline__ RsArgs rs_member(const RsArgs& a0, int b) {
  if (a0.batch <= 1) return a0;
  RsArgs a = a0;
  a.in += b * a0.in_bs;
  a.out += b * a0.out_bs;
  if (a.intensity) a.intensity += b * a0.inten_bs;
  a.t += b * a0.t_bs;
  a.hr += b * a0.hr_bs;
  a.hmid += b * a0.hr_bs;
  if (a.z_dev) a.z_dev += b;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (a.mask.st[i].m0) a.mask.st[i].m0 += b * a0.mask_bs;
    if (a.mask.st[i].m1) a.mask.st[i].m1 += b * a0.mask_bs;
  }
  return a;
}

// lag value of circular slot s on a P-point axis with N samples (0 => unused)
__device__ __forceinline__ bool lag_of(const double* g, int s, int n, int P, double* d) {
  if (s < n) { *d = g[s] - g[0]; return true; }
  if (s > P - n) { *d = g[P - s] - g[0]; return true; }
  return false;
}

// 1/sqrt(x) in float64 for x > 0: the hardware estimate (v_rsq_f64) refined by
// two Newton steps y <- y (3/2 - x y^2 / 2), which take any estimate within
// 2^-20 to the last bits of float64; the kernel values need r = x / sqrt(x) to
// ~1e-14 relative, for phases k r of up to ~1e6 rad.  One estimate and six
// fused multiply-adds replace a correctly rounded square root and a division.
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  y = y * fma(-hx * y, y, 1.5);
  return y;
}

// h(x, y; z) of wave_optics.py:291-297 with the phase k r reduced in float64.
__device__ __forceinline__ float2 rs_kernel_value(double X, double Y, double z, double k) {
  const double r2 = X * X + Y * Y + z * z;
  const double ir = rsqrt_f64(r2);
  const double r = r2 * ir;
  const double turns = k * r * 0.15915494309189535;   // k r / (2 pi)
  const float2 e = expi_turns(z > 0 ? turns : -turns);
  const double f = 0.15915494309189535 * z * ir * ir;   // (1/2pi) z / r^2
  const float fr = (float)(f * ir), fi = (float)(-f * k);   // f * (1/r - i k)
  return make_float2(fr * e.x - fi * e.y, fr * e.y + fi * e.x);
}

// dH/dz of rs_kernel_value (product rule, dr/dz = z/r, s = sign of the phase):
//   (1/2pi) e^{i s k r} [ (r^2 - 3z^2)/r^5 + s k^2 z^2/r^3 + i k (s z^2 - (r^2 - 2 z^2))/r^4 ].
__device__ __forceinline__ float2 rs_kernel_dz_value(double X, double Y, double z, double k) {
  const double r2 = X * X + Y * Y + z * z;
  const double ir = rsqrt_f64(r2), ir2 = ir * ir;
  const double r = r2 * ir;
  const double s = z > 0 ? 1.0 : -1.0;
  const double turns = k * r * 0.15915494309189535;
  const float2 e = expi_turns(s * turns);
  const double z2 = z * z;
  const double cr = (r2 - 3.0 * z2) * ir2 * ir2 * ir + s * k * k * z2 * ir2 * ir;
  const double ci = k * (s * z2 - (r2 - 2.0 * z2)) * ir2 * ir2;
  const float fr = (float)(0.15915494309189535 * cr), fi = (float)(0.15915494309189535 * ci);
  return make_float2(fr * e.x - fi * e.y, fr * e.y + fi * e.x);
}

// dH/dk of rs_kernel_value (k = 2 pi / lambda; r and z do not depend on k):
//   (1/2pi)(z/r^2) e^{i s k r} [ s k r + i (s - 1) ].
__device__ __forceinline__ float2 rs_kernel_dk_value(double X, double Y, double z, double k) {
  const double r2 = X * X + Y * Y + z * z;
  const double ir = rsqrt_f64(r2);
  const double r = r2 * ir;
  const double s = z > 0 ? 1.0 : -1.0;
  const double turns = k * r * 0.15915494309189535;
  const float2 e = expi_turns(s * turns);
  const double f = 0.15915494309189535 * z * ir * ir;
  const float fr = (float)(f * s * k * r), fi = (float)(f * (s - 1.0));
  return make_float2(fr * e.x - fi * e.y, fr * e.y + fi * e.x);
}

// h with its derivatives in z (DM bit 0) and k (DM bit 1) at one lag, the
// geometry (r, 1/r, the phase factor) formed once for all three; each value
// is the expression of rs_kernel_value / rs_kernel_dz_value /
// rs_kernel_dk_value.
template <int DM>
__device__ __forceinline__ void rs_kernel_with_derivatives(double X, double Y, double z, double k, float2& h,
                                                           float2& hz, float2& hk) {
  const double r2 = X * X + Y * Y + z * z;
  const double ir = rsqrt_f64(r2), ir2 = ir * ir;
  const double r = r2 * ir;
  const double s = z > 0 ? 1.0 : -1.0;
  const double turns = k * r * 0.15915494309189535;
  const float2 e = expi_turns(s * turns);
  auto rot = [&](float fr, float fi) { return make_float2(fr * e.x - fi * e.y, fr * e.y + fi * e.x); };
  const double f = 0.15915494309189535 * z * ir * ir;
  h = rot((float)(f * ir), (float)(-f * k));
  if constexpr ((DM & 1) != 0) {
    const double z2 = z * z;
    const double cr = (r2 - 3.0 * z2) * ir2 * ir2 * ir + s * k * k * z2 * ir2 * ir;
    const double ci = k * (s * z2 - (r2 - 2.0 * z2)) * ir2 * ir2;
    hz = rot((float)(0.15915494309189535 * cr), (float)(0.15915494309189535 * ci));
  }
  if constexpr ((DM & 2) != 0) hk = rot((float)(f * s * k * r), (float)(f * (s - 1.0)));
}

// The kept input spectra are written once and read once, by another launch:
// non-temporal accesses, so that they do not displace the tile lines the
// column pass shares through the caches.
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ float2 stream_load(const float2* p) {
  return make_float2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}

// Hr tile index of (row r, column c): P/8 row groups
__device__ __forceinline__ int64_t hr_idx(int r, int c, int P) {
  return ((int64_t)(c >> 2) * (P / 8) + (r >> 2)) * 16 + (r & 3) * 4 + (c & 3);
}

// Rows or columns per block for the one-FFT-per-group kernels: an FFT runs on
// P/16 threads (one radix-16 butterfly each per pass) in one LDS buffer of
// 35 KB at P = 4096, so four blocks share a CU and their pass barriers
// interleave; below a wave of threads per FFT a block takes several.
template <int P> struct Units {
  static constexpr int TPF = P / 16;
  static constexpr int CPB = TPF >= 64 ? 1 : 64 / TPF;
};

// One kernel row per group: H of rows 0..P/2-1 generated in float64 phase
// (the array is even in x: columns above P/2 mirror), row FFT, stored as Hr
// for kx <= P/2 (the transform of an even row is even).  DM != 0: the rows
// of the derivative kernels in the bit set DM (bit 0 dH/dz, bit 1 dH/dk) are
// formed from the same geometry, held in registers while H is transformed,
// then transformed in turn into dg[i] (a forward call that will be
// differentiated in z / lambda builds its derivative spectra this way).
template <int P, int DM = 0>
__global__ __launch_bounds__(Units<P>::CPB * (P / 16), 4) void rs_hrow_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value, CPB = Units<P>::CPB;
  constexpr int NCOL = (P / 2 + 1 + TPF - 1) / TPF;   // columns kx <= P/2 per thread
  __shared__ float2 lds[CPB * LS];
  const int g = threadIdx.x / TPF, t = threadIdx.x % TPF;
  float2* s = lds + g * LS;
  const int r = xcd_remap(blockIdx.x, gridDim.x) * CPB + g;      // rows 0..P/2-1
  const double z = a.z_dev ? *a.z_dev : a.z;
  double dy;
  const bool rowok = lag_of(a.y, r, a.n, P, &dy);
  float2 dv[2][NCOL];
#pragma unroll
  for (int i = 0; i < NCOL; ++i) {
    const int c = t + i * TPF;
    if (c > P / 2) break;
    double dx;
    float2 v = make_float2(0.f, 0.f), vz = v, vk = v;
    if (rowok && lag_of(a.x, c, a.n, P, &dx)) {
      if constexpr (DM == 0)
        v = a.kernel_dz == 1   ? rs_kernel_dz_value(dx, dy, z, a.k)
            : a.kernel_dz == 2 ? rs_kernel_dk_value(dx, dy, z, a.k)
                               : rs_kernel_value(dx, dy, z, a.k);
      else
        rs_kernel_with_derivatives<DM>(dx, dy, z, a.k, v, vz, vk);
    }
    s[lpad(c)] = v;
    if (c > 0 && c < P / 2) s[lpad(P - c)] = v;
    dv[0][i] = vz;
    dv[1][i] = vk;
  }
  __syncthreads();
  NoIO none;
  typename LastPass<P, TPF, true>::Type unused;
  {
    auto store = [&](int c, float2 v) {
      if (c <= P / 2) a.hr[hr_idx(r, c, P)] = v;
    };
    fft_io<P, TPF, false, SRC_LDS, DST_FN>(s, t, none, store, unused);
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    if constexpr (DM == 0) break;
    if (!(DM >> d & 1)) continue;
    __syncthreads();   // the last pass of other waves may still read LDS
#pragma unroll
    for (int i = 0; i < NCOL; ++i) {
      const int c = t + i * TPF;
      if (c > P / 2) break;
      s[lpad(c)] = dv[d][i];
      if (c > 0 && c < P / 2) s[lpad(P - c)] = dv[d][i];
    }
    __syncthreads();
    float2* dst = a.dg[d];
    auto store = [&](int c, float2 v) {
      if (c <= P / 2) dst[hr_idx(r, c, P)] = v;
    };
    fft_io<P, TPF, false, SRC_LDS, DST_FN>(s, t, none, store, unused);
  }
}

// One block per input row: the masks / Ez of each pixel are computed once by
// one thread for all components, component c is transformed by threads
// [c*TPF, (c+1)*TPF) in LDS buffer c.  Rows 4g..4g+3 run back to back on one
// XCD, so the four 32 B row slices of each 128 B tile meet in the same L2.
template <int P>
__global__ __launch_bounds__(3 * (P / 16)) void rs_urow_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int f = tid / TPF, t = tid % TPF;
  const int nrg = (a.n + 3) >> 2;
  const int row = xcd_remap(blockIdx.x, gridDim.x);
  const int nc = a.ncomp;
  const double z = a.z_dev ? *a.z_dev : a.z;
  for (int col = tid; col < P / 2; col += nthr) {
    float2 e0 = make_float2(0.f, 0.f), e1 = e0, e2 = e0;
    if (row < a.n && col < a.n) {
      const int64_t pix = (int64_t)row * a.n + col;
      e0 = a.in[pix];
      if (a.nin > 1) e1 = a.in[a.in_cs + pix];
      if (a.nin > 2) e2 = a.in[2 * a.in_cs + pix];
      apply_mask(a.mask, pix, e0, e1, e2);
      if (a.ez_mode != EZ_NONE) e2 = build_ez(a.ez_mode, e0, e1, a.x[col], a.y[row], z);
    }
    lds[lpad(col)] = e0;
    lds[lpad(col + P / 2)] = make_float2(0.f, 0.f);
    if (nc > 1) { lds[LS + lpad(col)] = e1; lds[LS + lpad(col + P / 2)] = make_float2(0.f, 0.f); }
    if (nc > 2) { lds[2 * LS + lpad(col)] = e2; lds[2 * LS + lpad(col + P / 2)] = make_float2(0.f, 0.f); }
  }
  __syncthreads();
  const int rg = row >> 2, rr = row & 3;
  float2* dst = a.t + (int64_t)f * nrg * 4 * P + rg * 16 + rr * 4;
  NoIO none;
  auto store = [&](int col, float2 v) { dst[(int64_t)(col >> 2) * nrg * 16 + (col & 3)] = v; };
  typename LastPass<P, TPF, true>::Type unused;
  fft_io<P, TPF, false, SRC_LDS, DST_FN>(lds + f * LS, t, none, store, unused);
}

// One input row per block of P/16 threads and one 35 KB buffer, so several
// blocks share a CU.  Thread t's first-pass inputs are the columns
// t + r P/16; the nonzero ones (columns < N <= P/2) are r < 8, so each thread
// computes the masks and Ez of exactly those 8 pixels once, keeps all
// components in registers, and the components are transformed one after the
// other from there (no LDS staging pass).
// ONE (inputs without masks or Ez construction: the adjoint's cotangent): one
// (row, component) unit per block, the components of a row on consecutive
// blocks, for grids whose row blocks alone leave resident slots empty.
template <int P, bool ONE = false>
__global__ __launch_bounds__(P / 16) void rs_urow1_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  constexpr int R0 = FftPlan<P>::radix(true, 0);
  static_assert(R0 == 16 && P / R0 == TPF, "first pass: one radix-16 butterfly per thread");
  constexpr int H = R0 / 2;   // inputs r < H hold the N <= P/2 pixels
  __shared__ float2 lds[LS];
  const int t = threadIdx.x;
  if constexpr (ONE) {
    const int unit = xcd_remap(blockIdx.x, gridDim.x);
    const int row = unit / a.ncomp, c = unit - row * a.ncomp;
    const int nrg = (a.n + 3) >> 2, rg = row >> 2, rr = row & 3;
    const float2* src = a.in + (int64_t)c * a.in_cs;
    Regs<1, R0> rin;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int col = t + r * TPF;
      rin.v[0][r] = (r < H && row < a.n && col < a.n) ? src[(int64_t)row * a.n + col] : make_float2(0.f, 0.f);
    }
    NoIO none;
    float2* dst = a.t + (int64_t)c * nrg * 4 * P + rg * 16 + rr * 4;
    auto store = [&](int col, float2 v) { dst[(int64_t)(col >> 2) * nrg * 16 + (col & 3)] = v; };
    typename LastPass<P, TPF, true>::Type unused;
    chain_io<P, TPF, false, true, 0, SRC_REG, DST_FN, true, false>(lds, t, none, store, rin, unused);
    return;
  }
  const int row = xcd_remap(blockIdx.x, gridDim.x);
  const int nrg = (a.n + 3) >> 2, rg = row >> 2, rr = row & 3;
  const int nc = a.ncomp;
  const double z = a.z_dev ? *a.z_dev : a.z;
  // all loads of the 8 pixels first (in flight together), then the masks and Ez
  const bool rowok = row < a.n, ez = a.ez_mode != EZ_NONE;
  PixRaw raw[H];
  int64_t pix[H];
  double xc[H];
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const int col = t + r * TPF;
    const bool ok = rowok && col < a.n;
    pix[r] = ok ? (int64_t)row * a.n + col : 0;
    xc[r] = ez ? a.x[ok ? col : 0] : 0.0;
  }
  pix_load_n<H>(a.in, a.in_cs, a.nin, a.mask, pix, raw);
  const double yr = ez ? a.y[rowok ? row : 0] : 0.0;
  float2 e[3][H];
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const int col = t + r * TPF;
    float2 e0 = make_float2(0.f, 0.f), e1 = e0, e2 = e0;
    if (rowok && col < a.n) {
      pix_masked(a.mask, raw[r], e0, e1, e2);
      if (ez) e2 = build_ez(a.ez_mode, e0, e1, xc[r], yr, z);
    }
    e[0][r] = e0;
    e[1][r] = e1;
    e[2][r] = e2;
  }
  NoIO none;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= nc) break;
    Regs<1, R0> rin;
#pragma unroll
    for (int r = 0; r < R0; ++r) rin.v[0][r] = r < H ? e[c][r] : make_float2(0.f, 0.f);
    float2* dst = a.t + (int64_t)c * nrg * 4 * P + rg * 16 + rr * 4;
    auto store = [&](int col, float2 v) { dst[(int64_t)(col >> 2) * nrg * 16 + (col & 3)] = v; };
    typename LastPass<P, TPF, true>::Type unused;
    chain_io<P, TPF, false, true, 0, SRC_REG, DST_FN, true, false>(lds, t, none, store, rin, unused);
  }
}

// Kernel spectrum.  Per column kx <= P/2: the Hr column (rows mirrored, row
// P/2 zero) is transformed along y and the spectrum, times dx dy / P^2
// (conjugated for the adjoint), replaces the Hr column in place for ky < P/2;
// ky = P/2 goes to hmid[col].  The array is even in y and in x, so ky > P/2
// and kx > P/2 are not formed (Hhat(P - ky, kx) = Hhat(ky, P - kx) = Hhat(ky, kx)).
// Blocks beyond the P/2 + 1 columns (grid padding for the XCD remap) idle.
template <int P>
__global__ __launch_bounds__(Units<P>::CPB * (P / 16), 4) void rs_hcol_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value, CPB = Units<P>::CPB;
  using HO = HandOff<P, TPF>;
  __shared__ float2 lds[CPB * LS];
  const int g = threadIdx.x / TPF, t = threadIdx.x % TPF;
  float2* s = lds + g * LS;
  const int col = xcd_remap(blockIdx.x, gridDim.x) * CPB + g;
  const bool act = col <= P / 2;
  if (CPB == 1 && !act) return;   // whole block
  const int cg = col >> 2, cc = col & 3;
  float2* hc = a.hr + (int64_t)cg * (P / 8) * 16 + cc;
  auto load = [&](int r) -> float2 {
    const int rr = r <= P / 2 ? r : P - r;
    return (!act || rr == P / 2) ? make_float2(0.f, 0.f) : hc[(rr >> 2) * 16 + (rr & 3) * 4];
  };
  typename HO::Type u;
  fft_forward_to_regs<P, TPF, SRC_FN>(s, t, load, u);
#pragma unroll
  for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
    for (int q = 0; q < HO::R; ++q) {
      const int ky = HO::index(t, b, q);
      float2 h = u.v[b][q] * a.scale;
      if (a.adjoint && !a.hstored) h = conjf2(h);
      if (!act) continue;
      if (ky < P / 2) hc[(ky >> 2) * 16 + (ky & 3) * 4] = h;
      else if (ky == P / 2) a.hmid[col] = h;
    }
}

// Sum of v over the TPF threads of each group (groups of TPF < 64 share a
// wave, aligned to TPF lanes), fixed order; every thread of the block calls
// it.  `red` holds CPB * (TPF / 64) values when TPF > 64.
template <int TPF>
__device__ __forceinline__ double group_sum_f64(double v, double* red, int g, int t) {
  constexpr int W = TPF < 64 ? TPF : 64;
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if constexpr (TPF > 64) {
    constexpr int NW = TPF / 64;
    if ((t & 63) == 0) red[g * NW + (t >> 6)] = v;
    __syncthreads();
    v = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) v += red[g * NW + w];
  }
  return v;
}

// spec_mode 2 of the column pass, after its multiply loop has accumulated
// acc[i] = sum over the thread's ky of Re(conj(G) dg[i] U) (G the spectrum in
// registers, the cotangent's; U the kept input spectrum; dg[i] the derivative
// spectrum, even in kx and ky like Hhat): the unit's sums, fixed order, into
// gpart[i][unit] for i in the bit set GB.
template <int TPF, int CPB, int GB>
__device__ __forceinline__ void rs_col_dot_sums(const RsArgs& a, const double (&acc)[2], int unit, int g, int t) {
  constexpr int NW = TPF > 64 ? TPF / 64 : 1;
  __shared__ double red[2][CPB * NW];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (!(GB >> i & 1)) continue;
    const double v = group_sum_f64<TPF>(acc[i], red[i], g, t);
    if (t == 0) a.gpart[i][unit] = v;
  }
}

// Column pass.  One (column, component) unit per thread group: the T column
// (rows 0..N-1, zero padded) is transformed along y straight from the tiles,
// multiplied in registers by the kernel spectrum of its column (read from
// the in-place Hhat, mirrored for ky > P/2), inverse-transformed, and rows
// 0..N-1 stored back in place.  A block holds one 4096-point FFT (35 KB of
// LDS), so four blocks share a CU and their pass barriers interleave.  The
// units of a column and of its tile neighbours run back to back on one XCD,
// so the Hhat column and the 128 B tile lines are fetched once into its L2.
// SPEC 1 also keeps the column's spectrum; SPEC = 2 + GB, GB in 1..3, adds
// its gradient dots for the derivative spectra in the bit set GB (spec_mode 2
// of RsArgs).
template <int P, int SPEC>
__global__ __launch_bounds__(Units<P>::CPB * (P / 16), 4) void rs_col_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  if (a.qf_out && blockIdx.x == 0 && threadIdx.x == 0)
    a.qf_out[blockIdx.y] = rs_quality_factor(a.z_dev ? *a.z_dev : a.z, a.qf_rmax, a.qf_dr, a.qf_wl);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value, CPB = Units<P>::CPB;
  using HO = HandOff<P, TPF>;
  __shared__ float2 lds[CPB * LS];
  const int g = threadIdx.x / TPF, t = threadIdx.x % TPF;
  float2* s = lds + g * LS;
  const int unit = xcd_remap(blockIdx.x, gridDim.x) * CPB + g;
  const int col = unit / a.ncomp, comp = unit - col * a.ncomp;
  const bool active = col < P;
  const int nrg = (a.n + 3) >> 2;
  const int nr = nrg * 4;
  const int cg = col >> 2, cc = col & 3;
  const int hcol = col <= P / 2 ? col : P - col;   // Hhat is even in kx
  const float2* hc = a.hr + (int64_t)(hcol >> 2) * (P / 8) * 16 + (hcol & 3);
  float2* tc = a.t + (int64_t)comp * nrg * 4 * P + (int64_t)cg * nrg * 16 + cc;
  auto load = [&](int r) -> float2 {
    return (active && r < nr) ? tc[(r >> 2) * 16 + (r & 3) * 4] : make_float2(0.f, 0.f);
  };
  typename HO::Type u;
  constexpr int GB = SPEC >= 3 ? SPEC - 2 : 0;
  // the gradient dots (SPEC >= 3) read the kept input spectrum U: up to
  // P = 2048 its loads issue before the forward transform and complete behind
  // it (at 4096 they would not fit beside the transform's registers)
  constexpr bool UPRE = GB != 0 && P <= 2048;
  typename HO::Type uk;
  if constexpr (UPRE) {
    const float2* sp = a.spec + (int64_t)unit * P + t;
#pragma unroll
    for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
      for (int q = 0; q < HO::R; ++q) uk.v[b][q] = active ? stream_load(sp + (b * HO::R + q) * TPF) : make_float2(0.f, 0.f);
  }
  // rows >= N <= P/2 are zero padding, and the rows kept are < N: the first
  // forward pass skips the upper half of its inputs, the last inverse pass
  // the upper half of its outputs
  fft_forward_to_regs<P, TPF, SRC_FN, true>(s, t, load, u);
  if constexpr (SPEC == 1) {
    if (active) {
      float2* sp = a.spec + (int64_t)unit * P + t;
#pragma unroll
      for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
        for (int q = 0; q < HO::R; ++q) stream_store(sp + (b * HO::R + q) * TPF, u.v[b][q]);
    }
  }
  const float hs = (a.adjoint && a.hstored) ? -1.f : 1.f;   // stored spectrum: conjugate for the adjoint
  double acc[2] = {0.0, 0.0};
  if (active) {
    // dg is read in the same loop as the kernel spectrum, at the same Hhat offsets
#pragma unroll
    for (int b = 0; b < HO::BPT; ++b)
#pragma unroll
      for (int q = 0; q < HO::R; ++q) {
        const int ky = HO::index(t, b, q);
        const int kk = ky <= P / 2 ? ky : P - ky;
        const int64_t ho = (int64_t)(hcol >> 2) * (P / 8) * 16 + (hcol & 3) + (kk >> 2) * 16 + (kk & 3) * 4;
        float2 h = kk == P / 2 ? a.hmid[hcol] : hc[(kk >> 2) * 16 + (kk & 3) * 4];
        h.y *= hs;
        if constexpr (GB != 0) {
          const float2 uv = UPRE ? uk.v[b][q] : stream_load(a.spec + (int64_t)unit * P + t + (b * HO::R + q) * TPF);
          const float2 gv = u.v[b][q];
          const int64_t di = kk == P / 2 ? (int64_t)(P / 2) * P + hcol : ho;
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            if (!(GB >> i & 1)) continue;
            const float2 w = cmul(a.dg[i][di], uv);
            acc[i] += (double)(gv.x * w.x + gv.y * w.y);
          }
        }
        u.v[b][q] = cmul(u.v[b][q], h);
      }
  }
  auto store = [&](int r, float2 v) {
    if (active && r < nr) tc[(r >> 2) * 16 + (r & 3) * 4] = v;
  };
  fft_inverse_from_regs<P, TPF, DST_FN, true>(s, t, u, store);
  if constexpr (GB != 0) rs_col_dot_sums<TPF, CPB, GB>(a, acc, unit, g, t);
}

// Per block 2 rows (one per thread group): the row halves of
// each component's tiles are staged into LDS with 64 B loads (the block
// holding the other two rows of a tile runs next on the same XCD and finds
// the line in L2), inverse-transformed in LDS, and the first N samples stored;
// |E|^2 is accumulated over the components.  Two 35 KB buffers per block let
// two blocks share a CU.
template <int P>
__global__ __launch_bounds__(2 * FftThreads<P>::value) void rs_final_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = FftThreads<P>::value, LS = LdsLen<P>::value;
  constexpr int NT = 2 * TPF;
  constexpr int PPT = (P / 2 + TPF - 1) / TPF;
  __shared__ float2 lds[2 * LS];
  const int tid = threadIdx.x, f = tid / TPF, t = tid % TPF;
  float2* s = lds + f * LS;
  const int nrg = (a.n + 3) >> 2;
  const int pair = xcd_remap(blockIdx.x, gridDim.x);
  const int rg = pair >> 1, r0 = (pair & 1) * 2;
  const int row = rg * 4 + r0 + f;
  float inten[PPT];
#pragma unroll
  for (int i = 0; i < PPT; ++i) inten[i] = 0.f;
  for (int c = 0; c < a.ncomp; ++c) {
    const float2* tc = a.t + (int64_t)c * nrg * 4 * P;
#pragma unroll 4
    for (int e = tid; e < 2 * P; e += NT) {
      const int cgp = e >> 3, rr = (e >> 2) & 1, cc = e & 3;
      lds[rr * LS + lpad(cgp * 4 + cc)] = tc[((int64_t)cgp * nrg + rg) * 16 + (r0 + rr) * 4 + cc];
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

// One output row per block of P/16 threads and one 35 KB buffer, so four
// blocks share a CU: the components are inverse-transformed one after the
// other, each read straight from its tiles into the first pass (32 B of each
// 128 B line; the blocks of the tile's other rows run next on the same XCD),
// each last pass left in registers, where the first N samples are stored and
// |E|^2 is accumulated over the components.
// ONE: one (component, row) unit per block (no |E|^2 to accumulate across
// components), component-major so a tile's four rows still run back to back:
// ncomp N blocks divide into the resident slots evenly where N row blocks
// (2.67 rounds at P = 4096) leave a partly idle last round.
template <int P, bool ONE = false>
__global__ __launch_bounds__(P / 16) void rs_final1_kernel(RsArgs a0) {
  const RsArgs a = rs_member(a0, blockIdx.y);
  constexpr int TPF = P / 16, LS = LdsLen<P>::value;
  using LP = LastPass<P, TPF, false>;
  __shared__ float2 lds[LS];
  const int t = threadIdx.x;
  const int unit = xcd_remap(blockIdx.x, gridDim.x);
  const int c0 = ONE ? unit / a.n : 0;
  const int row = ONE ? unit - c0 * a.n : unit;
  if (row >= a.n || c0 >= a.ncomp) return;   // whole block
  const int nrg = (a.n + 3) >> 2, rg = row >> 2, rr = row & 3;
  // output columns < N <= P/2 are slots q < R/2 of the last pass
  constexpr int QH = LP::R / 2;
  float inten[LP::BPT][QH];
#pragma unroll
  for (int b = 0; b < LP::BPT; ++b)
#pragma unroll
    for (int q = 0; q < QH; ++q) inten[b][q] = 0.f;
  NoIO none;
  for (int c = c0; c < (ONE ? c0 + 1 : a.ncomp); ++c) {
    const float2* tc = a.t + (int64_t)c * nrg * 4 * P + rg * 16 + rr * 4;
    auto load = [&](int col) -> float2 { return tc[(int64_t)(col >> 2) * nrg * 16 + (col & 3)]; };
    typename LP::Type o, unused;
    // the kept columns are < N <= P/2: the last pass forms the lower half only
    chain_io<P, TPF, true, false, 0, SRC_FN, DST_REG, false, true>(lds, t, load, none, unused, o);
    float2* oc = a.out + (int64_t)c * a.out_cs + (int64_t)row * a.n;
#pragma unroll
    for (int b = 0; b < LP::BPT; ++b) {
      if (!LP::valid(t, b)) continue;
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        const int col = LP::index(t, b, q);
        if (col < a.n) {
          const float2 v = o.v[b][q];
          oc[col] = v;
          if constexpr (!ONE) inten[b][q] += v.x * v.x + v.y * v.y;
        }
      }
    }
  }
  if (!ONE && a.intensity) {
#pragma unroll
    for (int b = 0; b < LP::BPT; ++b) {
      if (!LP::valid(t, b)) continue;
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        const int col = LP::index(t, b, q);
        if (col < a.n) a.intensity[(int64_t)row * a.n + col] = inten[b][q];
      }
    }
  }
}

// out[i] = sum of the n per-unit partials of gradient i (block i of 1024
// threads; a null partial array skips its block), fixed order.
__global__ __launch_bounds__(1024) void rs_grad_reduce_kernel(const double* p0, const double* p1, int n,
                                                              double* out) {
  const double* p = blockIdx.x == 0 ? p0 : p1;
  if (!p) return;   // whole block
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += p[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < 16; ++w) v += red[w];
    out[blockIdx.x] = v;
  }
}

// Partial sums written by rs_col_kernel<P, 2> for ncomp components (its
// padded unit count).
template <int P> inline int rs_col_units(int ncomp) {
  constexpr int CPB = Units<P>::CPB;
  return (P * ncomp + CPB - 1) / CPB * CPB;
}

// Dynamic-LDS limits of the multi-buffer row kernels (up to three buffers of
// P points), set once per device by the library's initialisation.
template <int P> inline hipError_t rs_set_attribute() {
  return hipFuncSetAttribute((const void*)rs_urow_kernel<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             3 * LdsLen<P>::value * (int)sizeof(float2));
}
inline hipError_t rs_set_attributes() {
  hipError_t e = hipSuccess;
  for (hipError_t r : {rs_set_attribute<64>(), rs_set_attribute<128>(), rs_set_attribute<256>(),
                       rs_set_attribute<512>()})
    if (r != hipSuccess) e = r;
  return e;
}

template <int P>
inline hipError_t rs_launch(const RsArgs& a, hipStream_t st) {
  constexpr int CPB = Units<P>::CPB;
  const dim3 blk(CPB * (P / 16));
  const int rows = ((a.n + 3) >> 2) * 4;
  const int nb = a.batch > 1 ? a.batch : 1;
  // a forward call that keeps its spectra for a gradient builds the derivative
  // spectra with its own kernel rows (one float64 geometry for all of them)
  const int dm = (a.dg[0] ? 1 : 0) | (a.dg[1] ? 2 : 0);
  const bool with_h = a.spec_mode == 1 && dm && !a.hreuse;
  if (!a.hreuse) {
    KTimer t(K_RS_HROW, st);
    if (with_h && dm == 1)
      hipLaunchKernelGGL((rs_hrow_kernel<P, 1>), dim3(P / 2 / CPB, nb), blk, 0, st, a);
    else if (with_h && dm == 2)
      hipLaunchKernelGGL((rs_hrow_kernel<P, 2>), dim3(P / 2 / CPB, nb), blk, 0, st, a);
    else if (with_h)
      hipLaunchKernelGGL((rs_hrow_kernel<P, 3>), dim3(P / 2 / CPB, nb), blk, 0, st, a);
    else
      hipLaunchKernelGGL((rs_hrow_kernel<P, 0>), dim3(P / 2 / CPB, nb), blk, 0, st, a);
  }
  const int ncb = (P / 2 + 1 + CPB - 1) / CPB;   // columns kx <= P/2, grid padded to whole XCD rounds
  if ((a.spec_mode == 2 && !a.dg_built) || (a.spec_mode == 1 && dm && !with_h)) {
    // derivative spectra (dH/dz, dH/dk) with the spectrum passes, unconjugated
    for (int i = 0; i < 2; ++i) {
      if (!a.dg[i]) continue;
      RsArgs d = a;
      d.hr = a.dg[i];
      d.hmid = a.dg[i] + (int64_t)(P / 2) * P;
      d.kernel_dz = i + 1;
      d.adjoint = 0;
      d.hstored = 1;
      KTimer t(K_RS_HROW, st);
      hipLaunchKernelGGL((rs_hrow_kernel<P, 0>), dim3(P / 2 / CPB, nb), blk, 0, st, d);
      hipLaunchKernelGGL(rs_hcol_kernel<P>, dim3((ncb + 7) / 8 * 8, nb), blk, 0, st, d);
    }
  }
  if constexpr (P >= 1024) {
    KTimer t(K_RS_UROW, st);
    bool one = false;
    if constexpr (P <= 2048) {
      // plain inputs (no masks, no Ez) over fewer row blocks than resident
      // slots: one block per (row, component)
      one = a.ncomp > 1 && a.nin == a.ncomp && a.ez_mode == EZ_NONE && !a.mask.st[0].kind && !a.mask.st[1].kind &&
            rows * nb < 2048;
      if (one) hipLaunchKernelGGL((rs_urow1_kernel<P, true>), dim3(rows * a.ncomp, nb), dim3(P / 16), 0, st, a);
    }
    if (!one) hipLaunchKernelGGL((rs_urow1_kernel<P, false>), dim3(rows, nb), dim3(P / 16), 0, st, a);
  } else {
    KTimer t(K_RS_UROW, st);
    hipLaunchKernelGGL(rs_urow_kernel<P>, dim3(rows, nb), dim3(a.ncomp * (P / 16)),
                       (size_t)a.ncomp * LdsLen<P>::value * sizeof(float2), st, a);
  }
  if (!a.hreuse) {
    KTimer t(K_RS_HCOL, st);
    hipLaunchKernelGGL(rs_hcol_kernel<P>, dim3((ncb + 7) / 8 * 8, nb), blk, 0, st, a);
    for (int i = 0; with_h && i < 2; ++i) {
      if (!a.dg[i]) continue;
      RsArgs d = a;   // the derivative row transforms, unconjugated
      d.hr = a.dg[i];
      d.hmid = a.dg[i] + (int64_t)(P / 2) * P;
      d.adjoint = 0;
      d.hstored = 1;
      hipLaunchKernelGGL(rs_hcol_kernel<P>, dim3((ncb + 7) / 8 * 8, nb), blk, 0, st, d);
    }
  }
  {
    KTimer t(K_RS_COL, st);
    const dim3 grid(rs_col_units<P>(a.ncomp) / CPB, nb);
    const int gb = (a.dg[0] ? 1 : 0) | (a.dg[1] ? 2 : 0);
    if (a.spec_mode == 1)
      hipLaunchKernelGGL((rs_col_kernel<P, 1>), grid, blk, 0, st, a);
    else if (a.spec_mode == 2 && gb == 1)
      hipLaunchKernelGGL((rs_col_kernel<P, 3>), grid, blk, 0, st, a);
    else if (a.spec_mode == 2 && gb == 2)
      hipLaunchKernelGGL((rs_col_kernel<P, 4>), grid, blk, 0, st, a);
    else if (a.spec_mode == 2 && gb == 3)
      hipLaunchKernelGGL((rs_col_kernel<P, 5>), grid, blk, 0, st, a);
    else
      hipLaunchKernelGGL((rs_col_kernel<P, 0>), grid, blk, 0, st, a);
  }
  if (a.spec_mode == 2 && !a.partials_only) {
    KTimer t(K_LOSS, st);
    hipLaunchKernelGGL(rs_grad_reduce_kernel, dim3(2), dim3(1024), 0, st, a.gpart[0], a.gpart[1],
                       rs_col_units<P>(a.ncomp), a.grad_out);
  }
  {
    KTimer t(K_RS_FINAL, st);
    if constexpr (P >= 1024) {
      if (a.intensity)
        hipLaunchKernelGGL((rs_final1_kernel<P, false>), dim3(a.n, nb), dim3(P / 16), 0, st, a);
      else
        hipLaunchKernelGGL((rs_final1_kernel<P, true>), dim3(a.n * a.ncomp, nb), dim3(P / 16), 0, st, a);
    }
    else
      hipLaunchKernelGGL(rs_final_kernel<P>, dim3(rows / 2, nb), dim3(2 * FftThreads<P>::value), 0, st, a);
  }
  return hipGetLastError();
}

}  // namespace xl

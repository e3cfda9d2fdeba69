// LDS-staged Stockham FFT building blocks for gfx950 (64-wide waves).
//
// A workgroup runs FPW independent complex64 FFTs of a power-of-two length L
// (64 <= L <= 4096), each on TPF = FftThreads<L> threads.  Every pass is radix 16
// except a final radix 2/4/8 when log2(L) is not a multiple of 4.  Data lives
// in LDS between passes (natural order in, natural order out); each thread
// keeps its R points of a butterfly in registers.  LDS indices are padded by
// one slot per 16 complex values so the stride-16 writes of the first pass
// spread over the banks.
//
// Inter-pass twiddles come from a 4096-entry table e^{-2 pi i k/4096}
// generated once per device in float64 (xl_fft_init_kernel).  A butterfly
// loads w and forms the other powers by complex products.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xl {

constexpr int TWN = 4096;                 // twiddle table length == max FFT length
constexpr int TWN3 = 3072;                // table for the lengths 3 * 2^k
constexpr int MAX_FFT = 4096;

__device__ float2 g_tw[TWN];              // e^{-2 pi i k / TWN}, float64-generated
__device__ float2 g_tw3[TWN3];            // e^{-2 pi i k / TWN3}, float64-generated

__global__ void xl_fft_init_kernel() {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < TWN) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)TWN, &s, &c);
    g_tw[k] = make_float2((float)c, (float)s);
  }
  if (k < TWN3) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)TWN3, &s, &c);
    g_tw3[k] = make_float2((float)c, (float)s);
  }
}

// Lengths are 2^k (64..4096) or 3 * 2^k (768..3072); the latter start with one
// radix-3 pass and take their twiddles from the 3072-entry table.
template <int L> struct Tw {
  static constexpr bool TRI = L % 3 == 0;
  static constexpr int N = TRI ? TWN3 : TWN;
  __device__ static float2 at(int i) { return TRI ? g_tw3[i] : g_tw[i]; }
};

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// Pass sequence of an L-point FFT: forward [3,] 16, ..., 16, REM; the inverse
// runs the same radices in reverse order, so the last forward pass and the
// first inverse pass share one register layout.
template <int L> struct FftPlan {
  static constexpr bool TRI = L % 3 == 0;
  static constexpr int L2 = TRI ? L / 3 : L;
  static constexpr int LG = ilog2c(L2);
  static constexpr int N16 = LG / 4;
  static constexpr int REM = 1 << (LG % 4);
  static constexpr int NP = (TRI ? 1 : 0) + N16 + (REM > 1 ? 1 : 0);
  static constexpr int fwd_radix(int i) {
    if (TRI) {
      if (i == 0) return 3;
      i -= 1;
    }
    return i < N16 ? 16 : REM;
  }
  static constexpr int radix(bool fwd, int i) { return fwd ? fwd_radix(i) : fwd_radix(NP - 1 - i); }
  static constexpr int ns(bool fwd, int i) {
    int p = 1;
    for (int k = 0; k < i; ++k) p *= radix(fwd, k);
    return p;
  }
  // register layout at the forward/inverse hand-off
  static constexpr int RH = radix(true, NP - 1);
  static constexpr int NBH = L / RH;
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
template <int L> struct LdsLen { static constexpr int value = L + L / 16; };
// Threads per FFT: one radix-16 butterfly per thread and pass up to L = 2048;
// two at L = 4096, so that four 4096-point FFTs per block run on 512 threads
// with up to 256 VGPRs each instead of spilling at 1024 threads (128 VGPRs).
template <int L> struct FftThreads { static constexpr int value = L >= 4096 ? L / 32 : L / 16; };

__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 operator*(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
template <bool INV> __device__ __forceinline__ float2 twmul(float2 v, float2 w) {
  return INV ? cmulc(v, w) : cmul(v, w);
}

// cos / sin of 2 pi q / 16, q = 0..7
__device__ constexpr float kC16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
                                      0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f};
__device__ constexpr float kS16[8] = {0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f,
                                      1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f};

// v * e^{-+ 2 pi i q / 16} with q a compile-time constant after unrolling.
template <bool INV> __device__ __forceinline__ float2 rot16(float2 v, int q) {
  if (q == 0) return v;
  if (q == 4) return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
  const float c = kC16[q], s = INV ? kS16[q] : -kS16[q];
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// In-register DFT of R points (R = 1, 2, 4, 8, 16), natural order in and out.
template <int R, bool INV> struct DFT {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    DFT<R / 2, INV>::run(e);
    DFT<R / 2, INV>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const float2 t = rot16<INV>(o[k], k * (16 / R));
      v[k] = e[k] + t;
      v[k + R / 2] = e[k] - t;
    }
  }
};
template <bool INV> struct DFT<1, INV> {
  __device__ __forceinline__ static void run(float2*) {}
};
// DFT of R points whose upper half v[R/2..R-1] is zero (a zero-padded
// input): the bottom radix-2 stage of the recursion is a copy.
template <int R, bool INV> struct DFTHalf {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    DFTHalf<R / 2, INV>::run(e);
    DFTHalf<R / 2, INV>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const float2 t = rot16<INV>(o[k], k * (16 / R));
      v[k] = e[k] + t;
      v[k + R / 2] = e[k] - t;
    }
  }
};
template <bool INV> struct DFTHalf<2, INV> {
  __device__ __forceinline__ static void run(float2* v) { v[1] = v[0]; }
};
// DFT of R points of which only the lower half of the outputs is kept
// (a cropped output): the top radix-2 stage forms v[0..R/2-1] only.
template <int R, bool INV> struct DFTLow {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    DFT<R / 2, INV>::run(e);
    DFT<R / 2, INV>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) v[k] = e[k] + rot16<INV>(o[k], k * (16 / R));
  }
};

// X1,2 = a - (b + c)/2 -+ i (sqrt(3)/2)(b - c) (signs swapped for the inverse).
template <bool INV> struct DFT<3, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    const float2 a = v[0], t1 = v[1] + v[2], d = v[1] - v[2];
    const float2 t2 = a - t1 * 0.5f;
    const float h = 0.86602540378443865f;
    const float2 m = make_float2(-d.y * h, d.x * h);   // i h d
    v[0] = a + t1;
    v[1] = INV ? t2 + m : t2 - m;
    v[2] = INV ? t2 - m : t2 + m;
  }
};

// v[r] *= w^r for r < R (w1 = w, e^{-+2 pi i base / TWN}, one table entry);
// w^2, w^4, w^8 by squaring and the other powers as products formed next to
// their use, so that only the four bases stay live.  Loads cost more than the
// products here (each is a scattered L1 access per lane); the squarings add
// at most a few float32 ulps to w^8.
template <int R, bool INV> __device__ __forceinline__ void apply_twiddles_w(float2* v, const float2 w1) {
  if constexpr (R == 3) {
    v[1] = twmul<INV>(v[1], w1);
    v[2] = twmul<INV>(v[2], cmul(w1, w1));
  } else if constexpr (R >= 2) {
    v[1] = twmul<INV>(v[1], w1);
    if constexpr (R >= 4) {
      const float2 w2 = cmul(w1, w1);
      v[2] = twmul<INV>(v[2], w2);
      const float2 w3 = cmul(w1, w2);
      v[3] = twmul<INV>(v[3], w3);
      if constexpr (R >= 8) {
        const float2 w4 = cmul(w2, w2);
        v[4] = twmul<INV>(v[4], w4);
        v[5] = twmul<INV>(v[5], cmul(w1, w4));
        v[6] = twmul<INV>(v[6], cmul(w2, w4));
        const float2 w7 = cmul(w3, w4);
        v[7] = twmul<INV>(v[7], w7);
        if constexpr (R >= 16) {
          const float2 w8 = cmul(w4, w4);
          v[8] = twmul<INV>(v[8], w8);
          v[9] = twmul<INV>(v[9], cmul(w1, w8));
          v[10] = twmul<INV>(v[10], cmul(w2, w8));
          v[11] = twmul<INV>(v[11], cmul(w3, w8));
          v[12] = twmul<INV>(v[12], cmul(w4, w8));
          v[13] = twmul<INV>(v[13], cmul(cmul(w1, w4), w8));
          v[14] = twmul<INV>(v[14], cmul(cmul(w2, w4), w8));
          v[15] = twmul<INV>(v[15], cmul(w7, w8));
        }
      }
    }
  }
}

template <int R, bool INV, int L> __device__ __forceinline__ void apply_twiddles(float2* v, int base) {
  apply_twiddles_w<R, INV>(v, Tw<L>::at(base));
}

// One Stockham radix-R pass of an L-point FFT over TPF threads (thread t).
// GUARD: the block has more than TPF threads; the others only join the
// barriers.
template <int L, int R, int NS, int TPF, bool INV, bool GUARD = false>
__device__ __forceinline__ void stockham_pass(float2* s, int t) {
  constexpr int NB = L / R;
  constexpr int BPT = (NB + TPF - 1) / TPF;
  float2 v[BPT][R];
  const bool act = !GUARD || t < TPF;
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPF;
    if (act && (NB % TPF == 0 || j < NB)) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = s[lpad(j + r * NB)];
      if constexpr (NS > 1) apply_twiddles<R, INV, L>(v[b], (j % NS) * (Tw<L>::N / (NS * R)));
      DFT<R, INV>::run(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPF;
    if (act && (NB % TPF == 0 || j < NB)) {
      const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) s[lpad(idxD + r * NS)] = v[b][r];
    }
  }
  __syncthreads();
}

template <int L, int TPF, bool INV, int I = 0, bool GUARD = false>
__device__ __forceinline__ void fft_lds(float2* s, int t) {
  using PL = FftPlan<L>;
  if constexpr (I < PL::NP) {
    stockham_pass<L, PL::radix(true, I), PL::ns(true, I), TPF, INV, GUARD>(s, t);
    fft_lds<L, TPF, INV, I + 1, GUARD>(s, t);
  }
}

// Block-index remap so that consecutive logical blocks (which share 128-byte
// lines of the tiled intermediates) land on the same XCD: hardware dispatches
// block b to XCD b % 8.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int NX = 8;
  if (nblocks % NX) return b;
  return (b % NX) * (nblocks / NX) + b / NX;
}

// --- float64 phase helpers ---------------------------------------------------
// e^{i 2 pi t} for a float64 number of turns t: the integer part is removed in
// float64, the remainder in [-1/2, 1/2] goes to float32 sincospi.
__device__ __forceinline__ float2 expi_turns(double t) {
  const double f = t - rint(t);
  float s, c;
  sincospif((float)(2.0 * f), &s, &c);
  return make_float2(c, s);
}

}  // namespace xl

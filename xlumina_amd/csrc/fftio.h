// Stockham FFT passes with direct I/O at the ends of the pass chain.
//
// Each pass is the radix-R Stockham step of fft.h, but its source can be LDS,
// a caller functor (global loads, masks, on-the-fly generation) or registers,
// and its sink LDS, a caller functor (global stores of only the samples that
// are needed) or registers.  The forward chain uses radices [3,] 16,...,16,REM
// and the inverse chain the same in reverse order, so the last forward pass leaves element
// j + q*(L/R) in register v[b][q] (j = t + b*TPF) and the first inverse pass
// reads exactly that layout: a frequency-domain product (transfer function,
// chirp spectrum) is applied in registers between the two chains, with no
// LDS round trip.
//
// Functor interfaces:
//   load(int idx) -> float2        input sample idx in [0, L) (0 where padded)
//   store(int idx, float2 v)       output sample idx (the functor skips unneeded ones)
//   mul(int idx, int slot, float2 v) -> float2   frequency-domain product
#pragma once
#include "fft.h"

namespace xl {

template <int BPT, int R> struct Regs {
  float2 v[BPT][R];
};

enum { SRC_LDS = 0, SRC_FN = 1, SRC_REG = 2 };
enum { DST_LDS = 0, DST_FN = 1, DST_REG = 2 };

// One pass.  SRC_REG / DST_REG use `reg` (shape [BPT][R] of this pass).
// ZHI: the inputs r >= R/2 are zero (a first pass over a zero-padded upper
// half: indices >= L/2); they are neither loaded nor transformed.  LOW: only
// the outputs r < R/2 are formed and written (a last pass whose indices
// >= L/2 are cropped away).
// wb: the pass's twiddle bases, one per butterfly slot b, when PREW (loaded
// ahead by the caller); otherwise the pass loads them itself (NS > 1).
template <int L, int R, int NS, int TPF, bool INV, int SRC, int DST, bool ZHI = false, bool LOW = false,
          bool PREW = true, class Load, class Store, class RegT>
__device__ __forceinline__ void pass_io(float2* s, int t, Load& load, Store& store, RegT& reg, const float2* wb) {
  constexpr int NB = L / R;
  constexpr int BPT = (NB + TPF - 1) / TPF;
  static_assert(!(ZHI && LOW), "a pass is either the first or the last of a chain here");
  static_assert(!ZHI || NS == 1, "zero upper inputs only on the first pass");
  float2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPF;
    if (NB % TPF == 0 || j < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (ZHI && r >= R / 2) v[b][r] = make_float2(0.f, 0.f);
        else if constexpr (SRC == SRC_LDS) v[b][r] = s[lpad(j + r * NB)];
        else if constexpr (SRC == SRC_FN) v[b][r] = load(j + r * NB);
        else v[b][r] = reg.v[b][r];
      }
      if constexpr (NS > 1) {
        if constexpr (PREW) apply_twiddles_w<R, INV>(v[b], wb[b]);
        else apply_twiddles<R, INV, L>(v[b], (j % NS) * (Tw<L>::N / (NS * R)));
      }
      if constexpr (ZHI) DFTHalf<R, INV>::run(v[b]);
      else if constexpr (LOW) DFTLow<R, INV>::run(v[b]);
      else DFT<R, INV>::run(v[b]);
    }
  }
  // LDS may still be read by other threads (this pass or the previous phase)
  if constexpr (DST == DST_LDS) __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPF;
    if (NB % TPF == 0 || j < NB) {
      const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < (LOW ? R / 2 : R); ++r) {
        if constexpr (DST == DST_LDS) s[lpad(idxD + r * NS)] = v[b][r];
        else if constexpr (DST == DST_FN) store(idxD + r * NS, v[b][r]);
        else reg.v[b][r] = v[b][r];
      }
    }
  }
  if constexpr (DST == DST_LDS) __syncthreads();
}

struct NoIO {
  __device__ float2 operator()(int) const { return make_float2(0.f, 0.f); }
  __device__ void operator()(int, float2) const {}
};

// Twiddle bases of every pass of a chain (the table entry w of pass I for
// butterfly slot b of thread t), loaded together when the chain starts: a
// pass's twiddles then need no memory access behind its barrier.
template <int L, int TPF, bool FWD> struct TwSet {
  using P = FftPlan<L>;
  static constexpr int bpt(int I) { return (L / P::radix(FWD, I) + TPF - 1) / TPF; }
  static constexpr int off(int I) {
    int o = 0;
    for (int k = 1; k < I; ++k) o += bpt(k);
    return o;
  }
  static constexpr int N = off(P::NP) > 0 ? off(P::NP) : 1;
  float2 w[N];
  template <int I = 1> __device__ __forceinline__ void load(int t) {
    if constexpr (I < P::NP) {
      constexpr int R = P::radix(FWD, I), NS = P::ns(FWD, I), NB = L / R;
#pragma unroll
      for (int b = 0; b < bpt(I); ++b) {
        const int j = t + b * TPF;
        w[off(I) + b] = (NB % TPF == 0 || j < NB) ? Tw<L>::at((j % NS) * (Tw<L>::N / (NS * R))) : make_float2(1.f, 0.f);
      }
      load<I + 1>(t);
    }
  }
  template <int I> __device__ __forceinline__ const float2* at() const { return w + (I >= 1 ? off(I) : 0); }
};

// Pass chain i..NP-1 of the forward (FWD = true) or inverse plan (FftPlan).  The first pass reads SRC, the last writes DST, the
// others go through LDS.  Every pass that writes LDS first waits for all
// threads (so a chain may follow any earlier LDS use); a chain ending in
// DST_FN / DST_REG leaves LDS free without a final barrier.
// ZHI / LOW (see pass_io) apply to the first / last pass of the chain; both
// need a plan that starts (ZHI) or ends (LOW) with an unstrided radix-16 pass
// over the L/2 split, i.e. a power-of-two L.
template <int L, int TPF, bool INV, bool FWD, int I, int SRC, int DST, bool ZHI, bool LOW, bool TWPRE, class Load,
          class Store, class RegIn, class RegOut>
__device__ __forceinline__ void chain_pass(float2* s, int t, Load& load, Store& store, RegIn& rin, RegOut& rout,
                                           TwSet<L, TPF, FWD>& tw) {
  using P = FftPlan<L>;
  constexpr int R = P::radix(FWD, I), NS = P::ns(FWD, I);
  constexpr bool first = I == 0, last = I == P::NP - 1;
  constexpr int S = first ? SRC : SRC_LDS;
  constexpr int D = last ? DST : DST_LDS;
  static_assert(!(ZHI || LOW) || (!P::TRI && P::NP >= 2), "zero-half passes need a power-of-two plan");
  static_assert(!(ZHI && first) || R == 16, "ZHI needs a radix-16 first pass");
  static_assert(!(LOW && last) || (R == 16 && NS * R == L), "LOW needs an unstrided radix-16 last pass");
  // TWPRE: the bases were loaded when the chain started; otherwise each pass
  // loads its own
  const float2* wb = tw.template at<I>();
  if constexpr (first) {
    pass_io<L, R, NS, TPF, INV, S, D, ZHI, false, TWPRE>(s, t, load, store, rin, wb);
  } else if constexpr (last) {
    pass_io<L, R, NS, TPF, INV, S, D, false, LOW, TWPRE>(s, t, load, store, rout, wb);
  } else {
    pass_io<L, R, NS, TPF, INV, S, D, false, false, TWPRE>(s, t, load, store, rout, wb);
  }
  if constexpr (!last)
    chain_pass<L, TPF, INV, FWD, I + 1, SRC, DST, ZHI, LOW, TWPRE>(s, t, load, store, rin, rout, tw);
}

// The whole chain (I = 0).  TWPRE: the twiddle bases of all passes are loaded
// first (a few registers per pass: the kernels that cannot spare them keep
// loading each pass's bases at the pass).
template <int L, int TPF, bool INV, bool FWD, int I, int SRC, int DST, bool ZHI = false, bool LOW = false,
          bool TWPRE = true, class Load, class Store, class RegIn, class RegOut>
__device__ __forceinline__ void chain_io(float2* s, int t, Load& load, Store& store, RegIn& rin, RegOut& rout) {
  static_assert(I == 0, "a chain starts at its first pass");
  TwSet<L, TPF, FWD> tw;
  if constexpr (TWPRE) tw.load(t);
  chain_pass<L, TPF, INV, FWD, 0, SRC, DST, ZHI, LOW, TWPRE>(s, t, load, store, rin, rout, tw);
}

template <int L, int TPF> struct HandOff {
  using P = FftPlan<L>;
  static constexpr int R = P::RH;
  static constexpr int NB = P::NBH;
  static constexpr int BPT = (NB + TPF - 1) / TPF;
  using Type = Regs<BPT, R>;
  // element index held in slot (b, q)
  __device__ static int index(int t, int b, int q) { return t + b * TPF + q * NB; }
  __device__ static bool valid(int t, int b) { return NB % TPF == 0 || t + b * TPF < NB; }
};

// forward FFT: SRC -> registers (hand-off layout)
template <int L, int TPF, int SRC, bool ZHI = false, bool TWPRE = true, class Load>
__device__ __forceinline__ void fft_forward_to_regs(float2* s, int t, Load& load,
                                                    typename HandOff<L, TPF>::Type& out) {
  NoIO none;
  typename HandOff<L, TPF>::Type dummy;
  chain_io<L, TPF, false, true, 0, SRC, DST_REG, ZHI, false, TWPRE>(s, t, load, none, dummy, out);
}

// inverse FFT: registers (hand-off layout) -> DST
template <int L, int TPF, int DST, bool LOW = false, bool TWPRE = true, class Store>
__device__ __forceinline__ void fft_inverse_from_regs(float2* s, int t, typename HandOff<L, TPF>::Type& in,
                                                      Store& store) {
  NoIO none;
  chain_io<L, TPF, true, false, 0, SRC_REG, DST, false, LOW, TWPRE>(s, t, none, store, in, in);
}

// Output layout of the last pass of a plan: slot (b, q) holds index j + q*NS,
// j = t + b*TPF < NB (NS = NB = L / R_last).
template <int L, int TPF, bool FWD> struct LastPass {
  using P = FftPlan<L>;
  static constexpr int R = P::radix(FWD, P::NP - 1);
  static constexpr int NB = L / R;
  static constexpr int BPT = (NB + TPF - 1) / TPF;
  using Type = Regs<BPT, R>;
  __device__ static int index(int t, int b, int q) { return t + b * TPF + q * NB; }
  __device__ static bool valid(int t, int b) { return NB % TPF == 0 || t + b * TPF < NB; }
};

// full transform, SRC -> DST (no hand-off)
template <int L, int TPF, bool INV, int SRC, int DST, class Load, class Store>
__device__ __forceinline__ void fft_io(float2* s, int t, Load& load, Store& store,
                                       typename LastPass<L, TPF, !INV>::Type& out) {
  typename LastPass<L, TPF, !INV>::Type dummy;
  chain_io<L, TPF, INV, !INV, 0, SRC, DST>(s, t, load, store, dummy, out);
}

}  // namespace xl

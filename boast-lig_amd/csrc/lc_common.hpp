// lc_common.hpp -- device-side building blocks of the libconv HIP library.
//
// Filters are compile-time tap tables (lc_taps.h) so every tap becomes an
// immediate operand once the tap loops are unrolled.  Operator families:
//   Conv<F>   one value per step, y_i = sum_l f_l x_{i+l}
//             (ConvolutionFilter, BigDFT/GenericConvolution2.rb:124-234)
//   Dwt<W>    wavelet analysis, input step = sample pair (s_2j, s_2j+1),
//             output step = (L_i, H_i) (WaveletFilterDecompose :505-587)
//   Idwt<W>   wavelet synthesis, input step = (L_j, H_j), output step =
//             (s_2i, s_2i+1) (WaveletFilterRecompose :589-671)
// Boundary conditions: 0 periodic, 1 free grow, -1 free shrink
// (BoundaryConditions :684-740).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lc_taps.h"

namespace lc {

// FREE: BoundaryConditions::FREE ("discard", GenericConvolution2.rb:684-740):
// input [0, n) zero outside, output [0, n), for the 3D convolution kernels
// only (libconv.f90:245 'F' nogrow, the Kinetic step of the local Hamiltonian)
enum Bc : int { PER = 0, GROW = 1, SHRINK = -1, FREE = 2 };

#define LC_AS_D(v) v,
#define LC_AS_F(v) v##f,

// ---------------------------------------------------------------- filters
#define LC_DEFINE_FILTER(NAME, TAG)                                        \
  struct NAME {                                                            \
    static constexpr int LO = LC_##TAG##_LO;                               \
    static constexpr int N = LC_##TAG##_N;                                 \
    static constexpr int UP = LO + N - 1;                                  \
    static constexpr double d[N] = {LC_TAPS_##TAG(LC_AS_D)};               \
    static constexpr float f[N] = {LC_TAPS_##TAG(LC_AS_F)};                \
  };
LC_DEFINE_FILTER(FMD, MD)
LC_DEFINE_FILTER(FIMD, IMD)
LC_DEFINE_FILTER(FD2, D2)
LC_DEFINE_FILTER(FD1, D1)
LC_DEFINE_FILTER(FLP, LP)
LC_DEFINE_FILTER(FHP, HP)
LC_DEFINE_FILTER(FCLPP, CLPP)
LC_DEFINE_FILTER(FCHPP, CHPP)
LC_DEFINE_FILTER(FCLPG, CLPG)
LC_DEFINE_FILTER(FCHPG, CHPG)
#undef LC_DEFINE_FILTER

// fused multiply-add in the storage precision (the float overload keeps
// FP32 data in FP32 FMAs; energies call the double one explicitly)
__host__ __device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R, class F>
__host__ __device__ __forceinline__ constexpr R tap(int k) {
  if constexpr (sizeof(R) == 8) return F::d[k]; else return F::f[k];
}

// Wavelet split filters (WaveletFilter#split_filters,
// GenericConvolution2.rb:386-410).  A wavelet (low, high, centre) is split
// into even/odd halves of LEN/2 taps at pair offsets [LO, UP]; synthesis
// uses the reversed filters with centre LEN - centre - 1.
template <class Low, class High, int CENTER, bool RECOMPOSE>
struct WSplit {
  static constexpr int LEN = Low::N;
  static constexpr int M = LEN / 2;
  static constexpr int CNTR = RECOMPOSE ? (LEN - CENTER - 1) : CENTER;
  static constexpr int LO = -(CNTR / 2);
  static constexpr int UP = M + LO - 1;
  static constexpr int W = M;
  // k-th tap (k = l - LO) of the four split filters
  template <typename R> __host__ __device__ static constexpr R le(int k) {
    return RECOMPOSE ? tap<R, Low>(LEN - 1 - 2 * k) : tap<R, Low>(2 * k);
  }
  template <typename R> __host__ __device__ static constexpr R lo(int k) {
    return RECOMPOSE ? tap<R, Low>(LEN - 2 - 2 * k) : tap<R, Low>(2 * k + 1);
  }
  template <typename R> __host__ __device__ static constexpr R he(int k) {
    return RECOMPOSE ? tap<R, High>(LEN - 1 - 2 * k) : tap<R, High>(2 * k);
  }
  template <typename R> __host__ __device__ static constexpr R ho(int k) {
    return RECOMPOSE ? tap<R, High>(LEN - 2 - 2 * k) : tap<R, High>(2 * k + 1);
  }
};

// ---------------------------------------------------------------- operators
// Common interface:
//   LO, UP, W        step offsets of the input window, W = UP - LO + 1
//   IN, OUT          values per input / output step
//   compute(win, o)  win[s][c], s = 0..W-1 (step i+LO+s), c < IN
//   DOT              the operator supports the dot_in / a_x terms
template <class F>
struct Conv {
  static constexpr int LO = F::LO, UP = F::UP, W = F::N, IN = 1, OUT = 1;
  static constexpr bool DOT = true;
  template <typename R>
  __device__ __forceinline__ static void compute(const R (*win)[1], R* out) {
    R t = R(0);
#pragma unroll
    for (int s = 0; s < W; ++s) t = fmar(win[s][0], tap<R, F>(s), t);
    out[0] = t;
  }
};

template <class S>
struct Dwt {  // input (s_2j, s_2j+1), output (L_i, H_i)
  static constexpr int LO = S::LO, UP = S::UP, W = S::W, IN = 2, OUT = 2;
  static constexpr bool DOT = false;
  template <typename R>
  __device__ __forceinline__ static void compute(const R (*win)[2], R* out) {
    R lt = R(0), ht = R(0);
#pragma unroll
    for (int s = 0; s < W; ++s) {
      lt = fmar(win[s][0], S::template le<R>(s), lt);
      ht = fmar(win[s][0], S::template he<R>(s), ht);
      lt = fmar(win[s][1], S::template lo<R>(s), lt);
      ht = fmar(win[s][1], S::template ho<R>(s), ht);
    }
    out[0] = lt;
    out[1] = ht;
  }
};

template <class S>
struct Idwt {  // input (L_j, H_j), output (s_2i, s_2i+1)
  static constexpr int LO = S::LO, UP = S::UP, W = S::W, IN = 2, OUT = 2;
  static constexpr bool DOT = false;
  template <typename R>
  __device__ __forceinline__ static void compute(const R (*win)[2], R* out) {
    R et = R(0), ot = R(0);
#pragma unroll
    for (int s = 0; s < W; ++s) {
      et = fmar(win[s][0], S::template lo<R>(s), et);
      ot = fmar(win[s][0], S::template le<R>(s), ot);
      et = fmar(win[s][1], S::template ho<R>(s), et);
      ot = fmar(win[s][1], S::template he<R>(s), ot);
    }
    out[0] = et;
    out[1] = ot;
  }
};

using Sym8Dec = WSplit<FLP, FHP, 8, false>;
using Sym8Rec = WSplit<FLP, FHP, 8, true>;
using CombPDec = WSplit<FCLPP, FCHPP, -LC_CLPP_LO, false>;
using CombPRec = WSplit<FCLPP, FCHPP, -LC_CLPP_LO, true>;
using CombGDec = WSplit<FCLPG, FCHPG, -LC_CLPG_LO, false>;
using CombGRec = WSplit<FCLPG, FCHPG, -LC_CLPG_LO, true>;

// ---------------------------------------------------------------- layout
// Memory layout of one processed line (stride L between steps):
//   Conv input  value j          at (j - ox)
//   Dwt  input  sample c of j    at c + 2 (j - ox)
//   Idwt input  band b of j      at (j - ox) + nx b
//   Conv output i                at (i - oy)
//   Dwt  output band b of i      at (i - oy) + ny b
//   Idwt output sample c of i    at c + 2 (i - oy)
// (dims of x/y in ConvolutionFilter / WaveletFilter{Decompose,Recompose}
//  get_{input,output}_dim*, GenericConvolution2.rb:205-225,:533-585,:617-669)
enum Kind : int { KCONV = 0, KDWT = 1, KIDWT = 2 };

template <class Op> struct KindOf;
template <class F> struct KindOf<Conv<F>> { static constexpr int v = KCONV; };
template <class S> struct KindOf<Dwt<S>> { static constexpr int v = KDWT; };
template <class S> struct KindOf<Idwt<S>> { static constexpr int v = KIDWT; };

// element offset (in units of the line stride) of component c of input step j
template <int K>
__host__ __device__ __forceinline__ long in_elem(long jrel, int c, int nx) {
  if (K == KCONV) return jrel;
  if (K == KDWT) return 2 * jrel + c;
  return jrel + (long)nx * c;
}
template <int K>
__host__ __device__ __forceinline__ long out_elem(long irel, int c, int ny) {
  if (K == KCONV) return irel;
  if (K == KDWT) return irel + (long)ny * c;
  return 2 * irel + c;
}

__host__ __device__ __forceinline__ int pmod(int a, int b) {
  int r = a % b;
  return r < 0 ? r + b : r;
}

// Epilogue flags (the dispatcher's variant selection,
// GenericConvolution2.rb:1650-1699 and :173-203)
// F_POT: after the epilogue, rho += t^2 and t = pot t (the potential step of
// the local Hamiltonian fused into the magic filter's store)
enum Flags : int { F_A = 1, F_AX = 2, F_AY = 4, F_DOT = 8, F_POT = 16 };

}  // namespace lc

// lc_fused_wav.hip -- single-pass 3D wavelet transforms on gfx950.
//
// One level of the separable 3D DWT (analysis) or IDWT (synthesis): along
// every dim the transform maps input steps j (two values each: a sample
// pair for the DWT, the L and H bands for the IDWT) to output steps i (two
// values each) with a W-step window,
//   out[i][c'] = sum_{s<W} sum_{c<2} C[c'][c](s) * in[i + LO + s][c],
// the split filters of WaveletFilterDecompose / WaveletFilterRecompose
// (BigDFT/GenericConvolution2.rb:386-410, :453-462, :505-671).  The three
// per-dim sweeps of the libconv composition (:2160-2240) become one pass:
// per input z-step the two z-planes of the (TX+W-1) x (TY+W-1)-step tile go
// to LDS, x-tasks filter rows into LDS, every lane filters one column along
// y for both bands and feeds its register window along z, which emits both
// z bands of the output step.  Data layout per dim (storage extents 2 nx and
// 2 ny values):
//   DWT   input  sample c of step j at c + 2 (j - ox),  output band b of
//         step i at (i - oy) + ny b
//   IDWT  input  band b of step j at (j - ox) + nx b,   output sample c of
//         step i at c + 2 (i - oy)
// Periodic dims wrap, free dims read zero outside the stored input, exactly
// as the 1D sweeps.  Epilogue: y = a * t + a_y * y.
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "lc_fused.hpp"

namespace lc {

// C[c'][c](s): weight of input value c of window step s in output value c'
template <class Op>
struct WavCoef;
template <class S>
struct WavCoef<Dwt<S>> {
  static constexpr int K = KDWT;
  template <typename R>
  __device__ __forceinline__ static constexpr R c(int co, int ci, int s) {
    return co == 0 ? (ci == 0 ? S::template le<R>(s) : S::template lo<R>(s))
                   : (ci == 0 ? S::template he<R>(s) : S::template ho<R>(s));
  }
};
template <class S>
struct WavCoef<Idwt<S>> {
  static constexpr int K = KIDWT;
  template <typename R>
  __device__ __forceinline__ static constexpr R c(int co, int ci, int s) {
    return co == 0 ? (ci == 0 ? S::template lo<R>(s) : S::template ho<R>(s))
                   : (ci == 0 ? S::template le<R>(s) : S::template he<R>(s));
  }
};

// both output values of one step from a window v(s, c), s-major order (the
// order of Dwt::compute / Idwt::compute)
template <typename R, class Op, class V>
__device__ __forceinline__ void wav_step(const V& v, R& o0, R& o1) {
  using C = WavCoef<Op>;
  R t0 = R(0), t1 = R(0);
#pragma unroll
  for (int s = 0; s < Op::W; ++s) {
    t0 = __builtin_fma(v(s, 0), C::template c<R>(0, 0, s), t0);
    t1 = __builtin_fma(v(s, 0), C::template c<R>(1, 0, s), t1);
    t0 = __builtin_fma(v(s, 1), C::template c<R>(0, 1, s), t0);
    t1 = __builtin_fma(v(s, 1), C::template c<R>(1, 1, s), t1);
  }
  o0 = t0;
  o1 = t1;
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
__device__ __forceinline__ void sepw3d_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr) {
  constexpr int K = WavCoef<Op>::K;
  constexpr int W = Op::W, LO = Op::LO, UP = Op::UP;
  constexpr int NT = 2 * TX * TY;
  constexpr int TWS = TX + W - 1;          // input steps per tile row
  constexpr int AROW = (2 * TWS) | 1;      // LDS row stride (odd)
  constexpr int TH = TY + W - 1;           // input steps per tile column
  constexpr int NR = 2 * TH;               // tile rows (both values of a step)
  constexpr int BW = (2 * TX) | 1;         // x-filtered row stride (odd)
  constexpr int NA = NR * 2 * TWS;         // tile values per z-plane
  constexpr int NAT = (NA + NT - 1) / NT;
  constexpr int NSEG = TX / QX;
  constexpr int NXT = 2 * NR * NSEG;       // x-tasks of both z-planes
  constexpr int WS = W - 1 + BZ;
  static_assert(TX % QX == 0 && BZ % AD == 0, "tile shape");
  __shared__ R A[2][NR * AROW];
  __shared__ R B[2][NR * BW];

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long rs_in = 2L * p.d0.nx, ps_in = 4L * p.d0.nx * p.d1.nx;
  const long rs_out = 2L * p.d0.ny, ps_out = 4L * p.d0.ny * p.d1.ny;

  // ---- this lane's output column: value c0 of x-step i0l, y-step i1l
  const int col = t % (2 * TX), i1l = t / (2 * TX);
  const int c0 = K == KIDWT ? (col & 1) : col / TX;
  const int i0l = K == KIDWT ? (col >> 1) : col % TX;
  const int bcol = c0 * TX + i0l;
  const bool out_ok = (X0 + i0l < p.d0.olo + p.d0.nout) && (Y0 + i1l < p.d1.olo + p.d1.nout);
  const long yoff_xy = out_elem<K>(X0 + i0l - p.d0.oy, c0, p.d0.ny) +
                       rs_out * out_elem<K>(Y0 + i1l - p.d1.oy, 0, p.d1.ny);
  const long ystep1 = rs_out * (out_elem<K>(0, 1, p.d1.ny) - out_elem<K>(0, 0, p.d1.ny));

  // ---- tile values of this lane (same xy positions in both z-planes)
  int aoff[NAT], aslot[NAT];
#pragma unroll
  for (int k = 0; k < NAT; ++k) {
    const int e = t + k * NT;
    const int rr = e / (2 * TWS), cc = e % (2 * TWS);
    // memory order along x: DWT interleaves the pair, IDWT stores bands
    const int s0 = K == KDWT ? (cc >> 1) : cc % TWS;
    const int v0 = K == KDWT ? (cc & 1) : cc / TWS;
    const int v1 = rr / TH, s1 = rr % TH;
    const int j0 = fsrc(p.d0, X0 + LO + s0), j1 = fsrc(p.d1, Y0 + LO + s1);
    const bool ok = e < NA && j0 >= 0 && j1 >= 0;
    aoff[k] = ok ? (int)(in_elem<K>(j0, v0, p.d0.nx) + rs_in * in_elem<K>(j1, v1, p.d1.nx)) : -1;
    aslot[k] = e < NA ? rr * AROW + v0 * TWS + s0 : -1;
  }
  auto load_step = [&](int j2, R (*dst)[NAT]) {
    const int s2 = fsrc(p.d2, j2);
    const R* b0 = xa + (s2 >= 0 ? ps_in * in_elem<K>(s2, 0, p.d2.nx) : 0);
    const R* b1 = xa + (s2 >= 0 ? ps_in * in_elem<K>(s2, 1, p.d2.nx) : 0);
#pragma unroll
    for (int k = 0; k < NAT; ++k) {
      dst[0][k] = (s2 >= 0 && aoff[k] >= 0) ? b0[aoff[k]] : R(0);
      dst[1][k] = (s2 >= 0 && aoff[k] >= 0) ? b1[aoff[k]] : R(0);
    }
  };

  R areg[AD][2][NAT];
  const int plast = ze - 1 + UP;
  const int pfirst = zb + LO;
#pragma unroll
  for (int h = 0; h < AD; ++h)
    if (pfirst + h <= plast) load_step(pfirst + h, areg[h]);

  // x and y passes of input z-step j2 (tile in ar) -> pv[c2][c1']
  auto filter_step = [&](int j2, R (*ar)[NAT], R (*pv)[2]) {
#pragma unroll
    for (int k = 0; k < NAT; ++k)
      if (aslot[k] >= 0) {
        A[0][aslot[k]] = ar[0][k];
        A[1][aslot[k]] = ar[1][k];
      }
    if (j2 + AD <= plast) load_step(j2 + AD, ar);
    __syncthreads();
    for (int task = t; task < NXT; task += NT) {
      // consecutive lanes on consecutive rows (odd stride: no bank conflicts)
      const int c2 = task / (NR * NSEG), rem = task % (NR * NSEG);
      const int rr = rem % NR, seg = rem / NR;
      const R* row = &A[c2][rr * AROW + seg * QX];
      R* brow = &B[c2][rr * BW + seg * QX];
#pragma unroll
      for (int q = 0; q < QX; ++q) {
        R o0, o1;
        wav_step<R, Op>([&](int s, int c) { return row[c * TWS + q + s]; }, o0, o1);
        brow[q] = o0;
        brow[TX + q] = o1;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2)
      wav_step<R, Op>([&](int s, int c) { return B[c2][(c * TH + i1l + s) * BW + bcol]; },
                      pv[c2][0], pv[c2][1]);
  };

  R win[2][WS][2];  // [c1'][window step][c2]
  for (int q0 = 0; q0 < W - 1; q0 += AD) {
#pragma unroll
    for (int j = 0; j < AD; ++j) {
      const int q = q0 + j;
      if (q < W - 1) {
        R pv[2][2];
        filter_step(pfirst + q, areg[j], pv);
#pragma unroll
        for (int c1 = 0; c1 < 2; ++c1) {
#pragma unroll
          for (int s = 0; s < W - 2; ++s) {
            win[c1][s][0] = win[c1][s + 1][0];
            win[c1][s][1] = win[c1][s + 1][1];
          }
          win[c1][W - 2][0] = pv[0][c1];
          win[c1][W - 2][1] = pv[1][c1];
        }
      }
    }
  }
  for (int z0 = zb; z0 < ze; z0 += BZ) {
#pragma unroll
    for (int u = 0; u < BZ; ++u) {
      const int z = z0 + u;
      if (z < ze) {
        R pv[2][2];
        filter_step(z + UP, areg[((W - 1) % AD + u) % AD], pv);
        const long zo0 = ps_out * out_elem<K>(z - p.d2.oy, 0, p.d2.ny);
        const long zo1 = ps_out * out_elem<K>(z - p.d2.oy, 1, p.d2.ny);
#pragma unroll
        for (int c1 = 0; c1 < 2; ++c1) {
          win[c1][W - 1 + u][0] = pv[0][c1];
          win[c1][W - 1 + u][1] = pv[1][c1];
          R o0, o1;
          wav_step<R, Op>([&](int s, int c) { return win[c1][u + s][c]; }, o0, o1);
          if (p.flags & F_A) {
            o0 = o0 * p.a0;
            o1 = o1 * p.a0;
          }
          if (out_ok) {
            R* y0 = ya + yoff_xy + c1 * ystep1 + zo0;
            R* y1 = ya + yoff_xy + c1 * ystep1 + zo1;
            if (p.flags & F_AY) {
              o0 = __builtin_fma(p.a_y, *y0, o0);
              o1 = __builtin_fma(p.a_y, *y1, o1);
            }
            *y0 = o0;
            *y1 = o1;
          }
        }
      }
    }
#pragma unroll
    for (int c1 = 0; c1 < 2; ++c1)
#pragma unroll
      for (int s = 0; s < W - 1; ++s) {
        win[c1][s][0] = win[c1][s + BZ][0];
        win[c1][s][1] = win[c1][s + BZ][1];
      }
  }
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
__global__ __launch_bounds__(2 * TX * TY) void sepw3d_fused(FusedArgs<R> p) {
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sepw3d_body<R, Op, TX, TY, QX, BZ, AD>(p, bx, by, zc, arr);
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
__global__ __launch_bounds__(2 * TX * TY) void sepw3d_batch(const FusedArgs<R>* __restrict__ tab,
                                                             const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  sepw3d_body<R, Op, TX, TY, QX, BZ, AD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0);
}

// ====================================================================== host
// Compiled variants: tile TX x TY steps, QX outputs per x-task, BZ steps per
// window slide, AD prefetch depth (2 TX TY lanes per block).
#define LC_WAV_VARIANTS(X) \
  X(0, 16, 8, 4, 2, 2) X(1, 8, 8, 4, 2, 2) X(2, 32, 4, 4, 2, 2) X(3, 16, 8, 4, 1, 1)

int wav_variant_count() { return 4; }

std::string wav_kernel_name(int prec, int form, int id, int v, bool batch) {
  (void)id;
  const char* r = prec == 4 ? "float" : "double";
  const char* op = form == FORM_IDWT ? "Idwt<lc::WSplit<lc::FLP, lc::FHP, 8, true> >"
                                     : "Dwt<lc::WSplit<lc::FLP, lc::FHP, 8, false> >";
  char buf[200];
  buf[0] = 0;
  switch (v) {
#define X(I, TX, TY, QX, BZ, AD)                                                                   \
  case I:                                                                                         \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d>", batch ? "sepw3d_batch" : "sepw3d_fused", \
             r, op, TX, TY, QX, BZ, AD);                                                          \
    break;
    LC_WAV_VARIANTS(X)
#undef X
  }
  return buf;
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
static FusedArgs<R> wav_args(const Box3& b, int narr, const R* x, R* y, R a, R a_y) {
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], Op::LO, Op::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], Op::LO, Op::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], Op::LO, Op::UP);
  p.nti = 8L * b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = 8L * b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  p.a0 = a; p.a_y = a_y;
  p.flags = (a != R(1) ? F_A : 0) | (a_y != R(0) ? F_AY : 0);
  return p;
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
static hipError_t launch_sepw3d(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st) {
  FusedArgs<R> p = wav_args<R, Op, TX, TY, QX, BZ, AD>(b, narr, x, y, a, a_y);
  const long tiles = (long)p.tiles_x * p.tiles_y;
  // z chunks no shorter than 2 W steps
  const long want = 2L * num_cu();
  long chunks = (want + tiles * std::max(narr, 1) - 1) / (tiles * std::max(narr, 1));
  chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / (2 * Op::W))));
  p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sepw3d_fused<R, Op, TX, TY, QX, BZ, AD><<<grid, 2 * TX * TY, 0, st>>>(p);
  return hipGetLastError();
}

template <typename R, class Op>
static hipError_t sepw3d_variant(int v, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                                 hipStream_t st) {
  switch (v) {
#define X(I, TX, TY, QX, BZ, AD) \
  case I: return launch_sepw3d<R, Op, TX, TY, QX, BZ, AD>(b, narr, x, y, a, a_y, st);
    LC_WAV_VARIANTS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

// The symicomb8 family is not instantiated: its single-pass kernel (420
// VGPRs, spilling to AGPRs) returned wrong high bands on the GPU and then
// faulted, while a CPU emulation of its exact address arithmetic stays in
// bounds and matches the oracle.  symicomb8 3D transforms run as three 1D
// sweeps instead (lc_api.cpp).
template <typename R>
hipError_t sepw3d_v(int v, int kind, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                    hipStream_t st) {
  if (id != WAV_SYM8) return hipErrorNotSupported;
  if (kind == KDWT) return sepw3d_variant<R, Dwt<Sym8Dec>>(v, b, narr, x, y, a, a_y, st);
  if (kind == KIDWT) return sepw3d_variant<R, Idwt<Sym8Rec>>(v, b, narr, x, y, a, a_y, st);
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t sepw3d(int kind, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                  hipStream_t st) {
  const int v = select_variant(sizeof(R), kind == KIDWT ? FORM_IDWT : FORM_DWT, id, b);
  return sepw3d_v<R>(v, kind, id, b, narr, x, y, a, a_y, st);
}

// ---- batches of independent 3D sym8 wavelet transforms
template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
static hipError_t wav_stage_build(const BatchItem<R>* it, int count, BatchStage* s) {
  std::vector<FusedArgs<R>> tab(count);
  std::vector<int> first(count + 1, 0);
  long tiles_all = 0;
  for (int o = 0; o < count; ++o) {
    tab[o] = wav_args<R, Op, TX, TY, QX, BZ, AD>(it[o].b, 1, it[o].x, it[o].y, (R)it[o].a[0], (R)it[o].a_y);
    tiles_all += (long)tab[o].tiles_x * tab[o].tiles_y;
  }
  const long want = 2L * num_cu();
  const long chunks = std::max(1L, (want + tiles_all - 1) / std::max(1L, tiles_all));
  long total = 0;
  for (int o = 0; o < count; ++o) {
    FusedArgs<R>& p = tab[o];
    const long c = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / (2 * Op::W))));
    p.zchunk = (int)((p.d2.nout + c - 1) / c);
    first[o] = (int)total;
    total += (long)p.tiles_x * p.tiles_y * ((p.d2.nout + p.zchunk - 1) / p.zchunk);
  }
  first[count] = (int)total;
  if (total > 0x7fffffffL) return hipErrorInvalidValue;
  s->blocks = (int)total;
  return batch_upload(tab.data(), sizeof(FusedArgs<R>) * count, first.data(), count, s);
}

template <typename R, class Op>
static hipError_t wav_stage_t(bool build, const BatchItem<R>* it, int count, BatchStage* s, hipStream_t st) {
  if (build) {
    switch (s->variant) {
#define X(I, TX, TY, QX, BZ, AD) \
  case I: return wav_stage_build<R, Op, TX, TY, QX, BZ, AD>(it, count, s);
      LC_WAV_VARIANTS(X)
#undef X
    }
    return hipErrorInvalidValue;
  }
  if (s->blocks == 0) return hipSuccess;
  const FusedArgs<R>* tab = static_cast<const FusedArgs<R>*>(s->tab);
  switch (s->variant) {
#define X(I, TX, TY, QX, BZ, AD)                                                                     \
  case I:                                                                                           \
    sepw3d_batch<R, Op, TX, TY, QX, BZ, AD><<<s->blocks, 2 * TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
    LC_WAV_VARIANTS(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t wav_stage(bool build, int kind, int id, int bc, const BatchItem<R>* it, int count,
                     BatchStage* s, hipStream_t st) {
  (void)bc;
  if (id != WAV_SYM8) return hipErrorNotSupported;
  if (kind == KDWT) return wav_stage_t<R, Dwt<Sym8Dec>>(build, it, count, s, st);
  if (kind == KIDWT) return wav_stage_t<R, Idwt<Sym8Rec>>(build, it, count, s, st);
  return hipErrorInvalidValue;
}

template hipError_t sepw3d<double>(int, int, const Box3&, int, const double*, double*, double, double,
                                   hipStream_t);
template hipError_t sepw3d<float>(int, int, const Box3&, int, const float*, float*, float, float,
                                  hipStream_t);
template hipError_t sepw3d_v<double>(int, int, int, const Box3&, int, const double*, double*, double, double,
                                     hipStream_t);
template hipError_t sepw3d_v<float>(int, int, int, const Box3&, int, const float*, float*, float, float,
                                    hipStream_t);
template hipError_t wav_stage<double>(bool, int, int, int, const BatchItem<double>*, int, BatchStage*,
                                      hipStream_t);
template hipError_t wav_stage<float>(bool, int, int, int, const BatchItem<float>*, int, BatchStage*,
                                     hipStream_t);

}  // namespace lc

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

#include "lc_body_wav.hpp"
#include "lc_body_wavq.hpp"

namespace lc {

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD, int WPE>
__global__ __launch_bounds__(2 * TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sepw3d_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[sepw3d_lds_bytes<R, Op, TX, TY, QX, BZ, AD>()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sepw3d_body<R, Op, TX, TY, QX, BZ, AD>(p, bx, by, zc, arr, lds);
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD, int WPE>
__global__ __launch_bounds__(2 * TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sepw3d_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[sepw3d_lds_bytes<R, Op, TX, TY, QX, BZ, AD>()];
  sepw3d_body<R, Op, TX, TY, QX, BZ, AD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// the plane-at-a-time form (lc_body_wavq.hpp): x-waves and y-waves, one
// barrier per input plane
template <typename R, class Op, int TX, int TY, int QX, int PD, int WPE, int KB>
__global__ __launch_bounds__((WavqGeo<R, Op, TX, TY, QX>::NT)) __attribute__((amdgpu_waves_per_eu(WPE))) void wavq_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[WavqGeo<R, Op, TX, TY, QX>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  if (p.flags & F_AY)
    wavq_body<R, Op, TX, TY, QX, PD, true, false, KB>(p, bx, by, zc, arr, lds);
  else
    wavq_body<R, Op, TX, TY, QX, PD, false, false, KB>(p, bx, by, zc, arr, lds);
}

template <typename R, class Op, int TX, int TY, int QX, int PD, int WPE, int KB>
__global__ __launch_bounds__((WavqGeo<R, Op, TX, TY, QX>::NT)) __attribute__((amdgpu_waves_per_eu(WPE))) void wavq_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[WavqGeo<R, Op, TX, TY, QX>::lds_bytes()];
  if (p.flags & F_AY)
    wavq_body<R, Op, TX, TY, QX, PD, true, false, KB>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
  else
    wavq_body<R, Op, TX, TY, QX, PD, false, false, KB>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// ====================================================================== host
// Compiled variants: tile TX x TY steps, QX outputs per x-task, BZ steps per
// window slide, AD prefetch depth (2 TX TY lanes per block), waves per EU,
// z chunks of a single call (0: ~2 blocks per CU, chunks >= 2 W steps;
// ZC < 0: at most -ZC blocks on any CU, chunks >= W steps)
#define LC_WAV_VARIANTS(X)                                                                            \
  X(0, 16, 8, 4, 2, 2, 1, 0) X(1, 8, 8, 4, 2, 2, 1, 0) X(2, 32, 4, 4, 2, 2, 1, 0) X(3, 16, 8, 4, 1, 1, 1, 0) \
  X(4, 16, 8, 4, 1, 1, 2, 0) X(5, 16, 8, 4, 1, 1, 1, -2) X(6, 16, 8, 4, 2, 2, 1, -2)

// plane-at-a-time variants: tile TX x TY steps, QX outputs per x-task, PD
// planes of prefetch, waves per EU, z chunks as above, scheduling fence
// interval of the filter reads (0: none)
#define LC_WAVQ_VARIANTS(X)                                                                                  \
  X(7, 16, 8, 4, 2, 1, -2, 0) X(8, 16, 8, 4, 2, 1, -3, 0) X(9, 32, 8, 4, 2, 1, -2, 0)                           \
  X(10, 16, 16, 4, 2, 1, -2, 0) X(11, 16, 8, 4, 4, 1, -2, 0) X(12, 16, 8, 4, 2, 4, -2, 2)                        \
  X(13, 16, 16, 4, 2, 4, -2, 2) X(14, 16, 8, 4, 2, 4, -2, 1) X(15, 16, 8, 4, 2, 1, -2, 2)                         \
  X(16, 8, 8, 2, 2, 1, -2, 0) X(17, 8, 8, 2, 2, 4, -2, 1)

int wav_variant_count() { return 18; }

// output planes reached through 32-bit buffer records (the plane-at-a-time
// variants store with buffer stores)
static bool wavq_fits(const Box3& b, int prec) {
  return wav_plane_fits(b, prec) && 4L * b.ny[0] * b.ny[1] * prec <= kPlaneMax;
}

std::string wav_kernel_name(int prec, int form, int id, int v, bool batch) {
  (void)id;
  const char* r = prec == 4 ? "float" : "double";
  const char* op = form == FORM_IDWT ? "Idwt<lc::WSplit<lc::FLP, lc::FHP, 8, true> >"
                                     : "Dwt<lc::WSplit<lc::FLP, lc::FHP, 8, false> >";
  char buf[200];
  buf[0] = 0;
  switch (v) {
#define X(I, TX, TY, QX, BZ, AD, WPE, ZC)                                                                   \
  case I:                                                                                         \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sepw3d_batch" : "sepw3d_fused", \
             r, op, TX, TY, QX, BZ, AD, WPE);                                                        \
    break;
    LC_WAV_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, PD, WPE, ZC, KB)                                                                   \
  case I:                                                                                                   \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "wavq_batch" : "wavq_fused", r, op, \
             TX, TY, QX, PD, WPE, KB);                                                                      \
    break;
    LC_WAVQ_VARIANTS(X)
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

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD, int WPE, int ZC>
static hipError_t launch_sepw3d(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st) {
  if (!wav_plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
  FusedArgs<R> p = wav_args<R, Op, TX, TY, QX, BZ, AD>(b, narr, x, y, a, a_y);
  const long tiles = (long)p.tiles_x * p.tiles_y;
  if (ZC < 0) {
    // at most -ZC blocks on any CU, at least one where the grid allows
    const long per = tiles * std::max(narr, 1);
    long chunks = (long)(-ZC) * num_cu() / per;
    if (chunks * per < num_cu()) chunks = (num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / Op::W)));
    p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
  } else {
    // z chunks no shorter than 2 W steps (W steps when the grid is too
    // small to give every CU a block)
    p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), std::max(1, Op::W / 2), num_cu());
  }
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sepw3d_fused<R, Op, TX, TY, QX, BZ, AD, WPE><<<grid, 2 * TX * TY, 0, st>>>(p);
  return hipGetLastError();
}

// z chunk of a single call: ZC < 0: at most -ZC blocks on any CU, at least
// one where the grid allows, chunks >= W steps; else ~2 blocks per CU,
// chunks >= 2 W steps
template <class Op, int ZC>
static int wav_zchunk(long tiles, int narr, int nout) {
  if (ZC < 0) {
    const long per = tiles * std::max(narr, 1);
    long chunks = (long)(-ZC) * num_cu() / per;
    if (chunks * per < num_cu()) chunks = (num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, nout / Op::W)));
    return (int)((nout + chunks - 1) / chunks);
  }
  return pick_zchunk(tiles, nout, std::max(narr, 1), std::max(1, Op::W / 2), num_cu());
}

template <typename R, class Op, int TX, int TY, int QX, int PD, int WPE, int ZC, int KB>
static hipError_t launch_wavq(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st) {
  if constexpr (!wavq_shape_ok<R, TX, TY, QX>()) {
    return hipErrorInvalidValue;  // a shape of the other precision
  } else {
  if (!wavq_fits(b, sizeof(R))) return hipErrorInvalidValue;
  FusedArgs<R> p = wav_args<R, Op, TX, TY, QX, 1, 1>(b, narr, x, y, a, a_y);
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = wav_zchunk<Op, ZC>(tiles, narr, p.d2.nout);
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  wavq_fused<R, Op, TX, TY, QX, PD, WPE, KB><<<grid, WavqGeo<R, Op, TX, TY, QX>::NT, 0, st>>>(p);
  return hipGetLastError();
  }
}

template <typename R, class Op>
static hipError_t sepw3d_variant(int v, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                                 hipStream_t st) {
  switch (v) {
#define X(I, TX, TY, QX, BZ, AD, WPE, ZC) \
  case I: return launch_sepw3d<R, Op, TX, TY, QX, BZ, AD, WPE, ZC>(b, narr, x, y, a, a_y, st);
    LC_WAV_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, PD, WPE, ZC, KB) \
  case I: return launch_wavq<R, Op, TX, TY, QX, PD, WPE, ZC, KB>(b, narr, x, y, a, a_y, st);
    LC_WAVQ_VARIANTS(X)
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
template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD, int WPE>
static hipError_t wav_stage_build(const BatchItem<R>* it, int count, BatchStage* s) {
  std::vector<FusedArgs<R>> tab(count);
  std::vector<int> first(count + 1, 0);
  long tiles_all = 0;
  for (int o = 0; o < count; ++o)
    if (!wav_plane_fits(it[o].b, sizeof(R))) return hipErrorInvalidValue;
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
#define X(I, TX, TY, QX, BZ, AD, WPE, ZC) \
  case I: return wav_stage_build<R, Op, TX, TY, QX, BZ, AD, WPE>(it, count, s);
      LC_WAV_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, PD, WPE, ZC, KB)                                  \
  case I:                                                                  \
    if (!wavq_shape_ok<R, TX, TY, QX>()) return hipErrorInvalidValue;     \
    for (int o = 0; o < count; ++o)                                        \
      if (!wavq_fits(it[o].b, sizeof(R))) return hipErrorInvalidValue;     \
    return wav_stage_build<R, Op, TX, TY, QX, 1, 1, WPE>(it, count, s);
      LC_WAVQ_VARIANTS(X)
#undef X
    }
    return hipErrorInvalidValue;
  }
  if (s->blocks == 0) return hipSuccess;
  const FusedArgs<R>* tab = static_cast<const FusedArgs<R>*>(s->tab);
  switch (s->variant) {
#define X(I, TX, TY, QX, BZ, AD, WPE, ZC)                                                                     \
  case I:                                                                                           \
    sepw3d_batch<R, Op, TX, TY, QX, BZ, AD, WPE><<<s->blocks, 2 * TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
    LC_WAV_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, PD, WPE, ZC, KB)                                                                 \
  case I:                                                                                                 \
    if constexpr (wavq_shape_ok<R, TX, TY, QX>())                                                          \
      wavq_batch<R, Op, TX, TY, QX, PD, WPE, KB><<<s->blocks, WavqGeo<R, Op, TX, TY, QX>::NT, 0, st>>>(    \
          tab, s->first, s->count);                                                                       \
    else                                                                                                  \
      return hipErrorInvalidValue;                                                                        \
    break;
    LC_WAVQ_VARIANTS(X)
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

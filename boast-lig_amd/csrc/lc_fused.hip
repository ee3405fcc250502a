// lc_fused.hip -- single-pass 3D libconv operators on gfx950.
//
// Both kernels march along the slowest dim (z) over an xy tile of the
// output, so every input value is read once from HBM and every output
// written once (the neighbouring tiles' halo reads hit L2 / Infinity Cache):
//
//   sum3d_fused  kinetic form y = a0 Fx x~ + a1 Fy x~ + a2 Fz x~ + a_x x~
//                + a_y y with the per-direction FP64 energies.  Per output
//                plane z the xy tile of x~(z) plus its cross-shaped filter
//                halo sits in LDS: x-tasks (QX consecutive outputs of a row
//                from one register window) write the x-terms to LDS, the
//                column owners compute the y-terms (CPT consecutive rows of
//                a column) and the z-terms from a register window of the
//                column's values along z, fed BZ planes ahead of use.
//   sep3d_fused  separable product y = a Fz Fy Fx x~ + a_y y.  Per input
//                plane p: the (TX+W-1) x (TY+W-1) tile goes to LDS, x-tasks
//                filter it along x into LDS, the column owners filter along
//                y straight into their register window along z and emit the
//                output plane p - UP.
// x~ is x wrapped along periodic dims and zero outside its box along free
// dims, exactly the per-dim semantics of the 1D sweeps
// (BigDFT/GenericConvolution2.rb:91-121 index ranges, :2160-2240 chains).
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include <algorithm>

#include "lc_fused.hpp"

#include "lc_body_conv.hpp"

namespace lc {

// ====================================================================== kernels
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD>
__global__ __launch_bounds__(TX * TY / CPT) void sum3d_fused(FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[sum3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, HD>()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sum3d_body<R, F, TX, TY, CPT, QX, BZ, HD>(p, bx, by, zc, arr,
                                           (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD>
__global__ __launch_bounds__(TX * TY / CPT) void sep3d_fused(FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[sep3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, AD>()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sep3d_body<R, F, TX, TY, CPT, QX, BZ, AD>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD>
__global__ __launch_bounds__(TX * TY / CPT) void sum3d_batch(const FusedArgs<R>* __restrict__ tab,
                                                              const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  const int bx = tile % p.tiles_x, by = tile / p.tiles_x;
  __shared__ __align__(16) unsigned char lds[sum3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, HD>()];
  sum3d_body<R, F, TX, TY, CPT, QX, BZ, HD>(p, bx, by, zc, 0, (long)local, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD>
__global__ __launch_bounds__(TX * TY / CPT) void sep3d_batch(const FusedArgs<R>* __restrict__ tab,
                                                              const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[sep3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, AD>()];
  sep3d_body<R, F, TX, TY, CPT, QX, BZ, AD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// ====================================================================== launch
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD>
static hipError_t launch_sum3d(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                               R a_y, double* dot, hipStream_t st, double* scratch,
                               size_t scratch_bytes) {
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
  p.a_x = a_x; p.a_y = a_y;
  p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
  p.nparts = tiles * zch;
  if (dot) {
    if (scratch_bytes < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
    p.part = scratch;
  }
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sum3d_fused<R, F, TX, TY, CPT, QX, BZ, HD><<<grid, TX * TY / CPT, 0, st>>>(p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD>
static hipError_t launch_sep3d(const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                               hipStream_t st) {
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = a; p.a_y = a_y;
  p.flags = (a != R(1) ? F_A : 0) | (a_y != R(0) ? F_AY : 0);
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sep3d_fused<R, F, TX, TY, CPT, QX, BZ, AD><<<grid, TX * TY / CPT, 0, st>>>(p);
  return hipGetLastError();
}

size_t fused_sum3d_parts(const Box3& b, int narr, int lo, int up, int w) {
  FDim g0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], lo, up);
  FDim g1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], lo, up);
  FDim g2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], lo, up);
  // upper bound over the variants: smallest tile, most z chunks
  const long tiles = (long)((g0.nout + 31) / 32) * ((g1.nout + 7) / 8);
  const long maxc = std::max(1, g2.nout / (4 * w)) + 1;
  (void)narr;
  return (size_t)(tiles * maxc);
}

// ====================================================================== variants
// Compiled variants of the two fused forms: tile TX x TY, CPT rows per lane,
// QX outputs per x-task, BZ planes per window slide, HD / AD prefetch depth.
// Indices 0..4 have batched forms; 5..7 are the wave-specialised kernels of
// lc_fused_ws.hip (single calls only).  The variant of a call comes from the
// tuner's table for its exact shape, else LIBCONV_{SUM3D,SEP3D}_VARIANT,
// else index 1 (fastest at 256^3 on MI355X).
#define LC_SUM_VARIANTS(X)                                                    \
  X(0, 32, 16, 2, 2, 4, 2) X(1, 32, 16, 2, 2, 4, 1) X(2, 32, 8, 2, 2, 4, 2)   \
  X(3, 64, 8, 2, 4, 4, 2) X(4, 32, 16, 2, 2, 2, 2)
#define LC_SEP_VARIANTS(X)                                                    \
  X(0, 32, 16, 2, 4, 4, 4) X(1, 32, 16, 2, 4, 4, 1) X(2, 32, 8, 2, 4, 4, 2)   \
  X(3, 64, 8, 2, 4, 4, 2) X(4, 32, 16, 2, 4, 4, 2)
#define LC_WS_VARIANTS(X) X(5, 32, 8) X(6, 64, 8) X(7, 32, 16)

template <typename R, class F>
static hipError_t sum3d_variant(int v, const Box3& b, int narr, const R* x, R* y, const double* a,
                                R a_x, R a_y, double* dot, hipStream_t st, double* sc, size_t sb) {
  switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, HD) \
  case I: return launch_sum3d<R, F, TX, TY, CPT, QX, BZ, HD>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY) \
  case I: return launch_sum3d_ws<R, F, TX, TY, 4, 4, 2>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb, 4, F::N);
    LC_WS_VARIANTS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename R, class F>
static hipError_t sep3d_variant(int v, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                                hipStream_t st) {
  switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, AD) \
  case I: return launch_sep3d<R, F, TX, TY, CPT, QX, BZ, AD>(b, narr, x, y, a, a_y, st);
    LC_SEP_VARIANTS(X)
#undef X
#define X(I, TX, TY) \
  case I: return launch_sep3d_ws<R, F, TX, TY, 4, 4, 2>(b, narr, x, y, a, a_y, st, 4, F::N);
    LC_WS_VARIANTS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

int variant_count(int form) { return form == FORM_SUM || form == FORM_SEP ? 8 : wav_variant_count(); }
int batch_variant_count(int form) { return form == FORM_SUM || form == FORM_SEP ? 5 : wav_variant_count(); }

static const char* filter_tag(int id) {
  switch (id) {
    case OP_MD: return "FMD";
    case OP_IMD: return "FIMD";
    case OP_D2: return "FD2";
    case OP_D1: return "FD1";
  }
  return "?";
}

std::string fused_kernel_name(int prec, int form, int id, int v, bool batch) {
  const char* r = prec == 4 ? "float" : "double";
  char buf[160];
  buf[0] = 0;
  const bool sum = form == FORM_SUM;
  switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, HD)                                                             \
  case I:                                                                                        \
    if (sum)                                                                                     \
      snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sum3d_batch" : \
               "sum3d_fused", r, filter_tag(id), TX, TY, CPT, QX, BZ, HD);                        \
    break;
    LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY) \
  case I: snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, 4, 4, 2>", sum ? "sum3d_ws" : "sep3d_ws", r, filter_tag(id), TX, TY); break;
    LC_WS_VARIANTS(X)
#undef X
  }
  if (!sum && v >= 0 && v < 5) {
    switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, AD)                                                             \
  case I:                                                                                        \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sep3d_batch" :  \
             "sep3d_fused", r, filter_tag(id), TX, TY, CPT, QX, BZ, AD);                          \
    break;
      LC_SEP_VARIANTS(X)
#undef X
    }
  }
  return buf;
}

static int env_variant(const char* name, int nvar) {
  const char* e = getenv(name);
  if (!e) return -1;
  const int v = atoi(e);
  return v >= 0 && v < nvar ? v : -1;
}

int select_variant(int prec, int form, int id, const Box3& b) {
  int v = tuned_variant(prec, form, id, b);
  if (v >= 0) return v;
  if (form == FORM_SUM || form == FORM_SEP) {
    v = env_variant(form == FORM_SUM ? "LIBCONV_SUM3D_VARIANT" : "LIBCONV_SEP3D_VARIANT", 8);
    return v >= 0 ? v : 1;
  }
  v = env_variant("LIBCONV_WAV3D_VARIANT", wav_variant_count());
  return v >= 0 ? v : 0;
}

template <typename R>
hipError_t sum3d_fused_v(int v, int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                         R a_y, double* dot, hipStream_t st, double* scratch, size_t sb) {
  switch (id) {
    case OP_D2: return sum3d_variant<R, FD2>(v, b, narr, x, y, a, a_x, a_y, dot, st, scratch, sb);
    case OP_D1: return sum3d_variant<R, FD1>(v, b, narr, x, y, a, a_x, a_y, dot, st, scratch, sb);
    case OP_MD: return sum3d_variant<R, FMD>(v, b, narr, x, y, a, a_x, a_y, dot, st, scratch, sb);
    case OP_IMD: return sum3d_variant<R, FIMD>(v, b, narr, x, y, a, a_x, a_y, dot, st, scratch, sb);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t sep3d_fused_v(int v, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                         hipStream_t st) {
  switch (id) {
    case OP_MD: return sep3d_variant<R, FMD>(v, b, narr, x, y, a, a_y, st);
    case OP_IMD: return sep3d_variant<R, FIMD>(v, b, narr, x, y, a, a_y, st);
    case OP_D2: return sep3d_variant<R, FD2>(v, b, narr, x, y, a, a_y, st);
    case OP_D1: return sep3d_variant<R, FD1>(v, b, narr, x, y, a, a_y, st);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t sum3d_fused_run(int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                           R a_y, double* dot, hipStream_t st, double* scratch, size_t sb) {
  const int v = select_variant(sizeof(R), FORM_SUM, id, b);
  return sum3d_fused_v<R>(v, id, b, narr, x, y, a, a_x, a_y, dot, st, scratch, sb);
}

template <typename R>
hipError_t sep3d_fused_run(int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st) {
  const int v = select_variant(sizeof(R), FORM_SEP, id, b);
  return sep3d_fused_v<R>(v, id, b, narr, x, y, a, a_y, st);
}

// ====================================================================== batches
hipError_t batch_upload(const void* tab, size_t tab_bytes, const int* first, int count, BatchStage* s) {
  hipError_t e = hipMalloc(&s->tab, std::max<size_t>(tab_bytes, 1));
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&s->first), sizeof(int) * (count + 1));
  if (e != hipSuccess) return e;
  e = hipMemcpy(s->tab, tab, tab_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  return hipMemcpy(s->first, first, sizeof(int) * (count + 1), hipMemcpyHostToDevice);
}

void batch_stage_free(BatchStage* s) {
  if (s->tab) (void)hipFree(s->tab);
  if (s->first) (void)hipFree(s->first);
  if (s->part) (void)hipFree(s->part);
  if (s->dots) (void)hipFree(s->dots);
  *s = BatchStage{};
}

// energies of item blockIdx.x: 3 sets of its blocks' partials, fixed order
__global__ __launch_bounds__(256) void batch_dots(const double* __restrict__ part,
                                                  const int* __restrict__ first,
                                                  double* const* __restrict__ dots) {
  __shared__ double red[4];
  const int o = blockIdx.x;
  double* dst = dots[o];
  if (!dst) return;  // uniform per block
  const long b0 = first[o], nb = first[o + 1] - first[o];
  const double* pp = part + 3 * b0;
  for (int k = 0; k < 3; ++k) {
    double v = 0.0;
    for (long i = threadIdx.x; i < nb; i += blockDim.x) v += pp[k * nb + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) dst[k] = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
}

template <typename R, class F, int TX, int TY>
static hipError_t fused_stage_build(int kind, const BatchItem<R>* it, int count, BatchStage* s) {
  std::vector<FusedArgs<R>> tab(count);
  std::vector<int> first(count + 1, 0);
  std::vector<double*> dots(count, nullptr);
  long tiles_all = 0;
  for (int o = 0; o < count; ++o) {
    const Box3& b = it[o].b;
    FusedArgs<R>& p = tab[o];
    p = FusedArgs<R>{};
    p.x = it[o].x; p.y = it[o].y;
    p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
    p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
    p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
    p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
    p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
    p.narr = 1;
    p.tiles_x = (p.d0.nout + TX - 1) / TX;
    p.tiles_y = (p.d1.nout + TY - 1) / TY;
    p.a0 = (R)it[o].a[0]; p.a1 = (R)it[o].a[1]; p.a2 = (R)it[o].a[2];
    p.a_x = (R)it[o].a_x; p.a_y = (R)it[o].a_y;
    if (kind == BK_SUM)
      p.flags = (p.a_x != R(0) ? F_AX : 0) | (p.a_y != R(0) ? F_AY : 0) | (it[o].dot ? F_DOT : 0);
    else
      p.flags = (p.a0 != R(1) ? F_A : 0) | (p.a_y != R(0) ? F_AY : 0);
    dots[o] = it[o].dot;
    if (it[o].dot) s->any_dot = true;
    tiles_all += (long)p.tiles_x * p.tiles_y;
  }
  // the same number of z chunks for every item: enough blocks for ~2 per CU
  const long chunks = std::max(1L, (2L * num_cu() + tiles_all - 1) / std::max(1L, tiles_all));
  long total = 0;
  for (int o = 0; o < count; ++o) {
    FusedArgs<R>& p = tab[o];
    const long c = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / (4 * F::N))));
    p.zchunk = std::max(1, (int)((p.d2.nout + c - 1) / c));
    const long nb = (long)p.tiles_x * p.tiles_y * ((p.d2.nout + p.zchunk - 1) / p.zchunk);
    first[o] = (int)total;
    p.nparts = nb;
    total += nb;
  }
  first[count] = (int)total;
  if (total > 0x7fffffffL) return hipErrorInvalidValue;
  s->blocks = (int)total;
  if (kind == BK_SUM && s->any_dot) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->part), sizeof(double) * 3 * std::max(total, 1L));
    if (e != hipSuccess) return e;
    for (int o = 0; o < count; ++o)
      if (it[o].dot) tab[o].part = s->part + 3L * first[o];
    e = hipMalloc(reinterpret_cast<void**>(&s->dots), sizeof(double*) * count);
    if (e != hipSuccess) return e;
    e = hipMemcpy(s->dots, dots.data(), sizeof(double*) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  return batch_upload(tab.data(), sizeof(FusedArgs<R>) * count, first.data(), count, s);
}

template <typename R, class F>
static hipError_t fused_stage_t(bool build, int kind, const BatchItem<R>* it, int count, BatchStage* s,
                                hipStream_t st) {
  if (build) {
    if (kind == BK_SUM) {
      switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, HD) \
  case I: return fused_stage_build<R, F, TX, TY>(kind, it, count, s);
        LC_SUM_VARIANTS(X)
#undef X
      }
    } else {
      switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, AD) \
  case I: return fused_stage_build<R, F, TX, TY>(kind, it, count, s);
        LC_SEP_VARIANTS(X)
#undef X
      }
    }
    return hipErrorInvalidValue;
  }
  if (s->blocks == 0) return hipSuccess;
  const FusedArgs<R>* tab = static_cast<const FusedArgs<R>*>(s->tab);
  if (kind == BK_SUM) {
    switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, HD)                                                                   \
  case I:                                                                                              \
    sum3d_batch<R, F, TX, TY, CPT, QX, BZ, HD><<<s->blocks, TX * TY / CPT, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SUM_VARIANTS(X)
#undef X
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !s->any_dot) return e;
    batch_dots<<<s->count, 256, 0, st>>>(s->part, s->first, s->dots);
  } else {
    switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, AD)                                                                   \
  case I:                                                                                              \
    sep3d_batch<R, F, TX, TY, CPT, QX, BZ, AD><<<s->blocks, TX * TY / CPT, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SEP_VARIANTS(X)
#undef X
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

template <typename R>
hipError_t fused_stage(bool build, int kind, int id, const BatchItem<R>* it, int count, BatchStage* s,
                       hipStream_t st) {
  if (kind != BK_SEP && kind != BK_SUM) return hipErrorInvalidValue;
  switch (id) {
    case OP_MD: return fused_stage_t<R, FMD>(build, kind, it, count, s, st);
    case OP_IMD: return fused_stage_t<R, FIMD>(build, kind, it, count, s, st);
    case OP_D2: return fused_stage_t<R, FD2>(build, kind, it, count, s, st);
    case OP_D1: return fused_stage_t<R, FD1>(build, kind, it, count, s, st);
  }
  return hipErrorInvalidValue;
}

template hipError_t fused_stage<double>(bool, int, int, const BatchItem<double>*, int, BatchStage*,
                                        hipStream_t);
template hipError_t fused_stage<float>(bool, int, int, const BatchItem<float>*, int, BatchStage*,
                                       hipStream_t);

template hipError_t sum3d_fused_run<double>(int, const Box3&, int, const double*, double*, const double*,
                                            double, double, double*, hipStream_t, double*, size_t);
template hipError_t sum3d_fused_run<float>(int, const Box3&, int, const float*, float*, const double*,
                                           float, float, double*, hipStream_t, double*, size_t);
template hipError_t sep3d_fused_run<double>(int, const Box3&, int, const double*, double*, double, double,
                                            hipStream_t);
template hipError_t sep3d_fused_run<float>(int, const Box3&, int, const float*, float*, float, float,
                                           hipStream_t);
template hipError_t sum3d_fused_v<double>(int, int, const Box3&, int, const double*, double*, const double*,
                                          double, double, double*, hipStream_t, double*, size_t);
template hipError_t sum3d_fused_v<float>(int, int, const Box3&, int, const float*, float*, const double*,
                                         float, float, double*, hipStream_t, double*, size_t);
template hipError_t sep3d_fused_v<double>(int, int, const Box3&, int, const double*, double*, double,
                                          double, hipStream_t);
template hipError_t sep3d_fused_v<float>(int, int, const Box3&, int, const float*, float*, float, float,
                                         hipStream_t);

}  // namespace lc

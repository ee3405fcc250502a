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

#include "lc_fused_impl.hpp"

namespace lc {

size_t fused_sum3d_parts(const Box3& b, int narr, int lo, int up, int w) {
  FDim g0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], lo, up);
  FDim g1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], lo, up);
  FDim g2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], lo, up);
  // upper bound over the variants: tiles no smaller than 16 x 8, z chunks no
  // shorter than 16 planes (two-plane kinetic form) or 4 w (the others)
  const long tiles = (long)((g0.nout + 15) / 16) * ((g1.nout + 7) / 8);
  const long maxc = std::max(std::max(1, g2.nout / (4 * w)), g2.nout / 16) + 2;
  (void)narr;
  return (size_t)(tiles * maxc);
}

// variant dispatch per filter, instantiated in lc_fused_f*.hip
template <typename R, class F>
hipError_t sum3d_variant(int v, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                         double* dot, hipStream_t st, double* sc, size_t sb);
template <typename R, class F>
hipError_t sep3d_variant(int v, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                         const R* pot, R* rho);
template <typename R, class F>
hipError_t fused_stage_t(bool build, int kind, const BatchItem<R>* it, int count, BatchStage* s,
                         hipStream_t st);
#define LC_FUSED_EXTERN(R, F)                                                                            \
  extern template hipError_t sum3d_variant<R, F>(int, const Box3&, int, const R*, R*, const double*, R, R, \
                                                 double*, hipStream_t, double*, size_t);                 \
  extern template hipError_t sep3d_variant<R, F>(int, const Box3&, int, const R*, R*, R, R, hipStream_t, \
                                                 const R*, R*);                                          \
  extern template hipError_t fused_stage_t<R, F>(bool, int, const BatchItem<R>*, int, BatchStage*,       \
                                                 hipStream_t);
LC_FUSED_EXTERN(double, FMD)
LC_FUSED_EXTERN(double, FIMD)
LC_FUSED_EXTERN(double, FD2)
LC_FUSED_EXTERN(double, FD1)
LC_FUSED_EXTERN(float, FMD)
LC_FUSED_EXTERN(float, FIMD)
LC_FUSED_EXTERN(float, FD2)
LC_FUSED_EXTERN(float, FD1)

// variants 0..15 of the fused tables (5..7: wave-specialised)
int variant_count(int form) {
  if (form == FORM_SUM) return kFusedVariantCount;
  if (form == FORM_SEP) return kSepVariantCount;
  return wav_variant_count();
}
bool has_batch_form(int form, int v) {
  if (form == FORM_SEP) return v >= 0 && v < kSepVariantCount && (v < 5 || v > 7);
  if (form == FORM_SUM) return v >= 0 && v < kFusedVariantCount && (v < 5 || v > 7);
  return v >= 0 && v < wav_variant_count();
}

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
  if (form == FORM_SUM) {
    switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, HD, WPE)                                                                \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                      \
             batch ? "sum3d_batch" : "sum3d_fused", r, filter_tag(id), TX, TY, CPT, QX, BZ, HD, WPE); \
    break;
      LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC)                                                        \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                          \
             batch ? "sum3d2_batch" : "sum3d2_fused", r, filter_tag(id), TX, TY, QX, QY, QZ, WPE, KC); \
    break;
      LC_SUM2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, U, WPE, ZC)                                                       \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                          \
             batch ? "sump_batch" : "sump_fused", r, filter_tag(id), TX, TY, QX, QY, PD, U, WPE);    \
    break;
      LC_SUMP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC, PW)                                                   \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d, %d>",                      \
             batch ? "sum3d2w_batch" : "sum3d2w_fused", r, filter_tag(id), TX, TY, QX, QY, QZ, WPE, KC, PW); \
    break;
      LC_SUM2W_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PH)                                                       \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                          \
             batch ? "sumt_batch" : "sumt_fused", r, filter_tag(id), TX, TY, QX, QY, QZ, WPE, PH);   \
    break;
      LC_SUMT_VARIANTS(X)
#undef X
    }
  } else {
    switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, AD, WPE)                                                                \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                      \
             batch ? "sep3d_batch" : "sep3d_fused", r, filter_tag(id), TX, TY, CPT, QX, BZ, AD, WPE); \
    break;
      LC_SEP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PD)                                                             \
  case I:                                                                                                \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d, %d>",                               \
             batch ? "sep3d2_batch" : "sep3d2_fused", r, filter_tag(id), TX, TY, QX, QY, QZ, WPE, PD);   \
    break;
      LC_SEP2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                                              \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sepp_batch" : "sepp_fused", \
             r, filter_tag(id), TX, TY, QX, QY, PD, WPE);                                            \
    break;
      LC_SEPP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                                              \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sepr_batch" : "sepr_fused", \
             r, filter_tag(id), TX, TY, QX, QY, PD, WPE);                                            \
    break;
      LC_SEPR_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                                              \
  case I:                                                                                           \
    snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, %d, %d, %d, %d>", batch ? "sepq_batch" : "sepq_fused", \
             r, filter_tag(id), TX, TY, QX, QY, PD, WPE);                                            \
    break;
      LC_SEPQ_VARIANTS(X)
#undef X
    }
  }
  switch (v) {
#define X(I, TX, TY) \
  case I: snprintf(buf, sizeof buf, "%s<%s, lc::%s, %d, %d, 4, 4, 2>", form == FORM_SUM ? "sum3d_ws" : "sep3d_ws", r, filter_tag(id), TX, TY); break;
    LC_WS_VARIANTS(X)
#undef X
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
    v = env_variant(form == FORM_SUM ? "LIBCONV_SUM3D_VARIANT" : "LIBCONV_SEP3D_VARIANT", variant_count(form));
    if (v >= 0) return v;
    // defaults: the fastest variants on MI355X (MF in FP64: the 128-bit-LDS
    // kernel on 32 x 16 tiles, 4-10% ahead of the others from the 128^3
    // chain and the 157^3 magic filter of the local Hamiltonian at 64^3 pairs
    // to the 284^3 chain; in FP32: the 128-bit-LDS kernel on 32 x 32 tiles on
    // output planes of 200^2 points and more, the pipelined kernel with two
    // outputs per task on smaller ones; K: the two-plane kernel with at most
    // two blocks per CU and two planes per step, in FP64 with the column
    // planes two steps in flight);
    // the plane-buffer forms (16 and up) take planes of at most kPlaneMax
    // bytes
    const bool fit = plane_fits(b, prec);
    if (form == FORM_SEP) {
      if (!fit) return 12;
      const int g = FMD::N - 1;  // outputs per dim: grow adds the filter length
      const long plane = (long)(b.bc[0] == GROW ? b.n[0] + g : b.n[0]) * (b.bc[1] == GROW ? b.n[1] + g : b.n[1]);
      if (prec == 8) return 62;
      return plane >= 200L * 200 ? 68 : 41;
    }
    if (!fit) return prec == 4 ? 7 : 8;
    if (prec == 8) return 46;
    // FP32: two planes per step on output planes of 150^2 points and more
    // (256^3 chain: 2% ahead; 171^3 grown output: 4%), four on smaller ones
    // (128^3: 9% ahead)
    const int gk = FD2::N - 1;
    const long kplane = (long)(b.bc[0] == GROW ? b.n[0] + gk : b.n[0]) * (b.bc[1] == GROW ? b.n[1] + gk : b.n[1]);
    return kplane >= 150L * 150 ? 31 : 32;
  }
  v = env_variant("LIBCONV_WAV3D_VARIANT", wav_variant_count());
  if (v >= 0) return v;
  // the plane-at-a-time kernel (its outputs go through 32-bit buffer
  // records): FP32 on 16 x 8 steps, FP64 on 8 x 8 steps with two outputs per
  // x-task (on 64^3 pairs 17% faster than the step-at-a-time kernel, on 96^3
  // within 4% either way); else one z-step per slide
  if (wav_plane_fits(b, prec) && 4L * b.ny[0] * b.ny[1] * prec <= kPlaneMax) return prec == 4 ? 7 : 16;
  return 5;
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
                         hipStream_t st, const R* pot, R* rho) {
  switch (id) {
    case OP_MD: return sep3d_variant<R, FMD>(v, b, narr, x, y, a, a_y, st, pot, rho);
    case OP_IMD: return sep3d_variant<R, FIMD>(v, b, narr, x, y, a, a_y, st, pot, rho);
    case OP_D2: return sep3d_variant<R, FD2>(v, b, narr, x, y, a, a_y, st, pot, rho);
    case OP_D1: return sep3d_variant<R, FD1>(v, b, narr, x, y, a, a_y, st, pot, rho);
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
hipError_t sep3d_fused_run(int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                           const R* pot, R* rho) {
  int v = select_variant(sizeof(R), FORM_SEP, id, b);
  if (pot && v >= 5 && v <= 7)  // the wave-specialised forms have no potential step
    v = plane_fits(b, sizeof(R)) ? (sizeof(R) == 4 ? 28 : 33) : 12;
  // with the potential step the 128-bit-LDS kernels take their general
  // epilogue (stores inside the y-waves' branch, conservative wait counts):
  // the pipelined kernel is faster there (the local Hamiltonian's 157^3
  // magic filter: 0.203 against 0.226 ms for the whole chain)
  if (pot && v >= 62 && v <= 72) v = sizeof(R) == 4 ? 41 : 45;
  return sep3d_fused_v<R>(v, id, b, narr, x, y, a, a_y, st, pot, rho);
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

hipError_t launch_batch_dots(const double* part, const int* first, double* const* dots, int count,
                             hipStream_t st) {
  batch_dots<<<count, 256, 0, st>>>(part, first, dots);
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
                                            hipStream_t, const double*, double*);
template hipError_t sep3d_fused_run<float>(int, const Box3&, int, const float*, float*, float, float,
                                           hipStream_t, const float*, float*);
template hipError_t sum3d_fused_v<double>(int, int, const Box3&, int, const double*, double*, const double*,
                                          double, double, double*, hipStream_t, double*, size_t);
template hipError_t sum3d_fused_v<float>(int, int, const Box3&, int, const float*, float*, const double*,
                                         float, float, double*, hipStream_t, double*, size_t);
template hipError_t sep3d_fused_v<double>(int, int, const Box3&, int, const double*, double*, double,
                                          double, hipStream_t, const double*, double*);
template hipError_t sep3d_fused_v<float>(int, int, const Box3&, int, const float*, float*, float, float,
                                         hipStream_t, const float*, float*);

}  // namespace lc

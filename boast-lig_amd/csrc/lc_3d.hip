// lc_3d.hip -- multi-dimensional libconv operators on gfx950.
//
//   sum3d  kinetic form y = sum_d a_d F_d x + a_x x + a_y y, where x is
//          wrapped along periodic dims and zero outside its box along free
//          dims (for periodic BCs: the :kinetic chain of
//          GenericConvolutionOperator, BigDFT/GenericConvolution2.rb:2160-2240,
//          and kinetic_per_ref, BigDFT/Kinetic.rb:981-1061), with the
//          per-direction energies dot_d = <x, a_d F_d x> accumulated in FP64.
//   sep3d  separable product y = a F2 F1 F0 x + a_y y (MagicFilter,
//          transposing composition, GenericConvolution2.rb:2160-2240 with
//          :transpose => 1; the final layout equals the input's axis order).
#include <stdlib.h>

#include <algorithm>

#include "lc_kernels.hpp"

namespace lc {

// LIBCONV_DIRECT=1 selects the unfused kernels (one lane per output for the
// sum, three sweeps for the product): a debugging aid, not a fast path.
bool direct_kernels() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("LIBCONV_DIRECT");
    v = (e && e[0] == '1') ? 1 : 0;
  }
  return v == 1;
}

// ------------------------------------------------------------ geometry
struct Dim3G {
  int n, bc, nx, ny, ox, oy, o_lo, next;  // next = number of output points
};

__host__ __device__ static inline Dim3G dim_geom(int n, int bc, int nx, int ny, int lo, int up) {
  Dim3G g;
  g.n = n; g.bc = bc; g.nx = nx; g.ny = ny;
  g.ox = bc == SHRINK ? lo : 0;
  g.oy = bc == GROW ? -up : 0;
  g.o_lo = bc == GROW ? -up : 0;
  g.next = bc == GROW ? n + (up - lo) : n;
  return g;
}

// value of x~ at logical index j along a dim (-1 when it is zero)
__device__ __forceinline__ int src_index(const Dim3G& g, int j) {
  if (g.bc == PER) return pmod(j, g.n);
  if (g.bc == GROW) return (j >= 0 && j < g.n) ? j : -1;
  return j - g.ox;
}

template <typename R>
__device__ __forceinline__ double block_sum3(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------ sum3d (direct)
// One lane per output point; the three line sums read x through the cache.
template <typename R, class F>
__global__ __launch_bounds__(256) void sum3d_direct(const R* __restrict__ x, R* __restrict__ y,
                                                      Dim3G g0, Dim3G g1, Dim3G g2, int narr,
                                                      long nti, long nto, R a0, R a1, R a2, R a_x,
                                                      R a_y, int flags, double* __restrict__ part) {
  __shared__ double red[4];
  const long npts = (long)g0.next * g1.next * g2.next;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int arr = blockIdx.y;
  double d0 = 0, d1 = 0, d2 = 0;
  if (t < npts) {
    const int i0 = g0.o_lo + (int)(t % g0.next);
    const int i1 = g1.o_lo + (int)((t / g0.next) % g1.next);
    const int i2 = g2.o_lo + (int)(t / ((long)g0.next * g1.next));
    const R* __restrict__ xa = x + arr * nti;
    R* __restrict__ ya = y + arr * nto;
    const long s1 = g0.nx, s2 = (long)g0.nx * g1.nx;
    const int c0 = src_index(g0, i0), c1 = src_index(g1, i1), c2 = src_index(g2, i2);
    const bool has_x = c0 >= 0 && c1 >= 0 && c2 >= 0;
    const R xc = has_x ? xa[c0 + c1 * s1 + c2 * s2] : R(0);
    R t0 = 0, t1 = 0, t2 = 0;
    if (c1 >= 0 && c2 >= 0) {
      const R* row = xa + c1 * s1 + c2 * s2;
#pragma unroll
      for (int l = 0; l < F::N; ++l) {
        const int j = src_index(g0, i0 + F::LO + l);
        if (j >= 0) t0 = fmar(row[j], tap<R, F>(l), t0);
      }
    }
    if (c0 >= 0 && c2 >= 0) {
      const R* col = xa + c0 + c2 * s2;
#pragma unroll
      for (int l = 0; l < F::N; ++l) {
        const int j = src_index(g1, i1 + F::LO + l);
        if (j >= 0) t1 = fmar(col[j * s1], tap<R, F>(l), t1);
      }
    }
    if (c0 >= 0 && c1 >= 0) {
      const R* col = xa + c0 + c1 * s1;
#pragma unroll
      for (int l = 0; l < F::N; ++l) {
        const int j = src_index(g2, i2 + F::LO + l);
        if (j >= 0) t2 = fmar(col[j * s2], tap<R, F>(l), t2);
      }
    }
    t0 = t0 * a0;
    t1 = t1 * a1;
    t2 = t2 * a2;
    if (has_x && (flags & F_DOT) && arr == (int)gridDim.y - 1) {
      d0 = (double)xc * (double)t0;
      d1 = (double)xc * (double)t1;
      d2 = (double)xc * (double)t2;
    }
    R tot = t0;
    if ((flags & F_AX) && has_x) tot = fmar(xc, a_x, tot);
    tot = t1 + tot;
    tot = t2 + tot;
    R* yp = ya + (i0 - g0.oy) + (long)(i1 - g1.oy) * g0.ny + (long)(i2 - g2.oy) * g0.ny * g1.ny;
    if (flags & F_AY) tot = fmar(a_y, *yp, tot);
    *yp = tot;
  }
  if (part) {
    const long nb = (long)gridDim.x;
    double s = block_sum3<R>(d0, red);
    if (threadIdx.x == 0 && arr == (int)gridDim.y - 1) part[blockIdx.x] = s;
    s = block_sum3<R>(d1, red);
    if (threadIdx.x == 0 && arr == (int)gridDim.y - 1) part[nb + blockIdx.x] = s;
    s = block_sum3<R>(d2, red);
    if (threadIdx.x == 0 && arr == (int)gridDim.y - 1) part[2 * nb + blockIdx.x] = s;
  }
}

template <typename R>
size_t sum3d_scratch_bytes(const Box3& b, int narr) {
  // the larger of the direct kernel's and the fused kernel's partials
  size_t best = 0;
  for (int id = 0; id < 4; ++id) {
    OpGeom gm = conv_geom(id);
    long npts = 1;
    for (int d = 0; d < 3; ++d) npts *= dim_geom(b.n[d], b.bc[d], b.nx[d], b.ny[d], gm.lo, gm.up).next;
    best = std::max(best, (size_t)((npts + 255) / 256));
    best = std::max(best, fused_sum3d_parts(b, narr, gm.lo, gm.up, gm.w));
    best = std::max(best, ws_sum3d_parts_bound(b, gm.lo, gm.up));
  }
  return sizeof(double) * 3 * best;
}

template <typename R, class F>
static hipError_t sum3d_t(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                          double* dot, hipStream_t st, double* scratch) {
  Dim3G g[3];
  for (int d = 0; d < 3; ++d) g[d] = dim_geom(b.n[d], b.bc[d], b.nx[d], b.ny[d], F::LO, F::UP);
  const long npts = (long)g[0].next * g[1].next * g[2].next;
  const long nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  const long nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  int flags = 0;
  if (a_x != R(0)) flags |= F_AX;
  if (a_y != R(0)) flags |= F_AY;
  if (dot) flags |= F_DOT;
  const unsigned nb = (unsigned)((npts + 255) / 256);
  if (narr == 0 || npts == 0) return hipSuccess;
  sum3d_direct<R, F><<<dim3(nb, narr), 256, 0, st>>>(x, y, g[0], g[1], g[2], narr, nti, nto,
                                                     (R)a[0], (R)a[1], (R)a[2], a_x, a_y, flags,
                                                     dot ? scratch : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, nb, dot, 3, nb, st);
}

template <typename R>
hipError_t sum3d(int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                 double* dot, hipStream_t st, double* scratch, size_t scratch_bytes) {
  if (dot && scratch_bytes < sum3d_scratch_bytes<R>(b, narr)) return hipErrorInvalidValue;
  if (!direct_kernels())
    return sum3d_fused_run<R>(id, b, narr, x, y, a, a_x, a_y, dot, st, scratch, scratch_bytes);
  switch (id) {
    case OP_D2: return sum3d_t<R, FD2>(b, narr, x, y, a, a_x, a_y, dot, st, scratch);
    case OP_D1: return sum3d_t<R, FD1>(b, narr, x, y, a, a_x, a_y, dot, st, scratch);
    case OP_MD: return sum3d_t<R, FMD>(b, narr, x, y, a, a_x, a_y, dot, st, scratch);
    case OP_IMD: return sum3d_t<R, FIMD>(b, narr, x, y, a, a_x, a_y, dot, st, scratch);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------ sep3d (sweeps)
template <typename R>
size_t sep3d_scratch_bytes(const Box3& b, int narr) {
  // intermediates after sweep 0 (ny0, nx1, nx2) and sweep 1 (ny0, ny1, nx2)
  const size_t w1 = (size_t)b.ny[0] * b.nx[1] * b.nx[2];
  const size_t w2 = (size_t)b.ny[0] * b.ny[1] * b.nx[2];
  return sizeof(R) * (size_t)std::max(narr, 1) * (w1 + w2);
}

template <typename R>
hipError_t sep3d(int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                 double* scratch, size_t scratch_bytes) {
  if (!direct_kernels()) return sep3d_fused_run<R>(id, b, narr, x, y, a, a_y, st);
  if (scratch_bytes < sep3d_scratch_bytes<R>(b, narr)) return hipErrorInvalidValue;
  R* w1 = reinterpret_cast<R*>(scratch);
  R* w2 = w1 + (size_t)std::max(narr, 1) * b.ny[0] * b.nx[1] * b.nx[2];
  OpGeom gm = conv_geom(id);
  int cur[3] = {b.nx[0], b.nx[1], b.nx[2]};
  const R* src = x;
  for (int s = 0; s < 3; ++s) {
    R* dst = s == 2 ? y : (s == 0 ? w1 : w2);
    SweepArgs<R> p{};
    p.x = src;
    p.y = dst;
    long L = 1, Rr = 1;
    for (int i = 0; i < s; ++i) L *= cur[i];
    for (int i = s + 1; i < 3; ++i) Rr *= cur[i];
    p.L = L;
    p.nr = Rr;
    p.narr = narr;
    p.ncols = L * Rr * narr;
    p.nx = cur[s];
    p.ny = b.ny[s];
    p.line_in = L * p.nx;
    p.line_out = L * p.ny;
    p.nti = p.line_in * Rr;
    p.nto = p.line_out * Rr;
    p.n = b.n[s];
    p.ox = b.bc[s] == SHRINK ? gm.lo : 0;
    p.oy = b.bc[s] == GROW ? -gm.up : 0;
    p.i_first = b.bc[s] == GROW ? -gm.up : 0;
    p.nout = b.bc[s] == GROW ? b.n[s] + gm.w - 1 : b.n[s];
    p.a = s == 2 ? a : R(1);
    p.a_y = a_y;
    p.flags = (s == 2 && a != R(1) ? F_A : 0) | (s == 2 && a_y != R(0) ? F_AY : 0);
    p.dot_part = nullptr;
    hipError_t e = sweep<R>(KCONV, id, b.bc[s], p, st, nullptr);
    if (e != hipSuccess) return e;
    cur[s] = b.ny[s];
    src = dst;
  }
  return hipSuccess;
}

template hipError_t sum3d<double>(int, const Box3&, int, const double*, double*, const double*, double,
                                  double, double*, hipStream_t, double*, size_t);
template hipError_t sum3d<float>(int, const Box3&, int, const float*, float*, const double*, float,
                                 float, double*, hipStream_t, double*, size_t);
template hipError_t sep3d<double>(int, const Box3&, int, const double*, double*, double, double,
                                  hipStream_t, double*, size_t);
template hipError_t sep3d<float>(int, const Box3&, int, const float*, float*, float, float, hipStream_t,
                                 double*, size_t);
template size_t sum3d_scratch_bytes<double>(const Box3&, int);
template size_t sum3d_scratch_bytes<float>(const Box3&, int);
template size_t sep3d_scratch_bytes<double>(const Box3&, int);
template size_t sep3d_scratch_bytes<float>(const Box3&, int);

}  // namespace lc

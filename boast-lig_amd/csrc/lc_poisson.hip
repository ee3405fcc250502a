// lc_poisson.hip -- the Poisson nabla operators (BigDFT/Poisson2.rb) on the GPU.
//
// First derivatives of order nord in {2, 4, 6, 8, 16} along the dims of a 3D
// grid, with the border rows of GenericConvolution2.rb:237-294 (PoissonFilter):
// on a periodic dim every output uses the centred row with wrap-around; on a
// non-periodic dim (bc -2) output i < m uses row i - m on inputs 0 .. 2m and
// output i >= n - m row i - n + m + 1 on inputs n - 1 - 2m .. n - 1.  The
// weights are the exact Lagrange weights of lc_taps.h.
//
// These stencils are short (nord + 1 taps) and carry ~2 flop per byte, so the
// kernels are HBM / L2 bound: one lane per output point, x fastest across the
// lanes of a wave (coalesced loads and stores); the taps of one output along
// a strided dim are loads of the same lanes at a fixed offset, which the L2
// serves: each XCD works through its own slab of every plane in z order
// (pblock), so the planes a z stencil needs are in its L2.  Each operator is
// one pass:
//   poisson_1d:          du = a D_idim u (+ a_y du)          (Poisson_broker)
//   div_u_i:             du = sum_k D_k u_k / h_k, optional cc
//   nabla_u:             du_k = D_k u / h_k
//   nabla_u_and_square:  the same and du2 = |du|^2
// The derivative sums run in increasing tap order, as the reference loops.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lc_common.hpp"
#include "lc_kernels.hpp"
#include "lc_taps.h"

namespace lc {

namespace {

#define LC_PQD(p, q) ((double)(p) / (double)(q)),
#define LC_PQF(p, q) (float)((double)(p) / (double)(q)),
__constant__ double kPoiD2[] = {LC_POISSON_2(LC_PQD)};
__constant__ double kPoiD4[] = {LC_POISSON_4(LC_PQD)};
__constant__ double kPoiD6[] = {LC_POISSON_6(LC_PQD)};
__constant__ double kPoiD8[] = {LC_POISSON_8(LC_PQD)};
__constant__ double kPoiD16[] = {LC_POISSON_16(LC_PQD)};
__constant__ float kPoiF2[] = {LC_POISSON_2(LC_PQF)};
__constant__ float kPoiF4[] = {LC_POISSON_4(LC_PQF)};
__constant__ float kPoiF6[] = {LC_POISSON_6(LC_PQF)};
__constant__ float kPoiF8[] = {LC_POISSON_8(LC_PQF)};
__constant__ float kPoiF16[] = {LC_POISSON_16(LC_PQF)};
#undef LC_PQD
#undef LC_PQF

template <typename R, int NORD>
__device__ __forceinline__ const R* ptab() {
  if constexpr (sizeof(R) == 8) {
    if constexpr (NORD == 2) return kPoiD2;
    else if constexpr (NORD == 4) return kPoiD4;
    else if constexpr (NORD == 6) return kPoiD6;
    else if constexpr (NORD == 8) return kPoiD8;
    else return kPoiD16;
  } else {
    if constexpr (NORD == 2) return kPoiF2;
    else if constexpr (NORD == 4) return kPoiF4;
    else if constexpr (NORD == 6) return kPoiF6;
    else if constexpr (NORD == 8) return kPoiF8;
    else return kPoiF16;
  }
}

// unscaled derivative at point i of a line of n points (stride s) starting at
// `line`; per: periodic
template <typename R, int NORD>
__device__ __forceinline__ R pderiv(const R* __restrict__ line, long s, int n, int i, bool per) {
  constexpr int M = NORD / 2, W = NORD + 1;
  const R* c = ptab<R, NORD>();
  R t = R(0);
  if (i >= M && i < n - M) {  // interior: centred row, no wrap
    const R* row = c + M * W;
#pragma unroll
    for (int j = -M; j <= M; ++j) t = fmar(row[j + M], line[(long)(i + j) * s], t);
  } else if (per) {
    const R* row = c + M * W;
#pragma unroll
    for (int j = -M; j <= M; ++j) t = fmar(row[j + M], line[(long)pmod(i + j, n) * s], t);
  } else if (i < M) {
    const R* row = c + i * W;  // row i - m
#pragma unroll
    for (int j = -M; j <= M; ++j) t = fmar(row[j + M], line[(long)(j + M) * s], t);
  } else {
    const R* row = c + (i - n + 2 * M + 1) * W;  // row i - n + m + 1
#pragma unroll
    for (int j = -M; j <= M; ++j) t = fmar(row[j + M], line[(long)(n - 1 - M + j) * s], t);
  }
  return t;
}

struct PGeom {
  int n0, n1, n2;
  long np;
  bool per0, per1, per2;
};

__device__ __forceinline__ void pcoords(const PGeom& g, long p, int& i0, int& i1, int& i2) {
  i0 = (int)(p % g.n0);
  const long q = p / g.n0;
  i1 = (int)(q % g.n1);
  i2 = (int)(q / g.n1);
}

template <typename R, int NORD>
__device__ __forceinline__ R dk(const PGeom& g, const R* __restrict__ f, long p, int k, int i0, int i1,
                                int i2) {
  const long s1 = g.n0, s2 = (long)g.n0 * g.n1;
  if (k == 0) return pderiv<R, NORD>(f + (p - i0), 1, g.n0, i0, g.per0);
  if (k == 1) return pderiv<R, NORD>(f + (p - (long)i1 * s1), s1, g.n1, i1, g.per1);
  return pderiv<R, NORD>(f + (p - (long)i2 * s2), s2, g.n2, i2, g.per2);
}

constexpr int kPThreads = 256;

// Block -> (plane z, 256-point segment of the plane).  The dispatcher hands
// consecutive block ids to the 8 XCDs in turn; XCD k serves the segments
// [k S / 8, (k + 1) S / 8) of every plane, plane after plane, so the 2m + 1
// planes a z stencil reads were loaded into that XCD's L2 by its own recent
// blocks.  Blocks beyond their XCD's share exit at once.
__device__ __forceinline__ bool pblock(const PGeom& g, int segs, long& p0, int& count) {
  const int b = blockIdx.x, x = b & 7;
  const long l = b >> 3;
  const int lo = (int)((long)segs * x / 8), hi = (int)((long)segs * (x + 1) / 8);
  const int c = hi - lo;
  if (c <= 0 || l >= (long)c * g.n2) return false;
  const long plane = (long)g.n0 * g.n1;
  const int z = (int)(l / c), seg = lo + (int)(l % c);
  p0 = (long)z * plane + (long)seg * kPThreads;
  count = (int)std::min<long>(kPThreads, plane - (long)seg * kPThreads);
  return true;
}

template <typename R, int NORD>
__global__ __launch_bounds__(kPThreads) void poisson_1d_kernel(PGeom g, int segs, int idim, const R* __restrict__ u,
                                                               R* __restrict__ du, R a, R a_y) {
  long p0;
  int count;
  if (!pblock(g, segs, p0, count) || (int)threadIdx.x >= count) return;
  const long p = p0 + threadIdx.x;
  int i0, i1, i2;
  pcoords(g, p, i0, i1, i2);
  R t = dk<R, NORD>(g, u, p, idim, i0, i1, i2) * a;
  if (a_y != R(0)) t = fmar(a_y, du[p], t);
  du[p] = t;
}

template <typename R, int NORD>
__global__ __launch_bounds__(kPThreads) void div_u_i_kernel(PGeom g, int segs, const R* __restrict__ u,
                                                            R* __restrict__ du, R* __restrict__ cc, R h0, R h1, R h2) {
  long p0;
  int count;
  if (!pblock(g, segs, p0, count) || (int)threadIdx.x >= count) return;
  const long p = p0 + threadIdx.x;
  const R* u0 = u;
  const R* u1 = u + g.np;
  const R* u2 = u + 2 * g.np;
  int i0, i1, i2;
  pcoords(g, p, i0, i1, i2);
  const R dx = dk<R, NORD>(g, u0, p, 0, i0, i1, i2) / h0;
  const R dy = dk<R, NORD>(g, u1, p, 1, i0, i1, i2) / h1;
  const R dz = dk<R, NORD>(g, u2, p, 2, i0, i1, i2) / h2;
  du[p] = (dx + dy) + dz;
  if (cc) {  // uniform
    const R x = u0[p], y = u1[p], z = u2[p];
    const R xy = dk<R, NORD>(g, u1, p, 0, i0, i1, i2) / h0;
    const R xz = dk<R, NORD>(g, u2, p, 0, i0, i1, i2) / h0;
    const R yz = dk<R, NORD>(g, u2, p, 1, i0, i1, i2) / h1;
    cc[p] = (((x * x) * dx + (y * y) * dy) + (z * z) * dz) + R(2) * ((x * y * xy + x * z * xz) + y * z * yz);
  }
}

template <typename R, int NORD>
__global__ __launch_bounds__(kPThreads) void nabla_u_kernel(PGeom g, int segs, const R* __restrict__ u,
                                                            R* __restrict__ du, R* __restrict__ du2, R h0, R h1, R h2) {
  long p0;
  int count;
  if (!pblock(g, segs, p0, count) || (int)threadIdx.x >= count) return;
  const long p = p0 + threadIdx.x;
  int i0, i1, i2;
  pcoords(g, p, i0, i1, i2);
  const R dx = dk<R, NORD>(g, u, p, 0, i0, i1, i2) / h0;
  const R dy = dk<R, NORD>(g, u, p, 1, i0, i1, i2) / h1;
  const R dz = dk<R, NORD>(g, u, p, 2, i0, i1, i2) / h2;
  du[p] = dx;
  du[g.np + p] = dy;
  du[2 * g.np + p] = dz;
  if (du2) du2[p] = (dx * dx + dy * dy) + dz * dz;  // uniform branch
}

PGeom pgeom(const int* n, int geocode, int bc_idim, int idim) {
  PGeom g;
  g.n0 = n[0]; g.n1 = n[1]; g.n2 = n[2];
  g.np = (long)n[0] * n[1] * n[2];
  if (idim >= 0) {  // one dim, its own bc
    g.per0 = g.per1 = g.per2 = bc_idim == 0;
  } else {          // div_u_i_ref's perx / pery / perz
    g.per0 = geocode != 1;
    g.per1 = geocode == 2;
    g.per2 = geocode != 1;
  }
  return g;
}

int psegs(const PGeom& g) { return (int)(((long)g.n0 * g.n1 + kPThreads - 1) / kPThreads); }

// 8 XCD shares of ceil(S / 8) segments times n2 planes
unsigned pgrid(const PGeom& g) { return (unsigned)(8L * ((psegs(g) + 7) / 8) * g.n2); }

}  // namespace

bool poisson_order_ok(int nord) { return nord == 2 || nord == 4 || nord == 6 || nord == 8 || nord == 16; }

#define LC_POISSON_DISPATCH(NORD_VAR, CALL) \
  switch (NORD_VAR) {                       \
    case 2: { constexpr int NO = 2; CALL; } break;   \
    case 4: { constexpr int NO = 4; CALL; } break;   \
    case 6: { constexpr int NO = 6; CALL; } break;   \
    case 8: { constexpr int NO = 8; CALL; } break;   \
    case 16: { constexpr int NO = 16; CALL; } break; \
    default: return hipErrorInvalidValue;            \
  }

template <typename R>
hipError_t poisson_1d(int nord, int idim, const int* n, int bc, const R* u, R* du, R a, R a_y, hipStream_t st) {
  const PGeom g = pgeom(n, 0, bc, idim);
  if (g.np == 0) return hipSuccess;
  LC_POISSON_DISPATCH(nord, (poisson_1d_kernel<R, NO><<<pgrid(g), kPThreads, 0, st>>>(g, psegs(g), idim, u, du, a, a_y)))
  return hipGetLastError();
}

template <typename R>
hipError_t div_u_i(int geocode, const int* n, const R* u, R* du, R* cc, int nord, const double* h, hipStream_t st) {
  const PGeom g = pgeom(n, geocode, 0, -1);
  if (g.np == 0) return hipSuccess;
  LC_POISSON_DISPATCH(nord, (div_u_i_kernel<R, NO><<<pgrid(g), kPThreads, 0, st>>>(g, psegs(g), u, du, cc, (R)h[0],
                                                                                       (R)h[1], (R)h[2])))
  return hipGetLastError();
}

template <typename R>
hipError_t nabla_u(int geocode, const int* n, const R* u, R* du, R* du2, int nord, const double* h, hipStream_t st) {
  const PGeom g = pgeom(n, geocode, 0, -1);
  if (g.np == 0) return hipSuccess;
  LC_POISSON_DISPATCH(nord, (nabla_u_kernel<R, NO><<<pgrid(g), kPThreads, 0, st>>>(g, psegs(g), u, du, du2, (R)h[0],
                                                                                       (R)h[1], (R)h[2])))
  return hipGetLastError();
}

#undef LC_POISSON_DISPATCH

template hipError_t poisson_1d<double>(int, int, const int*, int, const double*, double*, double, double,
                                       hipStream_t);
template hipError_t poisson_1d<float>(int, int, const int*, int, const float*, float*, float, float, hipStream_t);
template hipError_t div_u_i<double>(int, const int*, const double*, double*, double*, int, const double*,
                                    hipStream_t);
template hipError_t div_u_i<float>(int, const int*, const float*, float*, float*, int, const double*, hipStream_t);
template hipError_t nabla_u<double>(int, const int*, const double*, double*, double*, int, const double*,
                                    hipStream_t);
template hipError_t nabla_u<float>(int, const int*, const float*, float*, float*, int, const double*, hipStream_t);

}  // namespace lc

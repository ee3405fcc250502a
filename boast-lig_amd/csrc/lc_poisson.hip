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
//   nabla_u_square:      du2 only
//   nabla_u_epsilon:     du_k = eps D_k u / h_k
//   update_rhopol:       res = eta ((dlogeps . du) / (4 pi) - rhopol),
//                        rhopol += res, rhores2 = sum res^2
// The derivative sums run in increasing tap order, as the reference loops;
// the 1 / h_k factors are products with the rounded reciprocals (within an
// ulp of the quotients: a division is a dozen FP64 instructions).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <stdlib.h>

#include "lc_common.hpp"
#include "lc_fused.hpp"
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
// the same rows as compile-time constants: the z-marching kernels' centred
// taps then become instruction operands, not loads
template <int NORD>
struct PoiRows;
#define LC_POI_ROWS(NO)                                       \
  template <>                                                 \
  struct PoiRows<NO> {                                        \
    static constexpr double d[] = {LC_POISSON_##NO(LC_PQD)};  \
    static constexpr float f[] = {LC_POISSON_##NO(LC_PQF)};   \
  };
LC_POI_ROWS(2)
LC_POI_ROWS(4)
LC_POI_ROWS(6)
LC_POI_ROWS(8)
LC_POI_ROWS(16)
#undef LC_POI_ROWS
#undef LC_PQD
#undef LC_PQF

template <typename R, int NORD>
__device__ __forceinline__ constexpr R prow(int k) {
  if constexpr (sizeof(R) == 8) return PoiRows<NORD>::d[k]; else return PoiRows<NORD>::f[k];
}

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
  int z0, nz;  // planes the one-point kernels compute: z0 .. z0 + nz - 1
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
  if (c <= 0 || l >= (long)c * g.nz) return false;
  const long plane = (long)g.n0 * g.n1;
  const int z = g.z0 + (int)(l / c), seg = lo + (int)(l % c);
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
                                                            R* __restrict__ du, R* __restrict__ cc, R rh0, R rh1, R rh2) {
  long p0;
  int count;
  if (!pblock(g, segs, p0, count) || (int)threadIdx.x >= count) return;
  const long p = p0 + threadIdx.x;
  const R* u0 = u;
  const R* u1 = u + g.np;
  const R* u2 = u + 2 * g.np;
  int i0, i1, i2;
  pcoords(g, p, i0, i1, i2);
  const R dx = dk<R, NORD>(g, u0, p, 0, i0, i1, i2) * rh0;
  const R dy = dk<R, NORD>(g, u1, p, 1, i0, i1, i2) * rh1;
  const R dz = dk<R, NORD>(g, u2, p, 2, i0, i1, i2) * rh2;
  du[p] = (dx + dy) + dz;
  if (cc) {  // uniform
    const R x = u0[p], y = u1[p], z = u2[p];
    const R xy = dk<R, NORD>(g, u1, p, 0, i0, i1, i2) * rh0;
    const R xz = dk<R, NORD>(g, u2, p, 0, i0, i1, i2) * rh0;
    const R yz = dk<R, NORD>(g, u2, p, 1, i0, i1, i2) * rh1;
    cc[p] = (((x * x) * dx + (y * y) * dy) + (z * z) * dz) + R(2) * ((x * y * xy + x * z * xz) + y * z * yz);
  }
}

// What the gradient kernels do with the scaled derivatives (dx, dy, dz) of
// point p: MODE 2 (nabla_u) du_k = d_k, and du2 = (dx^2 + dy^2) + dz^2 when
// du2 is given; 3 (nabla_u_square) du2 only; 4 (nabla_u_epsilon) du_k =
// d_k eps; 5 (update_rhopol) res = ((e1 dx + e2 dy) + e3 dz) / (4 pi) with
// (e1, e2, e3) = dlogeps(:, p), res = (res - rhopol) eta, rhopol = res +
// rhopol, and res^2 returned for the block's partial sum.  The products and
// sums follow the reference's order (Poisson2.rb:2278-2550).
enum { GE_GRAD = 2, GE_SQUARE = 3, GE_EPS = 4, GE_RHOPOL = 5 };
template <typename R>
struct GradEpi {
  R* du;
  R* du2;
  const R* eps;
  const R* dlogeps;  // (3, n0, n1, n2), the component fastest
  R* rhopol;
  double* part;      // update_rhopol: one partial sum of res^2 per block
  R eta;
};

template <typename R, int MODE>
__device__ __forceinline__ double grad_epilogue(const GradEpi<R>& e, long np, long p, R dx, R dy, R dz) {
  if constexpr (MODE == GE_GRAD) {
    e.du[p] = dx;
    e.du[np + p] = dy;
    e.du[2 * np + p] = dz;
    if (e.du2) e.du2[p] = (dx * dx + dy * dy) + dz * dz;  // uniform branch
  } else if constexpr (MODE == GE_SQUARE) {
    e.du2[p] = (dx * dx + dy * dy) + dz * dz;
  } else if constexpr (MODE == GE_EPS) {
    const R ep = e.eps[p];
    e.du[p] = dx * ep;
    e.du[np + p] = dy * ep;
    e.du[2 * np + p] = dz * ep;
  } else {
    const R* d = e.dlogeps + 3 * p;
    const R oneo4pi = (R)0.07957747154594766788444188168625718101722982287022822437383367203;  // 1 / (4 pi)
    R res = (d[0] * dx + d[1] * dy) + d[2] * dz;
    res = res * oneo4pi;
    const R rho = e.rhopol[p];
    res = (res - rho) * e.eta;
    e.rhopol[p] = res + rho;
    return (double)(res * res);
  }
  return 0.0;
}

// sum of v over a block of NT threads in a fixed order (shuffle tree in each
// wave, then the waves in order); the result in thread 0
template <int NT>
__device__ __forceinline__ double pblock_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[w];
  return s;
}

template <typename R, int NORD, int MODE>
__global__ __launch_bounds__(kPThreads) void grad_point_kernel(PGeom g, int segs, const R* __restrict__ u,
                                                               GradEpi<R> ep, R rh0, R rh1, R rh2) {
  long p0;
  int count;
  const bool blk = pblock(g, segs, p0, count);
  if constexpr (MODE != GE_RHOPOL) {
    if (!blk || (int)threadIdx.x >= count) return;
  } else if (!blk) {  // uniform: the block's partial is 0
    if (threadIdx.x == 0) ep.part[blockIdx.x] = 0.0;
    return;
  }
  double sq = 0.0;
  if ((int)threadIdx.x < count) {
    const long p = p0 + threadIdx.x;
    int i0, i1, i2;
    pcoords(g, p, i0, i1, i2);
    const R dx = dk<R, NORD>(g, u, p, 0, i0, i1, i2) * rh0;
    const R dy = dk<R, NORD>(g, u, p, 1, i0, i1, i2) * rh1;
    const R dz = dk<R, NORD>(g, u, p, 2, i0, i1, i2) * rh2;
    sq = grad_epilogue<R, MODE>(ep, g.np, p, dx, dy, dz);
  }
  if constexpr (MODE == GE_RHOPOL) {
    __shared__ double red[kPThreads / 64];
    const double s = pblock_sum<kPThreads>(sq, red);
    if (threadIdx.x == 0) ep.part[blockIdx.x] = s;
  }
}

// ------------------------------------------------------------ z-marching form
// div_u_i and nabla_u as one z march per TX x TY tile of output columns.  Per
// plane z the tiles of the inputs (the tile plus an M-point halo along x, and
// along y where a y-derivative is taken) go to LDS from registers loaded two
// planes ahead; every lane owns one output column, takes its x- and
// y-stencils from LDS and the z-stencil of the last input from a register
// window of its column along z.  Each input value is read from HBM about
// once, and the stencil reads hit LDS, not L2.
//
// Non-periodic borders: an output within M points of the border uses the
// shifted row on inputs 0 .. 2m or n - 1 - 2m .. n - 1.  A tile holding such
// outputs along x or y shifts its LDS window to [0, T + 2m) or
// [n - T - 2m, n), which holds every input its outputs use, and its border
// lanes read their row from LDS too (the rows are copied to LDS once).  On a
// non-periodic z the 2m border planes are left to the one-point kernels, and
// the march covers planes m .. n - m - 1.  Taps, order and operations equal
// pderiv's, so both forms give identical bits.
constexpr int kZTX = 32;
// tile rows (every mode): 16 rows read half the y-halo rows per output that
// 8 rows do (update_rhopol at 256^3, order 16: 0.207 -> 0.184 ms; div_u_i
// with cc 0.285 -> 0.277 ms)
template <int MODE>
constexpr int pz_ty() { return 16; }

// MODE 0: div_u_i, 1: div_u_i with cc, 2..5: the gradient family (GE_*).
// Derivatives taken of input k (X, Y in LDS; Z in the register window).
template <int MODE, int K>
struct PNeed {
  static constexpr bool X = MODE == 0 ? K == 0 : true;
  static constexpr bool Y = MODE == 0 ? K == 1 : (MODE == 1 ? K >= 1 : true);
  static constexpr bool TILE = X || Y || MODE == 1;  // cc also needs the centre values
};

template <typename R, int NORD, int TY>
struct PzGeo {
  static constexpr int M = NORD / 2, W = NORD + 1;
  static constexpr int NT = kZTX * TY;  // lanes: one output column each
  static constexpr int TWP = (kZTX + 2 * M) | 1, TH = TY + 2 * M;
  static constexpr int TSZ = TH * TWP + 1;  // one tile plus a spare slot
  static constexpr int nload(bool x, bool y) { return (TY + (y ? 2 * M : 0)) * (kZTX + (x ? 2 * M : 0)); }
  static constexpr int nlane(bool x, bool y) { return (nload(x, y) + NT - 1) / NT; }
};

// first index of a tile's LDS window along one dim: the centred halo, or on
// a non-periodic dim the window [0, T + 2m) / [n - T - 2m, n) of a tile that
// holds border outputs (the low rule wins when both apply)
__device__ __forceinline__ int pz_window(int o0, int t, int m, int n, bool per) {
  if (per) return o0 - m;
  if (o0 < m) return 0;
  if (o0 + t > n - m) return n - t - 2 * m;
  return o0 - m;
}

// the lane's entries of input k's tile: LDS slot and element offset in a plane
template <typename R, int NORD, int MODE, int K>
struct PzTile {
  using G = PzGeo<R, NORD, pz_ty<MODE>()>;
  using N = PNeed<MODE, K>;
  static constexpr int NL = N::TILE ? G::nlane(N::X, N::Y) : 0;
  int slot[NL > 0 ? NL : 1], boff[NL > 0 ? NL : 1];
  R val[2][NL > 0 ? NL : 1];  // two planes in flight
  __device__ __forceinline__ void init(const PGeom& g, int X0, int Y0, int wx0, int wy0, int t) {
    constexpr int M = G::M;
    constexpr int NC = kZTX + (N::X ? 2 * M : 0), NLD = G::nload(N::X, N::Y);
    const int rx0 = N::X ? wx0 : X0, ry0 = N::Y ? wy0 : Y0;  // region origin
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      const int e = t + j * G::NT;
      const int gx = rx0 + e % NC, gy = ry0 + e / NC;
      const int sx = g.per0 ? pmod(gx, g.n0) : gx, sy = g.per1 ? pmod(gy, g.n1) : gy;
      // on a non-periodic dim only [0, n) is ever read
      const bool in = e < NLD && sx >= 0 && sx < g.n0 && sy >= 0 && sy < g.n1;
      slot[j] = in ? (gy - wy0) * G::TWP + (gx - wx0) : G::TSZ - 1;
      boff[j] = plane_boff<R>(in ? sx + g.n0 * sy : -1);
    }
  }
  __device__ __forceinline__ void load(int b, const R* u, long plane, int s2) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(u, plane, s2);
#pragma unroll
    for (int j = 0; j < NL; ++j) val[b][j] = plane_ld<R>(rs, boff[j]);
  }
  __device__ __forceinline__ void store(int b, R* T) const {
#pragma unroll
    for (int j = 0; j < NL; ++j) T[slot[j]] = val[b][j];
  }
};

// waves per SIMD asked of the compiler: 4 (<= 128 VGPRs; the FP64 form with
// cc at order 16 spills three values at that budget, on its 16-row tiles)
template <typename R, int MODE>
constexpr int pz_waves() { return 4; }

template <int MODE>
constexpr int pz_nt() { return kZTX * pz_ty<MODE>(); }  // lanes of a z-march block

template <typename R, int NORD, int MODE>
__global__ __launch_bounds__(pz_nt<MODE>()) __attribute__((amdgpu_waves_per_eu(pz_waves<R, MODE>()))) void poisson_zm_kernel(
    PGeom g, int tiles_x, int tiles, int zlo, int zhi, int zchunk, const R* __restrict__ u, R* __restrict__ o0,
    R* __restrict__ o1, GradEpi<R> ep, R rh0, R rh1, R rh2) {
  constexpr int TY = pz_ty<MODE>(), NT = pz_nt<MODE>();
  using G = PzGeo<R, NORD, TY>;
  constexpr int M = G::M, W = G::W, TWP = G::TWP, TSZ = G::TSZ;
  constexpr int NI = MODE >= GE_GRAD ? 1 : 3;
  __shared__ __align__(16) R lds[NI * TSZ];
  __shared__ R rows[W * W];  // every coefficient row (border lanes index it)
  for (int k = threadIdx.x; k < W * W; k += NT) rows[k] = prow<R, NORD>(k);
  const int lin = xcd_linear();
  const int tile = lin % tiles, zc = lin / tiles;
  const int X0 = (tile % tiles_x) * kZTX, Y0 = (tile / tiles_x) * TY;
  const int zb = zlo + zc * zchunk, ze = min(zb + zchunk, zhi);
  const int t = threadIdx.x, tx = t % kZTX, ty = t / kZTX;
  const int i0 = X0 + tx, i1 = Y0 + ty;
  const bool out_ok = i0 < g.n0 && i1 < g.n1;
  const long plane = (long)g.n0 * g.n1;
  const R* uz = u + (NI - 1) * g.np;  // the input taken along z
  const int wx0 = pz_window(X0, kZTX, M, g.n0, g.per0), wy0 = pz_window(Y0, TY, M, g.n1, g.per1);

  // the lane's stencils in the LDS window: first tap (row / column) and
  // coefficient row; border lanes of non-periodic dims take the shifted rows
  const int lr = i1 - wy0, lc = i0 - wx0;
  const bool bxl = !g.per0 && i0 < M, bxh = !g.per0 && i0 >= g.n0 - M;
  const bool byl = !g.per1 && i1 < M, byh = !g.per1 && i1 >= g.n1 - M;
  const int xs = (bxl ? 0 : bxh ? g.n0 - 1 - 2 * M : i0 - M) - wx0;
  const int ys = (byl ? 0 : byh ? g.n1 - 1 - 2 * M : i1 - M) - wy0;
  const int xrow = (bxl ? i0 : bxh ? i0 - g.n0 + 2 * M + 1 : M) * W;  // first tap of the row
  const int yrow = (byl ? i1 : byh ? i1 - g.n1 + 2 * M + 1 : M) * W;
  const bool bx = bxl || bxh, by = byl || byh;

  PzTile<R, NORD, MODE, 0> t0;
  PzTile<R, NORD, MODE, 1> t1;
  PzTile<R, NORD, MODE, 2> t2;
  t0.init(g, X0, Y0, wx0, wy0, t);
  if constexpr (NI == 3) {
    t1.init(g, X0, Y0, wx0, wy0, t);
    t2.init(g, X0, Y0, wx0, wy0, t);
  }
  FDim dz{};
  dz.n = g.n2; dz.nx = g.n2; dz.bc = g.per2 ? PER : FREE;
  ZCursor zc_tile(dz, zb), zc_col(dz, zb - M);
  auto load_tiles = [&](int b) {  // b static after unrolling
    const int s2 = zc_tile.src();
    t0.load(b, u, plane, s2);
    if constexpr (NI == 3) {
      t1.load(b, u + g.np, plane, s2);
      t2.load(b, u + 2 * g.np, plane, s2);
    }
    zc_tile.next();
  };
  const int cboff = plane_boff<R>(out_ok ? i0 + g.n0 * i1 : -1);
  auto col_next = [&]() -> R {
    const R v = plane_ld<R>(plane_rsrc(uz, plane, zc_col.src()), cboff);
    zc_col.next();
    return v;
  };
  // planes z + M (column) and z (tiles) are loaded two steps ahead of use
  R win[W];
#pragma unroll
  for (int k = 0; k < W - 1; ++k) win[k] = col_next();  // planes zb - M .. zb + M - 1
  R nxt[2];
  nxt[0] = col_next();
  nxt[1] = col_next();
  load_tiles(0);
  load_tiles(1);

  // x- / y-stencil of a tile at this lane (centred taps from scalar
  // registers; the shifted border rows per lane)
  auto sx = [&](const R* T) {
    const R* v = T + lr * TWP + xs;
    R s = R(0);
    if (!bx) {
#pragma unroll
      for (int j = 0; j < W; ++j) s = fmar(prow<R, NORD>(M * W + j), v[j], s);
    } else {
#pragma unroll
      for (int j = 0; j < W; ++j) s = fmar(rows[xrow + j], v[j], s);
    }
    return s;
  };
  auto sy = [&](const R* T) {
    const R* v = T + ys * TWP + lc;
    R s = R(0);
    if (!by) {
#pragma unroll
      for (int j = 0; j < W; ++j) s = fmar(prow<R, NORD>(M * W + j), v[j * TWP], s);
    } else {
#pragma unroll
      for (int j = 0; j < W; ++j) s = fmar(rows[yrow + j], v[j * TWP], s);
    }
    return s;
  };
  auto ctr = [&](const R* T) { return T[lr * TWP + lc]; };

  double sq = 0.0;  // update_rhopol: this lane's sum of res^2, z in order
  auto step = [&](int z, int b) {  // b = (z - zb) & 1, static after unrolling
    win[W - 1] = nxt[b];
    t0.store(b, lds);
    if constexpr (NI == 3) {
      t1.store(b, lds + TSZ);
      t2.store(b, lds + 2 * TSZ);
    }
    nxt[b] = col_next();
    load_tiles(b);
    __syncthreads();
    if (out_ok) {
      const long p = (long)z * plane + i0 + (long)g.n0 * i1;
      R dzv = R(0);
#pragma unroll
      for (int j = 0; j < W; ++j) dzv = fmar(prow<R, NORD>(M * W + j), win[j], dzv);
      if constexpr (MODE >= GE_GRAD) {
        const R dx = sx(lds) * rh0, dy = sy(lds) * rh1, dz2 = dzv * rh2;
        sq += grad_epilogue<R, MODE>(ep, g.np, p, dx, dy, dz2);
      } else {
        const R dx = sx(lds) * rh0;
        const R dy = sy(lds + TSZ) * rh1;
        const R dz2 = dzv * rh2;
        o0[p] = (dx + dy) + dz2;
        if constexpr (MODE == 1) {
          const R x = ctr(lds), y = ctr(lds + TSZ), zv = win[M];
          const R xy = sx(lds + TSZ) * rh0;
          const R xz = sx(lds + 2 * TSZ) * rh0;
          const R yz = sy(lds + 2 * TSZ) * rh1;
          o1[p] = (((x * x) * dx + (y * y) * dy) + (zv * zv) * dz2) +
                  R(2) * ((x * y * xy + x * zv * xz) + y * zv * yz);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < W - 1; ++k) win[k] = win[k + 1];
  };
  for (int z = zb; z < ze; z += 4) {  // four steps per trip: static buffers
    step(z, 0);
    if (z + 1 < ze) step(z + 1, 1);
    if (z + 2 < ze) step(z + 2, 0);
    if (z + 3 < ze) step(z + 3, 1);
  }
  if constexpr (MODE == GE_RHOPOL) {
    __shared__ double red[NT / 64];
    const double s = pblock_sum<NT>(sq, red);
    if (threadIdx.x == 0) ep.part[blockIdx.x] = s;
  }
}

PGeom pgeom(const int* n, int geocode, int bc_idim, int idim) {
  PGeom g;
  g.n0 = n[0]; g.n1 = n[1]; g.n2 = n[2];
  g.np = (long)n[0] * n[1] * n[2];
  g.z0 = 0;
  g.nz = n[2];
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

// 8 XCD shares of ceil(S / 8) segments times nz planes
unsigned pgrid(const PGeom& g) { return (unsigned)(8L * ((psegs(g) + 7) / 8) * g.nz); }

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

// z-march launch shape: tiles x z chunks of the planes zlo .. zhi - 1 (all,
// or on a non-periodic z the planes off its border).  The chunks put at most
// `resident` blocks (what one CU holds of the kernel) on any CU, so the grid
// runs in one even round, yet at least one block per CU where the grid
// allows; chunks no shorter than 2 stencil lengths.  False when the
// one-point kernels must do all (LIBCONV_POISSON_POINT set, planes beyond
// the buffer-record limit).
struct PzLaunch {
  int tiles_x, tiles, zlo, zhi, zchunk;
  unsigned blocks;
};
static bool pz_launch(const PGeom& g, int nord, int ty, int resident, PzLaunch& L) {
  if (getenv("LIBCONV_POISSON_POINT")) return false;
  if ((long)g.n0 * g.n1 * 8 > kPlaneMax) return false;
  const int m = nord / 2;
  L.zlo = g.per2 ? 0 : m;
  L.zhi = g.per2 ? g.n2 : g.n2 - m;
  const int nz = L.zhi - L.zlo;  // >= 1: n >= nord + 1
  L.tiles_x = (g.n0 + kZTX - 1) / kZTX;
  const long tiles = (long)L.tiles_x * ((g.n1 + ty - 1) / ty);
  long chunks = (long)std::max(1, resident) * num_cu() / tiles;
  if (chunks * tiles < num_cu()) chunks = (num_cu() + tiles - 1) / tiles;
  chunks = std::max(1L, std::min<long>(chunks, nz / (2 * (nord + 1))));
  L.zchunk = (int)((nz + chunks - 1) / chunks);
  const long nb = tiles * ((nz + L.zchunk - 1) / L.zchunk);
  if (nb > 0x7fffffffL) return false;
  L.tiles = (int)tiles;
  L.blocks = (unsigned)nb;
  return true;
}

// blocks of one z-march kernel that a CU holds (registers, LDS), cached
template <auto KER, int NT>
static int pz_resident() {
  static int nb = 0;
  if (!nb) {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, KER, NT, 0) != hipSuccess || v < 1) v = 1;
    nb = v;
  }
  return nb;
}

// one z-march launch; *done false when the one-point kernels must run
template <typename R, int NO, int MODE>
static hipError_t pz_run(const PGeom& g, const R* u, R* o0, R* o1, R r0, R r1, R r2, hipStream_t st,
                         bool* done) {
  PzLaunch L;
  *done = pz_launch(g, NO, pz_ty<MODE>(), pz_resident<poisson_zm_kernel<R, NO, MODE>, pz_nt<MODE>()>(), L);
  if (!*done) return hipSuccess;
  poisson_zm_kernel<R, NO, MODE><<<L.blocks, pz_nt<MODE>(), 0, st>>>(g, L.tiles_x, L.tiles, L.zlo, L.zhi, L.zchunk,
                                                                     u, o0, o1, GradEpi<R>{}, r0, r1, r2);
  return hipGetLastError();
}

// the planes of a non-periodic z the z march leaves out: 0 .. m - 1 and
// n - m .. n - 1, as two one-point launches (g.z0 / g.nz select the planes)
template <typename F>
static void pz_border_planes(const PGeom& g, int nord, F&& launch) {
  if (g.per2) return;
  PGeom b = g;
  b.nz = nord / 2;
  b.z0 = 0;
  launch(b);
  b.z0 = g.n2 - nord / 2;
  launch(b);
}

template <typename R>
hipError_t div_u_i(int geocode, const int* n, const R* u, R* du, R* cc, int nord, const double* h, hipStream_t st) {
  const PGeom g = pgeom(n, geocode, 0, -1);
  if (g.np == 0) return hipSuccess;
  const R r0 = (R)(1.0 / h[0]), r1 = (R)(1.0 / h[1]), r2 = (R)(1.0 / h[2]);
  auto point = [&](const PGeom& q) -> hipError_t {
    LC_POISSON_DISPATCH(nord, (div_u_i_kernel<R, NO><<<pgrid(q), kPThreads, 0, st>>>(q, psegs(q), u, du, cc, r0, r1, r2)))
    return hipGetLastError();
  };
  bool done = false;
  hipError_t e = hipSuccess;
  if (cc) {
    LC_POISSON_DISPATCH(nord, (e = pz_run<R, NO, 1>(g, u, du, cc, r0, r1, r2, st, &done)))
  } else {
    LC_POISSON_DISPATCH(nord, (e = pz_run<R, NO, 0>(g, u, du, cc, r0, r1, r2, st, &done)))
  }
  if (e != hipSuccess) return e;
  if (!done) return point(g);
  pz_border_planes(g, nord, [&](const PGeom& q) { if (e == hipSuccess) e = point(q); });
  return e;
}

// The launches of one gradient-family operator: the z march over the planes
// off a non-periodic z border and the one-point kernels on the border planes,
// or the one-point kernels on every plane.  Returns their block count, which
// is the number of partial sums update_rhopol writes (ep.part[0 .. count));
// dry: count only, nothing launched.
template <typename R, int NO, int MODE>
static long grad_launch(const PGeom& g, const R* u, const GradEpi<R>& ep, R r0, R r1, R r2, hipStream_t st,
                        bool dry) {
  long off = 0;
  auto point = [&](const PGeom& q) {
    const unsigned nb = pgrid(q);
    if (!dry) {
      GradEpi<R> e = ep;
      if (e.part) e.part += off;
      grad_point_kernel<R, NO, MODE><<<nb, kPThreads, 0, st>>>(q, psegs(q), u, e, r0, r1, r2);
    }
    off += nb;
  };
  PzLaunch L;
  if (!pz_launch(g, NO, pz_ty<MODE>(), pz_resident<poisson_zm_kernel<R, NO, MODE>, pz_nt<MODE>()>(), L)) {
    point(g);
    return off;
  }
  if (!dry)
    poisson_zm_kernel<R, NO, MODE><<<L.blocks, pz_nt<MODE>(), 0, st>>>(g, L.tiles_x, L.tiles, L.zlo, L.zhi,
                                                                       L.zchunk, u, nullptr, nullptr, ep, r0, r1, r2);
  off += L.blocks;
  pz_border_planes(g, NO, point);
  return off;
}

template <typename R, int MODE>
static hipError_t grad_op(int geocode, const int* n, const R* u, const GradEpi<R>& ep, int nord, const double* h,
                          hipStream_t st, long* blocks, bool dry) {
  const PGeom g = pgeom(n, geocode, 0, -1);
  *blocks = 0;
  if (g.np == 0) return hipSuccess;
  const R r0 = h ? (R)(1.0 / h[0]) : R(1), r1 = h ? (R)(1.0 / h[1]) : R(1), r2 = h ? (R)(1.0 / h[2]) : R(1);
  LC_POISSON_DISPATCH(nord, (*blocks = grad_launch<R, NO, MODE>(g, u, ep, r0, r1, r2, st, dry)))
  return dry ? hipSuccess : hipGetLastError();
}

template <typename R>
hipError_t nabla_u(int geocode, const int* n, const R* u, R* du, R* du2, int nord, const double* h, hipStream_t st) {
  GradEpi<R> ep{};
  ep.du = du;
  ep.du2 = du2;
  long nb;
  return grad_op<R, GE_GRAD>(geocode, n, u, ep, nord, h, st, &nb, false);
}

template <typename R>
hipError_t nabla_u_square(int geocode, const int* n, const R* u, R* ddu, int nord, const double* h, hipStream_t st) {
  GradEpi<R> ep{};
  ep.du2 = ddu;
  long nb;
  return grad_op<R, GE_SQUARE>(geocode, n, u, ep, nord, h, st, &nb, false);
}

template <typename R>
hipError_t nabla_u_epsilon(int geocode, const int* n, const R* u, R* du, const R* eps, int nord, const double* h,
                           hipStream_t st) {
  GradEpi<R> ep{};
  ep.du = du;
  ep.eps = eps;
  long nb;
  return grad_op<R, GE_EPS>(geocode, n, u, ep, nord, h, st, &nb, false);
}

template <typename R>
long update_rhopol_partials(int geocode, const int* n, int nord) {
  if (!poisson_order_ok(nord)) return -1;
  long nb = 0;
  (void)grad_op<R, GE_RHOPOL>(geocode, n, nullptr, GradEpi<R>{}, nord, nullptr, nullptr, &nb, true);
  return nb;
}

template <typename R>
hipError_t update_rhopol(int geocode, const int* n, const R* u, int nord, const double* h, R eta, const R* dlogeps,
                         R* rhopol, double* part, long nparts, double* rhores2, hipStream_t st) {
  GradEpi<R> ep{};
  ep.dlogeps = dlogeps;
  ep.rhopol = rhopol;
  ep.part = part;
  ep.eta = eta;
  long nb = 0;
  hipError_t e = grad_op<R, GE_RHOPOL>(geocode, n, u, ep, nord, h, st, &nb, true);
  if (e != hipSuccess) return e;
  if (nb > nparts) return hipErrorInvalidValue;  // the caller sized part by update_rhopol_partials
  e = grad_op<R, GE_RHOPOL>(geocode, n, u, ep, nord, h, st, &nb, false);
  if (e != hipSuccess) return e;
  return sum_partial_sets(part, nb, rhores2, 1, 0, st);
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
template hipError_t nabla_u_square<double>(int, const int*, const double*, double*, int, const double*, hipStream_t);
template hipError_t nabla_u_square<float>(int, const int*, const float*, float*, int, const double*, hipStream_t);
template hipError_t nabla_u_epsilon<double>(int, const int*, const double*, double*, const double*, int,
                                            const double*, hipStream_t);
template hipError_t nabla_u_epsilon<float>(int, const int*, const float*, float*, const float*, int, const double*,
                                           hipStream_t);
template long update_rhopol_partials<double>(int, const int*, int);
template long update_rhopol_partials<float>(int, const int*, int);
template hipError_t update_rhopol<double>(int, const int*, const double*, int, const double*, double, const double*,
                                          double*, double*, long, double*, hipStream_t);
template hipError_t update_rhopol<float>(int, const int*, const float*, int, const double*, float, const float*,
                                         float*, double*, long, double*, hipStream_t);

}  // namespace lc

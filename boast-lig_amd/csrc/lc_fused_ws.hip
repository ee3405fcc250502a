// lc_fused_ws.hip -- wave-specialised single-pass 3D libconv operators.
//
// Same arithmetic as lc_fused.hip (x~ wrapped / zero-extended per dim, the
// BigDFT/GenericConvolution2.rb:91-121 index rules), organised for
// occupancy: every block splits its waves by role so that no lane has to
// hold both a long window along z and an x/y filter window:
//
//   z-waves  one output column per lane (64 columns per wave).  A shifting
//            register window of the column's values along z gives the z
//            filter; the lane also combines the terms, applies the epilogue
//            and stores the output.
//   x-waves  QX consecutive outputs of one tile row per lane, from the LDS
//            tile (x filter).
//   y-waves  QY consecutive outputs of one tile column per lane, from LDS
//            (y filter).
//
//   sum3d_ws  y = a0 Fx x~ + a1 Fy x~ + a2 Fz x~ + a_x x~ + a_y y (kinetic)
//             with the per-direction FP64 energies; per output plane the
//             x-/y-waves stage the cross-shaped halo of x~(z) in LDS, the
//             z-waves write the tile centre from their windows.
//   sep3d_ws  y = a Fz Fy Fx x~ + a_y y (magic filter); per input plane all
//             waves stage the (TX+W-1) x (TY+W-1) tile, x-waves filter it
//             into B, y-waves filter B into the plane P, z-waves append P to
//             their windows one step later and emit the output plane.
// Two barriers per plane; LDS row strides are odd (conflict-free columns).
#include <stdlib.h>

#include <algorithm>

#include "lc_kernels.hpp"

namespace lc {

struct WDim {
  int n, bc, nx, ny, ox, oy, olo, nout;
};

__device__ __forceinline__ int wsrc(const WDim& g, int j) {
  if (g.bc == PER) return pmod(j, g.n);
  if (g.bc == GROW || g.bc == FREE) return (j >= 0 && j < g.n) ? j : -1;
  const int s = j - g.ox;
  return (s >= 0 && s < g.nx) ? s : -1;
}

template <typename R>
struct WsArgs {
  const R* x;
  R* y;
  WDim d0, d1, d2;
  long nti, nto;
  int narr, tiles_x, zchunk;
  R a0, a1, a2, a_x, a_y;
  int flags;
  double* part;
  long nparts;
};

// Block -> (tile, z chunk, array), XCD-contiguous (see lc_fused.hip).
__device__ __forceinline__ void ws_coords(int tiles_x, int& bx, int& by, int& zc, int& arr) {
  const int tiles = gridDim.x, nz = gridDim.y;
  const int nb = tiles * nz * gridDim.z;
  const int lin = blockIdx.x + tiles * (blockIdx.y + nz * blockIdx.z);
  const int xcd = lin & 7, k = lin >> 3;
  const int per = nb >> 3, rem = nb & 7;
  const int nl = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
  const int tile = nl % tiles;
  const int rest = nl / tiles;
  zc = rest % nz;
  arr = rest / nz;
  bx = tile % tiles_x;
  by = tile / tiles_x;
}

// sum of W taps from a register window, one FMA chain in tap order (the
// order of every other variant, so that all variants give identical bits)
template <typename R, class F>
__device__ __forceinline__ R ztaps(const R* w) {
  R s = R(0);
#pragma unroll
  for (int k = 0; k < F::N; ++k) s = fmar(w[k], tap<R, F>(k), s);
  return s;
}

// ===================================================================== sum3d
template <typename R, class F, int TX, int TY, int QX, int QY, int HD>
struct SumWs {
  static constexpr int W = F::N, HL = -F::LO, HR = F::UP;
  static constexpr int NCOL = TX * TY;
  static constexpr int NZW = NCOL / 64;
  static constexpr int NXT = TY * (TX / QX), NYT = TX * (TY / QY);
  static constexpr int NXW = (NXT + 63) / 64, NYW = (NYT + 63) / 64;
  static constexpr int NT = 64 * (NZW + NXW + NYW);
  static constexpr int TW = (TX + HL + HR) | 1, TH = TY + HL + HR;
  static constexpr int SW = TX | 1;
  static constexpr int NH = (TX + TY) * (HL + HR);
  static constexpr int NTT = 64 * (NXW + NYW);
  static constexpr int NHT = (NH + NTT - 1) / NTT;
  static_assert(NCOL % 64 == 0 && TX % QX == 0 && TY % QY == 0, "tile shape");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int HD>
__global__ __launch_bounds__((SumWs<R, F, TX, TY, QX, QY, HD>::NT)) void sum3d_ws(WsArgs<R> p) {
  using G = SumWs<R, F, TX, TY, QX, QY, HD>;
  constexpr int W = G::W, HL = G::HL, TW = G::TW, SW = G::SW, NZW = G::NZW, NXW = G::NXW;
  constexpr int NHT = G::NHT;
  __shared__ R T[G::TH * TW];
  __shared__ R SX[TY * SW];
  __shared__ R SY[TY * SW];
  __shared__ double red[G::NT / 64];

  const int t = threadIdx.x, wave = t >> 6;
  int bx, by, zc, arr;
  ws_coords(p.tiles_x, bx, by, zc, arr);
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;
  const bool is_z = wave < NZW;
  const bool is_x = !is_z && wave < NZW + NXW;
  const bool want_dot = (p.flags & F_DOT) && arr == p.narr - 1;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;

  auto plane = [&](int z, bool& ok) -> const R* {
    const int s2 = wsrc(p.d2, z);
    ok = s2 >= 0;
    return xa + (ok ? (long)s2 * sxy : 0);
  };

  if (is_z) {
    // ------------------------------------------------------------ z-wave
    const int col = t;  // waves 0..NZW-1
    const int tx = col % TX, ty = col / TX;
    const int s0 = wsrc(p.d0, X0 + tx), s1 = wsrc(p.d1, Y0 + ty);
    const int coff = (s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
    const bool out_ok = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);
    auto ld = [&](int z) -> R {
      bool ok;
      const R* b = plane(z, ok);
      return (ok && coff >= 0) ? b[coff] : R(0);
    };
    R win[W];
    R q[HD];
#pragma unroll
    for (int s = 0; s < W - 1; ++s) win[s] = ld(zb + F::LO + s);
#pragma unroll
    for (int h = 0; h < HD; ++h) q[h] = ld(zb + F::UP + h);
    R* const yp0 = ya + (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
    const long ys = (long)p.d0.ny * p.d1.ny;
    for (int z = zb; z < ze; ++z) {
      win[W - 1] = q[0];
#pragma unroll
      for (int h = 0; h < HD - 1; ++h) q[h] = q[h + 1];
      q[HD - 1] = ld(z + F::UP + HD);
      T[(ty + HL) * TW + tx + HL] = win[HL];
      __syncthreads();
      const R zt = ztaps<R, F>(win);
      __syncthreads();
      const R xc = win[HL];
      const bool has_x = out_ok && coff >= 0 && wsrc(p.d2, z) >= 0;
      R t0 = SX[ty * SW + tx] * p.a0;
      const R t1 = SY[ty * SW + tx] * p.a1;
      const R t2 = zt * p.a2;
      if (want_dot && has_x) {
        d0 = fmar((double)xc, (double)t0, d0);
        d1 = fmar((double)xc, (double)t1, d1);
        d2 = fmar((double)xc, (double)t2, d2);
      }
      if ((p.flags & F_AX) && has_x) t0 = fmar(xc, p.a_x, t0);
      R tot = t1 + t0;
      tot = t2 + tot;
      if (out_ok) {
        R* yp = yp0 + ys * (z - p.d2.oy);
        if (p.flags & F_AY) tot = fmar(p.a_y, *yp, tot);
        *yp = tot;
      }
#pragma unroll
      for (int s = 0; s < W - 1; ++s) win[s] = win[s + 1];
    }
  } else {
    // ------------------------------------------------------------ x/y-waves
    const int tt = t - 64 * NZW;  // 0 .. NTT-1
    int hslot[NHT], hoff[NHT];
#pragma unroll
    for (int k = 0; k < NHT; ++k) {
      const int e = tt + k * G::NTT;
      int r = 0, cc = 0;
      const bool valid = e < G::NH;
      if (e < TY * (HL + G::HR)) {
        r = e / (HL + G::HR);
        const int qq = e % (HL + G::HR);
        cc = qq < HL ? qq - HL : TX + (qq - HL);
      } else if (valid) {
        const int e2 = e - TY * (HL + G::HR);
        const int rr = e2 / TX;
        cc = e2 % TX;
        r = rr < HL ? rr - HL : TY + (rr - HL);
      }
      hslot[k] = valid ? (r + HL) * TW + (cc + HL) : -1;
      const int s0 = wsrc(p.d0, X0 + cc), s1 = wsrc(p.d1, Y0 + r);
      hoff[k] = (valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
    }
    auto ldh = [&](int z, R* dst) {
      bool ok;
      const R* b = plane(z, ok);
#pragma unroll
      for (int k = 0; k < NHT; ++k) dst[k] = (ok && hoff[k] >= 0) ? b[hoff[k]] : R(0);
    };
    R hq[HD][NHT];
#pragma unroll
    for (int h = 0; h < HD; ++h) ldh(zb + h, hq[h]);
    for (int z = zb; z < ze; ++z) {
#pragma unroll
      for (int k = 0; k < NHT; ++k)
        if (hslot[k] >= 0) T[hslot[k]] = hq[0][k];
#pragma unroll
      for (int h = 0; h < HD - 1; ++h)
#pragma unroll
        for (int k = 0; k < NHT; ++k) hq[h][k] = hq[h + 1][k];
      ldh(z + HD, hq[HD - 1]);
      __syncthreads();
      if (is_x) {
        const int task = tt;
        if (task < G::NXT) {
          const int r = task % TY, c0 = (task / TY) * QX;
          const R* row = T + (r + HL) * TW + c0;
          R acc[QX];
#pragma unroll
          for (int q = 0; q < QX; ++q) acc[q] = R(0);
#pragma unroll
          for (int k = 0; k < QX + W - 1; ++k) {
            const R v = row[k];
#pragma unroll
            for (int q = 0; q < QX; ++q)
              if (k - q >= 0 && k - q < W) acc[q] = fmar(v, tap<R, F>(k - q), acc[q]);
          }
#pragma unroll
          for (int q = 0; q < QX; ++q) SX[r * SW + c0 + q] = acc[q];
        }
      } else {
        const int task = tt - 64 * NXW;
        if (task < G::NYT) {
          const int c = task % TX, r0 = (task / TX) * QY;
          const R* colp = T + r0 * TW + c + HL;
          R acc[QY];
#pragma unroll
          for (int q = 0; q < QY; ++q) acc[q] = R(0);
#pragma unroll
          for (int k = 0; k < QY + W - 1; ++k) {
            const R v = colp[k * TW];
#pragma unroll
            for (int q = 0; q < QY; ++q)
              if (k - q >= 0 && k - q < W) acc[q] = fmar(v, tap<R, F>(k - q), acc[q]);
          }
#pragma unroll
          for (int q = 0; q < QY; ++q) SY[(r0 + q) * SW + c] = acc[q];
        }
      }
      __syncthreads();
    }
  }
  if (p.part) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      d0 += __shfl_xor(d0, o, 64);
      d1 += __shfl_xor(d1, o, 64);
      d2 += __shfl_xor(d2, o, 64);
    }
    const long bid = (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc;
    double* sr = red;
    for (int k = 0; k < 3; ++k) {
      __syncthreads();
      if ((t & 63) == 0) sr[wave] = k == 0 ? d0 : (k == 1 ? d1 : d2);
      __syncthreads();
      if (t == 0 && arr == p.narr - 1) {
        double s = 0.0;
        for (int w = 0; w < G::NT / 64; ++w) s += sr[w];
        p.part[k * p.nparts + bid] = s;
      }
    }
  }
}

// ===================================================================== sep3d
template <typename R, class F, int TX, int TY, int QX, int QY, int AD>
struct SepWs {
  static constexpr int W = F::N;
  static constexpr int TWD = TX + W - 1, TH = TY + W - 1;
  static constexpr int TWP = TWD | 1, BW = TX | 1;
  static constexpr int NA = TH * TWD;
  static constexpr int NCOL = TX * TY;
  static constexpr int NZW = NCOL / 64;
  static constexpr int NXT = TH * (TX / QX), NYT = TX * (TY / QY);
  static constexpr int NXW = (NXT + 63) / 64, NYW = (NYT + 63) / 64;
  static constexpr int NT = 64 * (NZW + NXW + NYW);
  static constexpr int NAT = (NA + NT - 1) / NT;
  static_assert(NCOL % 64 == 0 && TX % QX == 0 && TY % QY == 0, "tile shape");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int AD>
__global__ __launch_bounds__((SepWs<R, F, TX, TY, QX, QY, AD>::NT)) void sep3d_ws(WsArgs<R> p) {
  using G = SepWs<R, F, TX, TY, QX, QY, AD>;
  constexpr int W = F::N, TWD = G::TWD, TWP = G::TWP, BW = G::BW, NZW = G::NZW, NXW = G::NXW;
  constexpr int NAT = G::NAT, NT = G::NT;
  __shared__ R A[G::TH * TWP];
  __shared__ R B[G::TH * BW];
  __shared__ R P[TY * BW];

  const int t = threadIdx.x, wave = t >> 6;
  int bx, by, zc, arr;
  ws_coords(p.tiles_x, bx, by, zc, arr);
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;
  const int pfirst = zb + F::LO;
  const int np = (ze - zb) + W - 1;  // input planes of this chunk
  const bool is_z = wave < NZW;
  const bool is_x = !is_z && wave < NZW + NXW;

  // every lane stages NAT elements of each input tile (AD planes ahead)
  int aoff[NAT], aslot[NAT];
#pragma unroll
  for (int k = 0; k < NAT; ++k) {
    const int e = t + k * NT;
    const int s0 = wsrc(p.d0, X0 + F::LO + e % TWD), s1 = wsrc(p.d1, Y0 + F::LO + e / TWD);
    aoff[k] = (e < G::NA && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
    aslot[k] = e < G::NA ? (e / TWD) * TWP + e % TWD : -1;
  }
  auto lda = [&](int s, R* dst) {  // tile of input plane number s
    const int s2 = s < np ? wsrc(p.d2, pfirst + s) : -1;
    const R* b = xa + (s2 >= 0 ? (long)s2 * sxy : 0);
#pragma unroll
    for (int k = 0; k < NAT; ++k) dst[k] = (s2 >= 0 && aoff[k] >= 0) ? b[aoff[k]] : R(0);
  };
  R aq[AD][NAT];
#pragma unroll
  for (int h = 0; h < AD; ++h) lda(h, aq[h]);

  // z-wave state
  const int col = t, tx = col % TX, ty = col / TX;
  const bool out_ok = is_z && (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);
  R* const yp0 = ya + (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
  const long ys = (long)p.d0.ny * p.d1.ny;
  R win[W];
#pragma unroll
  for (int s = 0; s < W; ++s) win[s] = R(0);

  for (int s = 0; s <= np; ++s) {
    if (s < np) {
#pragma unroll
      for (int k = 0; k < NAT; ++k)
        if (aslot[k] >= 0) A[aslot[k]] = aq[0][k];
#pragma unroll
      for (int h = 0; h < AD - 1; ++h)
#pragma unroll
        for (int k = 0; k < NAT; ++k) aq[h][k] = aq[h + 1][k];
      lda(s + AD, aq[AD - 1]);
    }
    __syncthreads();
    if (is_x) {
      if (s < np) {
        const int task = t - 64 * NZW;
        if (task < G::NXT) {
          const int r = task % G::TH, c0 = (task / G::TH) * QX;
          const R* row = A + r * TWP + c0;
          R acc[QX];
#pragma unroll
          for (int q = 0; q < QX; ++q) acc[q] = R(0);
#pragma unroll
          for (int k = 0; k < QX + W - 1; ++k) {
            const R v = row[k];
#pragma unroll
            for (int q = 0; q < QX; ++q)
              if (k - q >= 0 && k - q < W) acc[q] = fmar(v, tap<R, F>(k - q), acc[q]);
          }
#pragma unroll
          for (int q = 0; q < QX; ++q) B[r * BW + c0 + q] = acc[q];
        }
      }
    } else if (is_z) {
      if (s >= 1) {
        // plane s-1 of the chunk, filtered along x and y, enters the window
#pragma unroll
        for (int k = 0; k < W - 1; ++k) win[k] = win[k + 1];
        win[W - 1] = P[ty * BW + tx];
        const int z = zb + (s - 1) - (W - 1);
        if (z >= zb && out_ok) {
          R sum = ztaps<R, F>(win);
          if (p.flags & F_A) sum = sum * p.a0;
          R* yp = yp0 + ys * (z - p.d2.oy);
          if (p.flags & F_AY) sum = fmar(p.a_y, *yp, sum);
          *yp = sum;
        }
      }
    }
    __syncthreads();
    if (!is_x && !is_z && s < np) {
      const int task = t - 64 * (NZW + NXW);
      if (task < G::NYT) {
        const int c = task % TX, r0 = (task / TX) * QY;
        const R* colp = B + r0 * BW + c;
        R acc[QY];
#pragma unroll
        for (int q = 0; q < QY; ++q) acc[q] = R(0);
#pragma unroll
        for (int k = 0; k < QY + W - 1; ++k) {
          const R v = colp[k * BW];
#pragma unroll
          for (int q = 0; q < QY; ++q)
            if (k - q >= 0 && k - q < W) acc[q] = fmar(v, tap<R, F>(k - q), acc[q]);
        }
#pragma unroll
        for (int q = 0; q < QY; ++q) P[(r0 + q) * BW + c] = acc[q];
      }
    }
  }
  (void)TWD;
}

// ===================================================================== launch
static WDim wdim(int n, int bc, int nx, int ny, int lo, int up) {
  WDim g;
  g.n = n; g.bc = bc; g.nx = nx; g.ny = ny;
  g.ox = bc == SHRINK ? lo : 0;
  g.oy = bc == GROW ? -up : 0;
  g.olo = bc == GROW ? -up : 0;
  g.nout = bc == GROW ? n + (up - lo) : n;
  return g;
}

static int ws_ncu() {
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess)
      ncu = pr.multiProcessorCount;
    if (ncu <= 0) ncu = 256;
  }
  return ncu;
}

// z chunks: about `per_cu` blocks per CU, chunks no shorter than `minz`
static int ws_zchunk(long tiles, int nz, int narr, int per_cu, int minz) {
  const long want = (long)per_cu * ws_ncu();
  long chunks = (want + tiles * narr - 1) / (tiles * narr);
  chunks = std::min<long>(chunks, std::max(1, nz / std::max(1, minz)));
  chunks = std::max<long>(chunks, 1);
  return (int)((nz + chunks - 1) / chunks);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int HD>
hipError_t launch_sum3d_ws(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                           double* dot, hipStream_t st, double* scratch, size_t sb, int per_cu,
                           int minz) {
  using G = SumWs<R, F, TX, TY, QX, QY, HD>;
  WsArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = wdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = wdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = wdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  const long tiles = (long)p.tiles_x * ((p.d1.nout + TY - 1) / TY);
  p.zchunk = ws_zchunk(tiles, p.d2.nout, std::max(narr, 1), per_cu, minz);
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
  p.a_x = a_x; p.a_y = a_y;
  p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
  p.nparts = tiles * zch;
  if (dot) {
    if (sb < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
    p.part = scratch;
  }
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sum3d_ws<R, F, TX, TY, QX, QY, HD><<<grid, G::NT, 0, st>>>(p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int AD>
hipError_t launch_sep3d_ws(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                           int per_cu, int minz) {
  using G = SepWs<R, F, TX, TY, QX, QY, AD>;
  WsArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = wdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = wdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = wdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  const long tiles = (long)p.tiles_x * ((p.d1.nout + TY - 1) / TY);
  p.zchunk = ws_zchunk(tiles, p.d2.nout, std::max(narr, 1), per_cu, minz);
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = a; p.a_y = a_y;
  p.flags = (a != R(1) ? F_A : 0) | (a_y != R(0) ? F_AY : 0);
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sep3d_ws<R, F, TX, TY, QX, QY, AD><<<grid, G::NT, 0, st>>>(p);
  return hipGetLastError();
}

size_t ws_sum3d_parts_bound(const Box3& b, int lo, int up) {
  WDim g0 = wdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], lo, up);
  WDim g1 = wdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], lo, up);
  WDim g2 = wdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], lo, up);
  // smallest tile 16 x 8, at most one chunk per output plane
  return (size_t)((g0.nout + 15) / 16) * ((g1.nout + 7) / 8) * (size_t)std::max(1, g2.nout);
}

#define LC_SUMWS(R, F, TX, TY, QX, QY, HD)                                                       \
  template hipError_t launch_sum3d_ws<R, F, TX, TY, QX, QY, HD>(                                 \
      const Box3&, int, const R*, R*, const double*, R, R, double*, hipStream_t, double*, size_t, \
      int, int);
#define LC_SEPWS(R, F, TX, TY, QX, QY, AD)                                                       \
  template hipError_t launch_sep3d_ws<R, F, TX, TY, QX, QY, AD>(const Box3&, int, const R*, R*, R, \
                                                                 R, hipStream_t, int, int);
#define LC_WS_ALL(R, F)               \
  LC_SUMWS(R, F, 32, 8, 4, 4, 2)      \
  LC_SUMWS(R, F, 64, 8, 4, 4, 2)      \
  LC_SUMWS(R, F, 32, 16, 4, 4, 2)     \
  LC_SEPWS(R, F, 32, 8, 4, 4, 2)      \
  LC_SEPWS(R, F, 64, 8, 4, 4, 2)      \
  LC_SEPWS(R, F, 32, 16, 4, 4, 2)
LC_WS_ALL(double, FD2)
LC_WS_ALL(double, FD1)
LC_WS_ALL(double, FMD)
LC_WS_ALL(double, FIMD)
LC_WS_ALL(float, FD2)
LC_WS_ALL(float, FD1)
LC_WS_ALL(float, FMD)
LC_WS_ALL(float, FIMD)

}  // namespace lc

// lc_body_conv.hpp -- device bodies of the single-pass 3D convolution
// operators (internal; lc_fused.hip describes the algorithms).  A body
// computes the outputs of one block (xy tile, z chunk, array) and takes its
// LDS from the caller, so the single-call, batched and persistent kernels
// share it.
#pragma once

#include "lc_fused.hpp"

namespace lc {

template <typename R>
__device__ __forceinline__ double wave_block_sum(double v, double* red, int nthreads) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (nthreads + 63) / 64; ++w) s += red[w];
  return s;
}

// ====================================================================== sum3d
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD>
constexpr size_t sum3d_lds_bytes() {
  return lds_align(sizeof(R) * 2 * (TY - F::LO + F::UP) * ((TX - F::LO + F::UP) | 1)) +
         lds_align(sizeof(R) * TY * (TX | 1)) + lds_align(sizeof(double) * (TX * TY / CPT / 64 + 1));
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD>
__device__ __forceinline__ void sum3d_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                           long bid, unsigned char* lds) {
  constexpr int NT = TX * TY / CPT;
  constexpr int W = F::N, HL = -F::LO, HR = F::UP;
  constexpr int TW = (TX + HL + HR) | 1;    // LDS tile row stride (odd: no bank conflicts)
  constexpr int SXW = TX | 1;                // x-term row stride
  constexpr int TH = TY + HL + HR;          // LDS tile rows
  constexpr int WS = W - 1 + BZ;            // z window per column
  constexpr int NH = (TX + TY) * (HL + HR); // halo elements per plane
  constexpr int NHT = (NH + NT - 1) / NT;   // halo elements per thread
  constexpr int NXT = TY * (TX / QX);       // x-tasks
  static_assert(TY % CPT == 0 && TX % QX == 0 && BZ % HD == 0, "tile shape");
  // LDS: double-buffered plane tile, x-terms, reduction slots
  R (*TB)[TH * TW] = reinterpret_cast<R (*)[TH * TW]>(lds);
  R* SX = reinterpret_cast<R*>(lds + lds_align(sizeof(R) * 2 * TH * TW));
  double* red = reinterpret_cast<double*>(lds + lds_align(sizeof(R) * 2 * TH * TW) +
                                          lds_align(sizeof(R) * TY * SXW));

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;

  // ---- owner columns: (tx, ty0 + c)
  const int tx = t % TX, ty0 = (t / TX) * CPT;
  int coff[CPT];   // xy offset of the column in x (or -1 when x~ is zero)
  bool cout_ok[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int s0 = fsrc(p.d0, X0 + tx), s1 = fsrc(p.d1, Y0 + ty0 + c);
    coff[c] = (s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
    cout_ok[c] = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty0 + c < p.d1.olo + p.d1.nout);
  }
  // ---- halo elements of this thread: LDS slot and xy offset
  int hslot[NHT], hoff[NHT];
#pragma unroll
  for (int k = 0; k < NHT; ++k) {
    const int e = t + k * NT;
    int r = 0, cc = 0;
    bool valid = e < NH;
    if (e < TY * (HL + HR)) {
      r = e / (HL + HR);
      const int q = e % (HL + HR);
      cc = q < HL ? q - HL : TX + (q - HL);
    } else if (valid) {
      const int e2 = e - TY * (HL + HR);
      const int rr = e2 / TX;
      cc = e2 % TX;
      r = rr < HL ? rr - HL : TY + (rr - HL);
    }
    hslot[k] = valid ? (r + HL) * TW + (cc + HL) : -1;
    const int s0 = fsrc(p.d0, X0 + cc), s1 = fsrc(p.d1, Y0 + r);
    hoff[k] = (valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
  }

  auto plane_base = [&](int z, bool& ok) -> const R* {
    const int s2 = fsrc(p.d2, z);
    ok = s2 >= 0;
    return xa + (ok ? (long)s2 * sxy : 0);
  };
  auto load_cols = [&](int z, R* dst) {
    bool ok;
    const R* b = plane_base(z, ok);
#pragma unroll
    for (int c = 0; c < CPT; ++c) dst[c] = (ok && coff[c] >= 0) ? b[coff[c]] : R(0);
  };
  auto load_halo = [&](int z, R* dst) {
    bool ok;
    const R* b = plane_base(z, ok);
#pragma unroll
    for (int k = 0; k < NHT; ++k) dst[k] = (ok && hoff[k] >= 0) ? b[hoff[k]] : R(0);
  };

  R win[CPT][WS];
  R stg[CPT][BZ];
  R hreg[HD][NHT];   // halo of planes z .. z+HD-1 in flight
  // prefill the window with planes zb+LO .. zb+UP+BZ-1
#pragma unroll
  for (int s = 0; s < WS; ++s) {
    R v[CPT];
    load_cols(zb + F::LO + s, v);
#pragma unroll
    for (int c = 0; c < CPT; ++c) win[c][s] = v[c];
  }
#pragma unroll
  for (int h = 0; h < HD; ++h) load_halo(zb + h, hreg[h]);
  const bool want_dot = (p.flags & F_DOT) && arr == p.narr - 1;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;

  for (int z0 = zb; z0 < ze; z0 += BZ) {
    // stage the next block's new planes (z0+UP+BZ .. z0+UP+2BZ-1)
#pragma unroll
    for (int u = 0; u < BZ; ++u) {
      R v[CPT];
      load_cols(z0 + F::UP + BZ + u, v);
#pragma unroll
      for (int c = 0; c < CPT; ++c) stg[c][u] = v[c];
    }
#pragma unroll
    for (int u = 0; u < BZ; ++u) {
      const int z = z0 + u;
      R* const T = TB[(z - zb) & 1];
      if (z < ze) {
        // ---- tile of plane z: centre from the window, halo from registers
#pragma unroll
        for (int c = 0; c < CPT; ++c) T[(ty0 + c + HL) * TW + tx + HL] = win[c][u + HL];
#pragma unroll
        for (int k = 0; k < NHT; ++k)
          if (hslot[k] >= 0) T[hslot[k]] = hreg[u % HD][k];
        if (z + HD < ze) load_halo(z + HD, hreg[u % HD]);
        __syncthreads();
        // ---- x-terms
        if (t < NXT) {
          // consecutive lanes on consecutive rows: odd strides, no conflicts
          const int r = t % TY, c0 = (t / TY) * QX;
          R wv[QX + W - 1];
#pragma unroll
          for (int k = 0; k < QX + W - 1; ++k) wv[k] = T[(r + HL) * TW + c0 + k];
#pragma unroll
          for (int q = 0; q < QX; ++q) {
            R s = R(0);
#pragma unroll
            for (int k = 0; k < W; ++k) s = fmar(wv[q + k], tap<R, F>(k), s);
            SX[r * SXW + c0 + q] = s;
          }
        }
        __syncthreads();
        // ---- y-terms and z-terms of the owned columns (the tile is double
        // buffered, so it stays valid until the next plane's first barrier)
        R yt[CPT], zt[CPT];
        {
          R wv[CPT + W - 1];
#pragma unroll
          for (int k = 0; k < CPT + W - 1; ++k) wv[k] = T[(ty0 + k) * TW + tx + HL];
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            R s = R(0);
#pragma unroll
            for (int k = 0; k < W; ++k) s = fmar(wv[c + k], tap<R, F>(k), s);
            yt[c] = s;
          }
        }
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          R s = R(0);
#pragma unroll
          for (int k = 0; k < W; ++k) s = fmar(win[c][u + k], tap<R, F>(k), s);
          zt[c] = s;
        }
        // ---- combine
        const bool zin = fsrc(p.d2, z) >= 0;
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          const R xc = win[c][u + HL];
          const bool has_x = zin && coff[c] >= 0 && cout_ok[c];  // output points only
          R t0 = SX[(ty0 + c) * SXW + tx] * p.a0;
          const R t1 = yt[c] * p.a1;
          const R t2 = zt[c] * p.a2;
          if (want_dot && has_x) {
            d0 = fmar((double)xc, (double)t0, d0);
            d1 = fmar((double)xc, (double)t1, d1);
            d2 = fmar((double)xc, (double)t2, d2);
          }
          if ((p.flags & F_AX) && has_x) t0 = fmar(xc, p.a_x, t0);
          R tot = t1 + t0;
          tot = t2 + tot;
          if (cout_ok[c]) {
            R* yp = ya + (X0 + tx - p.d0.oy) +
                    (long)p.d0.ny * ((Y0 + ty0 + c - p.d1.oy) + (long)p.d1.ny * (z - p.d2.oy));
            if (p.flags & F_AY) tot = fmar(p.a_y, *yp, tot);
            *yp = tot;
          }
        }
      }
    }
    // ---- slide the window by BZ planes
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
#pragma unroll
      for (int s = 0; s < W - 1; ++s) win[c][s] = win[c][s + BZ];
#pragma unroll
      for (int u = 0; u < BZ; ++u) win[c][W - 1 + u] = stg[c][u];
    }
  }
  if (p.part) {
    double s = wave_block_sum<R>(d0, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[bid] = s;
    s = wave_block_sum<R>(d1, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[p.nparts + bid] = s;
    s = wave_block_sum<R>(d2, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[2 * p.nparts + bid] = s;
  }
}

// ====================================================================== sep3d
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD>
constexpr size_t sep3d_lds_bytes() {
  return lds_align(sizeof(R) * (TY + F::N - 1) * ((TX + F::N - 1) | 1)) +
         lds_align(sizeof(R) * (TY + F::N - 1) * (TX | 1));
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD>
__device__ __forceinline__ void sep3d_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                           unsigned char* lds) {
  constexpr int NT = TX * TY / CPT;
  constexpr int W = F::N;
  constexpr int TW = TX + W - 1;            // input tile row length
  constexpr int TWP = TW | 1;               // its LDS stride (odd)
  constexpr int BW = TX | 1;                // x-filtered row stride (odd)
  constexpr int TH = TY + W - 1;            // input tile rows
  constexpr int WS = W - 1 + BZ;
  constexpr int NA = TH * TW;               // input tile elements
  constexpr int NAT = (NA + NT - 1) / NT;
  constexpr int NXT = TH * (TX / QX);       // x-tasks
  static_assert(TY % CPT == 0 && TX % QX == 0, "tile shape");
  R* A = reinterpret_cast<R*>(lds);                                   // input tile
  R* B = reinterpret_cast<R*>(lds + lds_align(sizeof(R) * TH * TWP));  // x-filtered tile

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;

  const int tx = t % TX, ty0 = (t / TX) * CPT;
  bool cout_ok[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c)
    cout_ok[c] = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty0 + c < p.d1.olo + p.d1.nout);
  // input tile element e -> logical (X0 + LO + e % TW, Y0 + LO + e / TW)
  int aoff[NAT], aslot[NAT];
#pragma unroll
  for (int k = 0; k < NAT; ++k) {
    const int e = t + k * NT;
    const int s0 = fsrc(p.d0, X0 + F::LO + e % TW), s1 = fsrc(p.d1, Y0 + F::LO + e / TW);
    aoff[k] = (e < NA && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1;
    aslot[k] = e < NA ? (e / TW) * TWP + e % TW : -1;
  }
  auto load_tile = [&](int z, R* dst) {
    const int s2 = fsrc(p.d2, z);
    const R* b = xa + (s2 >= 0 ? (long)s2 * sxy : 0);
#pragma unroll
    for (int k = 0; k < NAT; ++k) dst[k] = (s2 >= 0 && aoff[k] >= 0) ? b[aoff[k]] : R(0);
  };

  R win[CPT][WS];
  R areg[AD][NAT];  // input tiles of the next AD planes in flight
  const int plast = ze - 1 + F::UP;  // last input plane any output needs
  const int pfirst = zb + F::LO;
#pragma unroll
  for (int h = 0; h < AD; ++h)
    if (pfirst + h <= plast) load_tile(pfirst + h, areg[h]);
  // x and y passes of input plane `plane` (its tile in areg[slot]) into pv
  auto filter_plane = [&](int plane, R* ar, R* pv) {
#pragma unroll
    for (int k = 0; k < NAT; ++k)
      if (aslot[k] >= 0) A[aslot[k]] = ar[k];
    if (plane + AD <= plast) load_tile(plane + AD, ar);
    __syncthreads();
    for (int task = t; task < NXT; task += NT) {
      // consecutive lanes on consecutive rows: odd strides, no bank conflicts
      const int r = task % TH, c0 = (task / TH) * QX;
      R wv[QX + W - 1];
#pragma unroll
      for (int k = 0; k < QX + W - 1; ++k) wv[k] = A[r * TWP + c0 + k];
#pragma unroll
      for (int q = 0; q < QX; ++q) {
        R sum = R(0);
#pragma unroll
        for (int k = 0; k < W; ++k) sum = fmar(wv[q + k], tap<R, F>(k), sum);
        B[r * BW + c0 + q] = sum;
      }
    }
    __syncthreads();
    R wv[CPT + W - 1];
#pragma unroll
    for (int k = 0; k < CPT + W - 1; ++k) wv[k] = B[(ty0 + k) * BW + tx];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      R sum = R(0);
#pragma unroll
      for (int k = 0; k < W; ++k) sum = fmar(wv[c + k], tap<R, F>(k), sum);
      pv[c] = sum;
    }
  };
  // prefill: planes pfirst .. pfirst+W-2 enter the window's first W-1 slots
  // (plane pfirst + q sits in prefetch slot q % AD)
  for (int q0 = 0; q0 < W - 1; q0 += AD) {
#pragma unroll
    for (int j = 0; j < AD; ++j) {
      const int q = q0 + j;
      if (q < W - 1) {
        R pv[CPT];
        filter_plane(pfirst + q, areg[j], pv);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
#pragma unroll
          for (int s = 0; s < W - 2; ++s) win[c][s] = win[c][s + 1];
          win[c][W - 2] = pv[c];
        }
      }
    }
  }
  for (int z0 = zb; z0 < ze; z0 += BZ) {
#pragma unroll
    for (int u = 0; u < BZ; ++u) {
      const int z = z0 + u;
      if (z < ze) {
        R pv[CPT];
        // plane z + UP is plane number (W - 1) + (z - zb) after pfirst
        filter_plane(z + F::UP, areg[((W - 1) % AD + u) % AD], pv);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          win[c][W - 1 + u] = pv[c];
          R sum = R(0);
#pragma unroll
          for (int k = 0; k < W; ++k) sum = fmar(win[c][u + k], tap<R, F>(k), sum);
          if (p.flags & F_A) sum = sum * p.a0;
          if (cout_ok[c]) {
            R* yp = ya + (X0 + tx - p.d0.oy) +
                    (long)p.d0.ny * ((Y0 + ty0 + c - p.d1.oy) + (long)p.d1.ny * (z - p.d2.oy));
            if (p.flags & F_AY) sum = fmar(p.a_y, *yp, sum);
            if (p.flags & F_POT) sum = pot_step(p, yp - ya, sum);
            *yp = sum;
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
      for (int s = 0; s < W - 1; ++s) win[c][s] = win[c][s + BZ];
  }
}

}  // namespace lc

// lc_body_sepq.hpp -- device body of the 128-bit-LDS separable-product kernel
// (internal).  Same arithmetic and summation order as sep3d_body, sepp_body
// and sepr_body (x taps, then y, then z, each in increasing tap order), so
// all give identical bits.
//
// gfx950 serves 128-bit LDS reads (ds_read_b128) at twice the bytes per
// clock of 64-bit ones, and the magic-filter kernels are bound by LDS reads.
// This kernel keeps the structure of the ring-owning kernel (x-waves filter
// rows, y-waves filter columns into z rings and emit the outputs, one barrier
// per plane) with every filter read 16 bytes wide:
//
//   * the tile A[row][col] has a row stride of an odd number of 16-byte units,
//     so an x-task (QX outputs of a row) reads its row in 16-byte pieces and
//     consecutive lanes (consecutive rows) hit distinct banks;
//   * the x-filtered plane is stored transposed, B[col][row] (again an odd
//     number of 16-byte units per column), so a y-task (QY outputs of a
//     column) also reads 16-byte pieces along its column, and the x-task's
//     64-bit stores of one column group fall on consecutive rows.
//
// The LDS bytes read per output drop by about half against the pipelined
// kernel.
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

template <typename R>
struct Vec16;  // 16 bytes of R
template <> struct Vec16<double> { using T = double2; static constexpr int N = 2; };
template <> struct Vec16<float> { using T = float4; static constexpr int N = 4; };

// 16-byte LDS read of V consecutive values starting at p (16-byte aligned)
template <typename R>
__device__ __forceinline__ void lds_read16(const R* p, R* v) {
  using VT = typename Vec16<R>::T;
  const VT q = *reinterpret_cast<const VT*>(p);
  if constexpr (Vec16<R>::N == 2) {
    v[0] = q.x; v[1] = q.y;
  } else {
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  }
}

constexpr int odd_ceil(int v) { return v % 2 ? v : v + 1; }
constexpr int imax(int a, int b) { return a > b ? a : b; }
constexpr int igcd(int a, int b) { return b ? igcd(b, a % b) : a; }

// the shapes this kernel takes: whole 16-byte pieces per task start
template <typename R, int TX, int TY, int QX, int QY>
constexpr bool sepq_shape_ok() {
  return TX % QX == 0 && TY % QY == 0 && QX % Vec16<R>::N == 0 && QY % Vec16<R>::N == 0;
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD>
struct SepqGeo {
  static constexpr int W = F::N, V = Vec16<R>::N;
  static constexpr int TWD = TX + W - 1, TH = TY + W - 1;
  // rows of A = x-tasks per column group: the fewest rows >= TH that make
  // the x-tasks whole waves (TX / QX column groups of AR rows each)
  static constexpr int ARS = 64 / igcd(64, TX / QX);
  static constexpr int AR = (TH + ARS - 1) / ARS * ARS;
  static constexpr int RX = (QX + W - 1 + V - 1) / V;     // 16-byte reads per x-task
  static constexpr int RY = (QY + W - 1 + V - 1) / V;     // 16-byte reads per y-task
  // row stride of A and column stride of B: odd multiples of 16 bytes that
  // hold every value a task reads (the last 16-byte piece may run past the
  // filter window)
  static constexpr int TWP = V * odd_ceil((imax(TWD, TX - QX + RX * V) + V - 1) / V);
  static constexpr int BHT = V * odd_ceil((imax(AR, TY - QY + RY * V) + V - 1) / V);
  static constexpr int NXT = AR * (TX / QX);           // x-tasks per plane
  static constexpr int NYT = TX * (TY / QY);           // y-tasks per plane
  static constexpr int NWX = NXT / 64, NWY = NYT / 64;
  static constexpr int NT = 64 * (NWX + NWY);
  static constexpr int NA = TH * TWD;                  // tile values per plane
  static constexpr int NAP = (NA + NT - 1) / NT;       // per lane
  static constexpr size_t SA = lds_align(sizeof(R) * (AR * TWP + V));  // one half, spare slot
  static constexpr size_t SB = lds_align(sizeof(R) * TX * BHT);
  static constexpr size_t lds_bytes() { return 2 * (SA + SB); }
  static_assert(TX % QX == 0 && TY % QY == 0 && QX % V == 0 && QY % V == 0, "tile shape");
  static_assert(NXT % 64 == 0 && NYT % 64 == 0 && NT <= 1024, "whole waves per pass");
  static_assert(W % 2 == 0 && W % PD == 0, "static ring and prefetch slots");
};

// ME: the epilogue reads or writes memory besides the output (F_AY, F_POT).
// Without (ME false) every wave issues the same memory operations in every
// step: the y-waves' outputs are stored after the x/y branch, by all waves,
// as buffer stores that drop the x-waves' and the invalid lanes' values (out
// of range).  The compiler's wait counts are then exact, and a step waits for
// the tile loads it uses, not for the previous step's stores.
// NTS (ME false): streaming (non-temporal) output stores.
template <typename R, class F, int TX, int TY, int QX, int QY, int PD, bool ME = true, bool NTS = false>
__device__ __forceinline__ void sepq_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                          unsigned char* lds) {
  using G = SepqGeo<R, F, TX, TY, QX, QY, PD>;
  constexpr int W = G::W, V = G::V, NT = G::NT, TWD = G::TWD, TWP = G::TWP, BHT = G::BHT, AR = G::AR;
  constexpr int NA = G::NA, NAP = G::NAP, RX = G::RX, RY = G::RY;
  constexpr int HA = G::SA / sizeof(R), HB = G::SB / sizeof(R);
  R* const A = reinterpret_cast<R*>(lds);                // [2][AR][TWP] (+ spare slot)
  R* const B = reinterpret_cast<R*>(lds + 2 * G::SA);    // [2][TX][BHT]: transposed

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;
  const int pfirst = zb + F::LO;        // input plane of plane number 0
  const int np = (ze - zb) + W - 1;     // input planes of this chunk

  // tile values of this lane: element t + j NT of every plane
  int aoff[NAP], aslot[NAP];
#pragma unroll
  for (int j = 0; j < NAP; ++j) {
    const int f = t + j * NT;
    const bool valid = f < NA;
    const int s0 = fsrc(p.d0, X0 + F::LO + f % TWD), s1 = fsrc(p.d1, Y0 + F::LO + f / TWD);
    aoff[j] = plane_boff<R>((valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1);
    aslot[j] = valid ? (f / TWD) * TWP + f % TWD : AR * TWP;  // past the tile: the spare slot
  }
  auto load_plane = [&](int q, R* dst) {
    const int s2 = q < np ? fsrc(p.d2, pfirst + q) : -1;  // uniform
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, s2);
#pragma unroll
    for (int j = 0; j < NAP; ++j) dst[j] = plane_ld<R>(rs, aoff[j]);
  };

  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
  const bool xwave = wv < G::NWX;
  // x-task: row xr of column group xg; consecutive lanes take consecutive rows
  const int xt = wv * 64 + ln;
  const int xr = xt % AR, xg = xt / AR;
  const int xo = xr * TWP + xg * QX;            // row start in A (16-byte aligned)
  const int bo = (xg * QX) * BHT + xr;          // first output column, this row, in B
  // y-task: column yc, outputs yr0 .. yr0 + QY - 1; consecutive lanes take
  // consecutive columns
  const int yt = (wv - G::NWX) * 64 + ln;
  const int yc = yt % TX, yr0 = (yt / TX) * QY;
  const int yo = yc * BHT + yr0;                // column start in B (16-byte aligned)
  const long ocol = (X0 + yc - p.d0.oy) + (long)p.d0.ny * (Y0 + yr0 - p.d1.oy);
  const bool xin = X0 + yc < p.d0.olo + p.d0.nout;
  const int yrows = (p.d1.olo + p.d1.nout) - (Y0 + yr0);  // output rows left from yr0
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - p.d2.oy) * syo;
  ZCursor zc_ld(p.d2, pfirst + PD + 1);

  R ring[QY][W];
#pragma unroll
  for (int i = 0; i < QY; ++i)
#pragma unroll
    for (int k = 0; k < W; ++k) ring[i][k] = R(0);
  R buf[PD][NAP];
#pragma unroll
  for (int u = 0; u < PD; ++u) load_plane(u, buf[u]);
#pragma unroll
  for (int j = 0; j < NAP; ++j) A[aslot[j]] = buf[0][j];
  load_plane(PD, buf[0]);
  __syncthreads();

  const int nsteps = np + 1;
  for (int sb = 0;; sb += W) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int s = sb + u;
      // ---- tile of plane s + 1, then the load of plane s + 1 + PD
      {
        R* a1 = A + ((u + 1) & 1) * HA;
        R* bq = buf[(u + 1) % PD];
#pragma unroll
        for (int j = 0; j < NAP; ++j) a1[aslot[j]] = bq[j];
        const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, zc_ld.src());
#pragma unroll
        for (int j = 0; j < NAP; ++j) bq[j] = plane_ld<R>(rs, aoff[j]);
        zc_ld.next();
      }
      R outv[QY];  // ME false: the step's outputs of this lane
#pragma unroll
      for (int i = 0; i < QY; ++i) outv[i] = R(0);
      if (xwave) {
        // ---- x-task on plane s: 16-byte reads along the row
        const R* row = A + (u & 1) * HA + xo;
        R acc[QX];
#pragma unroll
        for (int i = 0; i < QX; ++i) acc[i] = R(0);
#pragma unroll
        for (int m = 0; m < RX; ++m) {
          R v[V];
          lds_read16(row + m * V, v);
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const int k = m * V + e;
#pragma unroll
            for (int i = 0; i < QX; ++i)
              if (k - i >= 0 && k - i < W) acc[i] = fmar(v[e], tap<R, F>(k - i), acc[i]);
          }
        }
        R* bd = B + (u & 1) * HB + bo;
#pragma unroll
        for (int i = 0; i < QX; ++i) bd[i * BHT] = acc[i];
      } else {
        // ---- y-task on plane s - 1: 16-byte reads along the column, into the z rings
        const R* colp = B + ((u + 1) & 1) * HB + yo;
        R acc[QY];
#pragma unroll
        for (int i = 0; i < QY; ++i) acc[i] = R(0);
#pragma unroll
        for (int m = 0; m < RY; ++m) {
          R v[V];
          lds_read16(colp + m * V, v);
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const int k = m * V + e;
#pragma unroll
            for (int i = 0; i < QY; ++i)
              if (k - i >= 0 && k - i < W) acc[i] = fmar(v[e], tap<R, F>(k - i), acc[i]);
          }
        }
#pragma unroll
        for (int i = 0; i < QY; ++i) ring[i][(u + W - 1) % W] = acc[i];
        // ---- z: the outputs whose planes s - W .. s - 1 are complete
        if constexpr (ME) {
          if (s >= W) {
#pragma unroll
            for (int i = 0; i < QY; ++i) {
              R sum = R(0);
#pragma unroll
              for (int k = 0; k < W; ++k) sum = fmar(ring[i][(u + k) % W], tap<R, F>(k), sum);
              if (p.flags & F_A) sum = sum * p.a0;
              if (xin && i < yrows) {
                const long o = ocol + (long)i * p.d0.ny + zo;
                R* yp = ya + o;
                if (p.flags & F_AY) sum = fmar(p.a_y, *yp, sum);
                if (p.flags & F_POT) sum = pot_step(p, o, sum);
                *yp = sum;
              }
            }
            zo += syo;
          }
        } else {
#pragma unroll
          for (int i = 0; i < QY; ++i) {
            R sum = R(0);
#pragma unroll
            for (int k = 0; k < W; ++k) sum = fmar(ring[i][(u + k) % W], tap<R, F>(k), sum);
            if (p.flags & F_A) sum = sum * p.a0;
            outv[i] = sum;
          }
        }
      }
      if constexpr (!ME) {
        // the outputs of planes s - W (before plane W: none), stored by every wave
        const bool em = s >= W;  // uniform
        const __amdgpu_buffer_rsrc_t ro = span_rsrc(ya + zo, em ? syo : 0);
#pragma unroll
        for (int i = 0; i < QY; ++i)
          plane_st<R, NTS>(ro, (!xwave && xin && i < yrows) ? (int)((ocol + (long)i * p.d0.ny) * (long)sizeof(R)) : kOffNone,
                      outv[i]);
        zo += em ? syo : 0;
      }
      __syncthreads();
      if (s + 1 == nsteps) return;  // the last output plane (q = np - 1) is stored
    }
  }
}

}  // namespace lc

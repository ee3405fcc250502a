// lc_body_sepr.hpp -- device body of the ring-owning separable-product kernel
// (internal).  Same arithmetic and summation order as sep3d_body and
// sepp_body (y = a Fz Fy Fx x~ + a_y y: x taps, then y, then z, each in
// increasing tap order), so all three give identical bits.
//
// A block owns a TX x TY output tile and marches along z one input plane per
// step, as the pipelined kernel does, but the z filter lives with the y
// filter: the lane that filters QY outputs of a tile column along y keeps the
// last W y-filtered planes of those QY output columns in registers and emits
// their z-filtered outputs itself.  In step s
//
//   * every lane writes its values of the tile of plane s + 1 to A[(s+1)&1]
//     (loaded PD steps ahead) and starts the load of plane s + 1 + PD;
//   * the x-waves filter the tile rows of plane s along x (A[s&1] -> B[s&1],
//     QX outputs per lane, halo rows included);
//   * the y-waves filter QY outputs of a column of plane s - 1 along y
//     (B[(s-1)&1] -> registers), append them to their z rings and emit the
//     outputs whose W planes are complete.
//
// Against the pipelined kernel this removes the third plane buffer (one LDS
// write and one read per output) and a pipeline stage, and it balances the
// waves: an x-wave runs QX W FMAs per plane and a y-wave 2 QY W (y and z),
// where the pipelined kernel gave the z filter to every lane on top of the
// x- or y-task.  One barrier per plane orders the pipeline; the step loop is
// unrolled W times so that ring slots, buffer halves and prefetch registers
// are static.
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

template <typename R, class F, int TX, int TY, int QX, int QY, int PD>
struct SeprGeo {
  static constexpr int W = F::N;
  static constexpr int TWD = TX + W - 1, TH = TY + W - 1;
  static constexpr int TWP = TWD | 1, BW = TX | 1;     // odd LDS row strides
  static constexpr int AR = (TH + 31) / 32 * 32;       // rows of A and B (x-tasks per group)
  static constexpr int NXT = AR * (TX / QX);           // x-tasks per plane
  static constexpr int NYT = TX * (TY / QY);           // y-tasks per plane
  static constexpr int NWX = NXT / 64, NWY = NYT / 64; // x-waves, y-waves
  static constexpr int NT = 64 * (NWX + NWY);          // lanes per block
  static constexpr int NA = TH * TWD;                  // tile values per plane
  static constexpr int NAP = (NA + NT - 1) / NT;       // per lane
  static constexpr size_t SA = lds_align(sizeof(R) * (AR * TWP + 1));  // one half, spare slot
  static constexpr size_t SB = lds_align(sizeof(R) * AR * BW);
  static constexpr size_t lds_bytes() { return 2 * (SA + SB); }
  static_assert(TX % QX == 0 && TY % QY == 0, "tile shape");
  static_assert(NXT % 64 == 0 && NYT % 64 == 0 && NT <= 1024, "whole waves per pass");
  static_assert(W % 2 == 0 && W % PD == 0, "static ring and prefetch slots");
  static_assert(SA % sizeof(R) == 0 && SB % sizeof(R) == 0, "halves");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int PD>
__device__ __forceinline__ void sepr_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                          unsigned char* lds) {
  using G = SeprGeo<R, F, TX, TY, QX, QY, PD>;
  constexpr int W = G::W, NT = G::NT, TWD = G::TWD, TH = G::TH, TWP = G::TWP, BW = G::BW, AR = G::AR;
  constexpr int NA = G::NA, NAP = G::NAP;
  constexpr int HA = G::SA / sizeof(R), HB = G::SB / sizeof(R);
  R* const A = reinterpret_cast<R*>(lds);                // [2][AR][TWP] (+ spare slot)
  R* const B = reinterpret_cast<R*>(lds + 2 * G::SA);    // [2][AR][BW]

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
  // plane number q of the chunk (planes >= np and zero padding read zero)
  auto load_plane = [&](int q, R* dst) {
    const int s2 = q < np ? fsrc(p.d2, pfirst + q) : -1;  // uniform
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, s2);
#pragma unroll
    for (int j = 0; j < NAP; ++j) dst[j] = plane_ld<R>(rs, aoff[j]);
  };

  // the wave decides the pass (uniform per wave)
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
  const bool xwave = wv < G::NWX;
  // x-task: row r of column group g; consecutive lanes take consecutive rows
  // (odd stride TWP: distinct banks), rows TH..AR-1 are spare rows nobody reads
  const int xt = wv * 64 + ln;
  const int xo = (xt % AR) * TWP + (xt / AR) * QX;
  const int bo = (xt % AR) * BW + (xt / AR) * QX;
  // y-task: column c, outputs r0 .. r0 + QY - 1 (consecutive lanes: columns)
  const int yt = (wv - G::NWX) * 64 + ln;
  const int yc = yt % TX, yr0 = (yt / TX) * QY;
  const int yo = yr0 * BW + yc;
  // output columns of a y-lane and the offset of its next output plane
  const long ocol = (X0 + yc - p.d0.oy) + (long)p.d0.ny * (Y0 + yr0 - p.d1.oy);
  const bool xin = X0 + yc < p.d0.olo + p.d0.nout;
  const int yrows = (p.d1.olo + p.d1.nout) - (Y0 + yr0);  // output rows left from r0
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - p.d2.oy) * syo;
  // source plane of the load issued in the next step (plane number PD + 1
  // first), advanced one plane per step
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

  // Planes past the chunk (s >= np, and the plane -1 the first y-step sees)
  // pass through the pipeline like the others: their values reach no stored
  // output, so no step has to test its plane.
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
      if (xwave) {
        // ---- x-task on plane s
        const R* row = A + (u & 1) * HA + xo;
        R acc[QX];
#pragma unroll
        for (int i = 0; i < QX; ++i) acc[i] = R(0);
#pragma unroll
        for (int k = 0; k < QX + W - 1; ++k) {
          const R v = row[k];
#pragma unroll
          for (int i = 0; i < QX; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
        R* bd = B + (u & 1) * HB + bo;
#pragma unroll
        for (int i = 0; i < QX; ++i) bd[i] = acc[i];
      } else {
        // ---- y-task on plane s - 1, into the z rings (plane q in slot q % W)
        const R* colp = B + ((u + 1) & 1) * HB + yo;
        R acc[QY];
#pragma unroll
        for (int i = 0; i < QY; ++i) acc[i] = R(0);
#pragma unroll
        for (int k = 0; k < QY + W - 1; ++k) {
          const R v = colp[k * BW];
#pragma unroll
          for (int i = 0; i < QY; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
#pragma unroll
        for (int i = 0; i < QY; ++i) ring[i][(u + W - 1) % W] = acc[i];
        // ---- z: the output whose planes s - W .. s - 1 are complete
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
      }
      __syncthreads();
      if (s + 1 == nsteps) return;  // the last output plane (q = np - 1) is stored
    }
  }
  (void)TH;
}

}  // namespace lc

// lc_body_sepp.hpp -- device body of the pipelined separable-product kernel
// (internal).  Same arithmetic and summation order as sep3d_body
// (y = a Fz Fy Fx x~ + a_y y), so both give identical bits.
//
// A block owns a TX x TY output tile and marches along z one input plane per
// step with three planes in flight.  In step s
//
//   * every lane writes its values of the tile of plane s + 1 to A[(s+1)&1]
//     (registers loaded PD steps ahead) and starts the load of plane s+1+PD,
//     whose source plane follows the z boundary rule incrementally;
//   * the first half of the waves filter the tile rows of plane s along x
//     (A[s&1] -> B[s&1], QX outputs per lane, halo rows included);
//   * the second half filter the columns of plane s - 1 along y
//     (B[(s-1)&1] -> P[(s-1)&1], QY outputs per lane);
//   * every lane appends the value of plane s - 2 at its output column,
//     P[s&1], to a ring of the last W planes in registers and emits the
//     output whose W planes are complete.
//
// Every buffer half is written and read in different steps, so ONE barrier
// per plane orders the whole pipeline, and the three filter passes of a step
// overlap.  The step loop is unrolled W times: ring slots, buffer halves and
// prefetch registers are then static, and a step tests nothing but whether
// it emits an output, so the scalar unit stays off the critical path.
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

template <typename R, class F, int TX, int TY, int QX, int QY, int PD>
struct SeppGeo {
  static constexpr int W = F::N;
  static constexpr int NT = TX * TY, NW = NT / 64;
  static constexpr int TWD = TX + W - 1, TH = TY + W - 1;
  static constexpr int TWP = TWD | 1, BW = TX | 1;
  static constexpr int NA = TH * TWD;                 // tile values per plane
  static constexpr int NAP = (NA + NT - 1) / NT;      // per lane
  static constexpr int NXT = TH * (TX / QX);          // x-tasks per plane
  static constexpr int NYT = TX * (TY / QY);          // y-tasks per plane
  static constexpr int NWX = NW / 2, NWY = NW - NW / 2;
  static constexpr int XPW = (NXT + NWX - 1) / NWX;   // x-tasks per x-wave
  static constexpr int YPW = (NYT + NWY - 1) / NWY;   // y-tasks per y-wave
  // x-tasks laid out one column group per half-wave when they fit so
  static constexpr bool HALF = TH <= 32 && 2 * NWX >= TX / QX;
  // every x- and y-lane has a task: the passes need no lane masks (the
  // x-lanes past the TH rows filter a spare row 31 nobody reads)
  static constexpr bool XFULL = HALF && TX / QX == 2 * NWX;
  static constexpr bool YFULL = NYT == 64 * NWY;
  static constexpr int AR = XFULL ? 32 : TH;          // rows of A and B
  static constexpr size_t SA = lds_align(sizeof(R) * (AR * TWP + 1));  // one half, spare slot
  static constexpr size_t SB = lds_align(sizeof(R) * AR * BW);
  static constexpr size_t SP = lds_align(sizeof(R) * TY * BW);
  static constexpr size_t lds_bytes() { return 2 * (SA + SB + SP); }
  static_assert(TX % QX == 0 && TY % QY == 0 && NT % 128 == 0, "tile shape");
  static_assert(XPW <= 64 && YPW <= 64, "one task per lane and plane");
  static_assert(W % 2 == 0 && W % PD == 0, "static ring and prefetch slots");
  static_assert(SA % sizeof(R) == 0 && SB % sizeof(R) == 0 && SP % sizeof(R) == 0, "halves");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int PD>
__device__ __forceinline__ void sepp_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                          unsigned char* lds) {
  using G = SeppGeo<R, F, TX, TY, QX, QY, PD>;
  constexpr int W = G::W, NT = G::NT, TWD = G::TWD, TH = G::TH, TWP = G::TWP, BW = G::BW;
  constexpr int NA = G::NA, NAP = G::NAP;
  constexpr int HA = G::SA / sizeof(R), HB = G::SB / sizeof(R), HP = G::SP / sizeof(R);
  R* const A = reinterpret_cast<R*>(lds);                    // [2][AR][TWP]
  R* const B = reinterpret_cast<R*>(lds + 2 * G::SA);        // [2][AR][BW]
  R* const P = reinterpret_cast<R*>(lds + 2 * (G::SA + G::SB));  // [2][TY][BW]

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;
  const int pfirst = zb + F::LO;        // input plane of plane number 0
  const int np = (ze - zb) + W - 1;     // input planes of this chunk

  const int tx = t % TX, ty = t / TX;
  const bool out_ok = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);

  // tile values of this lane: element t + j NT of every plane
  int aoff[NAP], aslot[NAP];
#pragma unroll
  for (int j = 0; j < NAP; ++j) {
    const int f = t + j * NT;
    const bool valid = f < NA;
    const int s0 = fsrc(p.d0, X0 + F::LO + f % TWD), s1 = fsrc(p.d1, Y0 + F::LO + f / TWD);
    aoff[j] = plane_boff<R>((valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1);
    aslot[j] = valid ? (f / TWD) * TWP + f % TWD : TH * TWP;  // past the tile: a spare slot
  }
  // plane number q of the chunk (planes >= np and zero padding read zero)
  auto load_plane = [&](int q, R* dst) {
    const int s2 = q < np ? fsrc(p.d2, pfirst + q) : -1;  // uniform
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, s2);
#pragma unroll
    for (int j = 0; j < NAP; ++j) dst[j] = plane_ld<R>(rs, aoff[j]);
  };

  // task of this lane: the wave's half decides the pass (uniform per wave)
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
  const bool xwave = wv < G::NWX;
  int xo = -1, bo = -1;  // x-task: row start in A, output start in B
  int yo = -1, po = -1;  // y-task: column start in B, output start in P
  if (wv < G::NWX) {
    if constexpr (G::HALF) {
      // one column group per half-wave (the LDS serves 32 lanes of a b64
      // access at a time): rows at the odd strides TWP and BW then hit
      // distinct banks, where two groups in one half-wave collide
      const int r = ln & 31, grp = 2 * wv + (ln >> 5);
      if (G::XFULL || (r < TH && grp < TX / QX)) {
        xo = r * TWP + grp * QX;
        bo = r * BW + grp * QX;
      }
    } else {
      const int task = wv * G::XPW + ln;
      if (ln < G::XPW && task < G::NXT) {
        const int r = task % TH, c0 = (task / TH) * QX;  // consecutive lanes: consecutive rows
        xo = r * TWP + c0;
        bo = r * BW + c0;
      }
    }
  } else {
    const int task = (wv - G::NWX) * G::YPW + ln;
    if (ln < G::YPW && task < G::NYT) {
      const int c = task % TX, r0 = (task / TX) * QY;
      yo = r0 * BW + c;
      po = r0 * BW + c;
    }
  }
  const int pz = ty * BW + tx;
  // output column of this lane and the offset of the next output plane
  const long ocol = (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - p.d2.oy) * syo;
  // source plane of the load issued in the next step (plane number PD + 1
  // first), advanced one plane per step
  ZCursor zc_ld(p.d2, pfirst + PD + 1);

  R ring[W];
#pragma unroll
  for (int k = 0; k < W; ++k) ring[k] = R(0);
  R buf[PD][NAP];
#pragma unroll
  for (int u = 0; u < PD; ++u) load_plane(u, buf[u]);
#pragma unroll
  for (int j = 0; j < NAP; ++j) A[aslot[j]] = buf[0][j];
  load_plane(PD, buf[0]);
  __syncthreads();

  // Planes past the chunk (s >= np, and the planes -1 and -2 the first
  // y- and z-steps see) pass through the pipeline like the others: their
  // values reach no stored output, and no step has to test its plane.
  const int nsteps = np + 2;
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
      // ---- x-task on plane s (x-waves) or y-task on plane s - 1 (y-waves)
      if (xwave && (G::XFULL || xo >= 0)) {
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
      }
      if (!xwave && (G::YFULL || yo >= 0)) {
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
        R* pd = P + ((u + 1) & 1) * HP + po;
#pragma unroll
        for (int i = 0; i < QY; ++i) pd[i * BW] = acc[i];
      }
      // ---- z: plane s - 2 joins the ring; emit the output ending there
      ring[(u + W - 2) % W] = P[(u & 1) * HP + pz];
      if (s >= W + 1) {  // planes s - 2 - (W - 1) .. s - 2 are all >= 0
        R sum = R(0);
#pragma unroll
        for (int k = 0; k < W; ++k) sum = fmar(ring[(u + W - 2 + 1 + k) % W], tap<R, F>(k), sum);
        if (p.flags & F_A) sum = sum * p.a0;
        if (out_ok) {
          R* yp = ya + (ocol + zo);
          if (p.flags & F_AY) sum = fmar(p.a_y, *yp, sum);
          if (p.flags & F_POT) sum = pot_step(p, ocol + zo, sum);
          *yp = sum;
        }
        zo += syo;
      }
      __syncthreads();
      if (s + 1 == nsteps) return;  // the last output plane (s - 2 = np - 1) is stored
    }
  }
}

}  // namespace lc

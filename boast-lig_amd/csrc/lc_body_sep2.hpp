// lc_body_sep2.hpp -- device body of the two-plane separable-product kernel
// (internal).  Same arithmetic as sep3d_body (y = a Fz Fy Fx x~ + a_y y),
// organised like the two-plane kinetic body:
//
//   * a block owns a TX x TY output tile and consumes QZ input planes per
//     step; their (TX + W - 1) x (TY + W - 1) tiles go to LDS from registers
//     loaded one step ahead;
//   * x-tasks (QX outputs of a tile row of one plane, rows of the filter halo
//     included) fill B, then y-tasks (QY outputs of a column of one plane)
//     fill P (a lane takes one task per phase for the 16-tap filters);
//   * every lane owns one output column and keeps the last W - 1 filtered
//     planes in registers; with the step's QZ planes appended it emits the
//     outputs whose W planes are complete.
// Three barriers per QZ planes.  The input tiles are loaded PD steps ahead
// (PD register buffers).  Summation orders equal sep3d_body's, so both give
// identical bits.
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

// PDA: prefetch depth in steps (1 or 2), plus 4 when P shares A's LDS (A is
// dead once the x-tasks have read it; one more barrier per step, before the
// next step writes A, buys 2/3 of the LDS and so two FP64 blocks per CU at
// four planes per step)
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int PDA = 1>
struct Sep2Geo {
  static constexpr int W = F::N;
  static constexpr int NT = TX * TY;
  static constexpr int TWD = TX + W - 1, TH = TY + W - 1;
  static constexpr int TWP = TWD | 1, BW = TX | 1;
  static constexpr int NA = TH * TWD;                 // tile values per plane
  static constexpr int NAP = (NA + NT - 1) / NT;      // per lane per plane
  static constexpr int NAT = QZ * NAP;                // per lane per step
  static constexpr int NXT = QZ * TH * (TX / QX);
  static constexpr int NYT = QZ * TX * (TY / QY);
  static constexpr bool ALIAS = (PDA & 4) != 0;
  static constexpr int PD = PDA & 3;
  static constexpr size_t lds_bytes() {
    return lds_align(sizeof(R) * QZ * TH * TWP) + lds_align(sizeof(R) * QZ * TH * BW) +
           (ALIAS ? 0 : lds_align(sizeof(R) * QZ * TY * BW));
  }
  static_assert(PD == 1 || PD == 2, "prefetch depth");
  static_assert(TY * BW <= TH * TWP, "P fits in A");
  static_assert(TX % QX == 0 && TY % QY == 0 && NT % 64 == 0, "tile shape");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int PDA>
__device__ __forceinline__ void sep3d2_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                            unsigned char* lds) {
  using G = Sep2Geo<R, F, TX, TY, QX, QY, QZ, PDA>;
  constexpr int PD = G::PD;
  constexpr int W = G::W, NT = G::NT, TWD = G::TWD, TH = G::TH, TWP = G::TWP, BW = G::BW;
  constexpr int NA = G::NA, NAP = G::NAP, NAT = G::NAT;
  R* A = reinterpret_cast<R*>(lds);                                          // [QZ][TH][TWP]
  R* B = reinterpret_cast<R*>(lds + lds_align(sizeof(R) * QZ * TH * TWP));  // [QZ][TH][BW]
  R* P = G::ALIAS ? A : B + lds_align(sizeof(R) * QZ * TH * BW) / sizeof(R);  // [QZ][TY][BW]

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;
  const int pfirst = zb + F::LO;        // input plane of step plane number 0
  const int np = (ze - zb) + W - 1;     // input planes of this chunk

  const int tx = t % TX, ty = t / TX;
  const bool out_ok = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);

  // tile values of this lane: element t + j NT of every plane of a step
  // (register k = q NAP + j holds plane q: the plane is known statically)
  int aoff[NAP], aslot[NAP];
#pragma unroll
  for (int j = 0; j < NAP; ++j) {
    const int f = t + j * NT;
    const bool valid = f < NA;
    const int s0 = fsrc(p.d0, X0 + F::LO + f % TWD), s1 = fsrc(p.d1, Y0 + F::LO + f / TWD);
    aoff[j] = plane_boff<R>((valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1);
    aslot[j] = valid ? (f / TWD) * TWP + f % TWD : -1;
  }
  // tiles of the step starting at plane number s0 (planes >= np read zero;
  // zero padding is read as out-of-range)
  auto load_step = [&](int s0, R* dst) {
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
      const int s2 = s0 + q < np ? fsrc(p.d2, pfirst + s0 + q) : -1;  // uniform
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, s2);
#pragma unroll
      for (int j = 0; j < NAP; ++j) dst[q * NAP + j] = plane_ld<R>(rs, aoff[j]);
    }
  };

  R win[W - 1 + QZ];  // the last W - 1 filtered planes, then the step's QZ
#pragma unroll
  for (int s = 0; s < W - 1 + QZ; ++s) win[s] = R(0);

  // one step: planes s0 .. s0 + QZ - 1 from the registers `buf`, which then
  // receive the step PD steps ahead
  auto step = [&](int s0, R* buf) {
#pragma unroll
    for (int q = 0; q < QZ; ++q)
#pragma unroll
      for (int j = 0; j < NAP; ++j)
        if (aslot[j] >= 0) A[q * TH * TWP + aslot[j]] = buf[q * NAP + j];
    if (s0 + PD * QZ < np) load_step(s0 + PD * QZ, buf);
    __syncthreads();
    // ---- x-tasks: QX outputs of a row (halo rows included) of one plane
    for (int task = t; task < G::NXT; task += NT) {
      const int r = task % (QZ * TH), c0 = (task / (QZ * TH)) * QX;  // consecutive lanes: consecutive rows
      const R* row = A + r * TWP + c0;
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
#pragma unroll
      for (int i = 0; i < QX; ++i) B[r * BW + c0 + i] = acc[i];
    }
    __syncthreads();
    // ---- y-tasks: QY outputs of a column of one plane
    for (int task = t; task < G::NYT; task += NT) {
      const int c = task % TX, rest = task / TX;
      const int q = rest % QZ, r0 = (rest / QZ) * QY;
      const R* colp = B + (q * TH + r0) * BW + c;
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
      for (int i = 0; i < QY; ++i) P[(q * TY + r0 + i) * BW + c] = acc[i];
    }
    __syncthreads();
    // ---- z: append the step's planes, emit the complete outputs
#pragma unroll
    for (int q = 0; q < QZ; ++q) win[W - 1 + q] = P[(q * TY + ty) * BW + tx];
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
      const int s = s0 + q;           // plane number just appended
      if (s >= W - 1 && s < np) {
        R sum = R(0);
#pragma unroll
        for (int k = 0; k < W; ++k) sum = fmar(win[q + k], tap<R, F>(k), sum);
        if (p.flags & F_A) sum = sum * p.a0;
        if (out_ok) {
          const int z = zb + s - (W - 1);
          R* yp = ya + (X0 + tx - p.d0.oy) +
                  (long)p.d0.ny * ((Y0 + ty - p.d1.oy) + (long)p.d1.ny * (z - p.d2.oy));
          if (p.flags & F_AY) sum = fmar(p.a_y, *yp, sum);
          if (p.flags & F_POT) sum = pot_step(p, yp - ya, sum);
          *yp = sum;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < W - 1; ++s) win[s] = win[s + QZ];
    if (G::ALIAS) __syncthreads();  // every lane has read P before the next step writes A
  };

  R areg[PD][NAT];  // prefetch depth PD: buffer u serves steps u, u + PD, ...
#pragma unroll
  for (int u = 0; u < PD; ++u) load_step(u * QZ, areg[u]);
  for (int s0 = 0; s0 < np; s0 += PD * QZ) {
#pragma unroll
    for (int u = 0; u < PD; ++u)
      if (s0 + u * QZ < np) step(s0 + u * QZ, areg[u]);
  }
}

}  // namespace lc

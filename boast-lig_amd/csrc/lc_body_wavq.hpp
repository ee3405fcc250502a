// lc_body_wavq.hpp -- device body of the plane-at-a-time 3D sym8 wavelet
// kernel (internal).  Same arithmetic and summation order as sepw3d_body (x
// pass, then y, then z, each over the window steps in increasing order and
// within a step value 0 before value 1), so both give identical bits.
//
// The structure is the magic-filter kernel's (lc_body_sepq.hpp), carried
// over to the two-value steps of the wavelet: the block works one input
// z-plane per barrier instead of one z-step (two planes) per two barriers.
//
//   * x-waves filter the rows of the plane's tile (both x-bands of QX output
//     steps per task) with 16-byte LDS reads; in LDS every row keeps its two
//     values per step as two runs (value 0, then value 1), each 16-byte
//     aligned, and the row stride is an odd number of 16-byte units;
//   * the x-filtered plane is stored transposed, B[col][row] (col = output
//     value of the x-bands, band-major), and y-waves filter one column each
//     for one output step (both y-bands) from the previous plane's B, then
//     push the two values into a 16-plane ring in registers;
//   * every second plane (the last value of an input step) the y-waves emit
//     the output step whose 8-step z window is complete: both z-bands of both
//     y-bands, four values per lane.
//
// ME: the epilogue reads y (F_AY).  Without (ME false) every wave issues the
// same memory operations in every plane: the outputs are stored after the
// x/y branch by all waves, as buffer stores that drop the x-waves' and the
// invalid lanes' values (offset out of range) and the planes that emit
// nothing (record count 0), so the compiler's wait counts are exact and a
// plane waits only for the tile loads it uses.
#pragma once

#include "lc_body_sepq.hpp"
#include "lc_body_wav.hpp"

namespace lc {

// the shapes this kernel takes: whole 16-byte pieces per x-task, whole waves
// of x- and y-tasks (some shapes exist in one precision only)
template <typename R, int TX, int TY, int QX>
constexpr bool wavq_shape_ok() {
  if (QX <= 0 || QX > TX || TX % QX != 0 || QX % Vec16<R>::N != 0) return false;
  const int nr = 2 * (TY + 7), ars = 64 / igcd(64, TX / QX);
  const int ar = (nr + ars - 1) / ars * ars;
  return (ar * (TX / QX)) % 64 == 0 && (2 * TX * TY) % 64 == 0;
}

template <typename R, class Op, int TX, int TY, int QX>
struct WavqGeo {
  static constexpr int W = Op::W, V = Vec16<R>::N;
  static constexpr int TWS = TX + W - 1, TH = TY + W - 1;  // input steps per tile row / column
  static constexpr int RX = (QX + W - 1 + V - 1) / V;       // 16-byte reads per run of an x-task
  // one run of a row (the values c of its TWS steps), 16-byte aligned and
  // holding every value a task's last 16-byte piece reads
  static constexpr int TWSP = V * ((imax(TWS, TX - QX + RX * V) + V - 1) / V);
  static constexpr int TWP = 2 * TWSP + ((2 * TWSP / V) % 2 ? 0 : V);  // odd number of 16-byte units
  static constexpr int NR = 2 * TH;                                     // tile rows: (value, step) of y
  static constexpr int ARS = 64 / igcd(64, TX / QX);
  static constexpr int AR = (NR + ARS - 1) / ARS * ARS;                 // x-task rows (whole waves)
  static constexpr int BHT = odd_ceil(AR);                              // column stride of B
  static constexpr int NXT = AR * (TX / QX);                            // x-tasks per plane
  static constexpr int NYT = 2 * TX * TY;                               // y-tasks per plane
  static constexpr int NWX = NXT / 64, NWY = NYT / 64;
  static constexpr int NT = 64 * (NWX + NWY);
  static constexpr int NA = NR * 2 * TWS;                               // tile values per plane
  static constexpr int NAP = (NA + NT - 1) / NT;                        // per lane
  static constexpr size_t SA = lds_align(sizeof(R) * (AR * TWP + V));  // one half, spare slot
  static constexpr size_t SB = lds_align(sizeof(R) * 2 * TX * BHT);
  static constexpr size_t lds_bytes() { return 2 * (SA + SB); }
  static_assert(W == 8, "sym8 split filters");
  static_assert(TX % QX == 0 && QX % V == 0, "tile shape");
  static_assert(NXT % 64 == 0 && NYT % 64 == 0 && NT <= 1024, "whole waves per pass");
};
// the prefetch slot of plane s is s % PD: static in the 16-plane unrolled loop
template <int PD>
constexpr bool wavq_pd_ok() { return PD >= 1 && (2 * 8) % PD == 0; }

// KB > 0: the x- and y-tasks consume their LDS reads as they arrive, with a
// scheduling fence every KB 16-byte pieces (x) or KB window steps (y), so that
// few values are in flight: fewer VGPRs, more waves per SIMD.  The FMAs run
// in the same order either way.
template <typename R, class Op, int TX, int TY, int QX, int PD, bool ME = true, bool NTS = false, int KB = 0>
__device__ __forceinline__ void wavq_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                          unsigned char* lds) {
  using G = WavqGeo<R, Op, TX, TY, QX>;
  static_assert(wavq_pd_ok<PD>(), "prefetch depth must divide the 16-plane ring");
  constexpr int K = WavCoef<Op>::K;
  constexpr int W = G::W, V = G::V, NT = G::NT, TWS = G::TWS, TWSP = G::TWSP, TWP = G::TWP;
  constexpr int TH = G::TH, AR = G::AR, BHT = G::BHT, NA = G::NA, NAP = G::NAP, RX = G::RX;
  constexpr int NP = 2 * W;  // ring planes: the z window of one output step
  constexpr int HA = G::SA / sizeof(R), HB = G::SB / sizeof(R);
  constexpr int LO = Op::LO;
  R* const A = reinterpret_cast<R*>(lds);              // [2][AR][TWP] (+ spare slot)
  R* const B = reinterpret_cast<R*>(lds + 2 * G::SA);  // [2][2 TX][BHT]: transposed

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long rs_in = 2L * p.d0.nx, ps_in = 4L * p.d0.nx * p.d1.nx;
  const long rs_out = 2L * p.d0.ny, ps_out = 4L * p.d0.ny * p.d1.ny;
  const int sfirst = zb + LO;              // input z-step of plane 0 (planes 2 j + c)
  const int np = 2 * ((ze - zb) + W - 1);  // input planes of this chunk

  // ---- tile values of this lane: element t + j NT of every plane
  int aoff[NAP], aslot[NAP];
#pragma unroll
  for (int j = 0; j < NAP; ++j) {
    const int f = t + j * NT;
    const int rr = f / (2 * TWS), cc = f % (2 * TWS);
    // memory order along x: the DWT interleaves the pair, the IDWT stores bands
    const int s0 = K == KDWT ? (cc >> 1) : cc % TWS;
    const int v0 = K == KDWT ? (cc & 1) : cc / TWS;
    const int v1 = rr / TH, s1 = rr % TH;
    const int j0 = fsrc(p.d0, X0 + LO + s0), j1 = fsrc(p.d1, Y0 + LO + s1);
    const bool ok = f < NA && j0 >= 0 && j1 >= 0;
    aoff[j] = plane_boff<R>(ok ? (int)(in_elem<K>(j0, v0, p.d0.nx) + rs_in * in_elem<K>(j1, v1, p.d1.nx)) : -1);
    aslot[j] = f < NA ? rr * TWP + v0 * TWSP + s0 : AR * TWP;  // past the tile: the spare slot
  }
  // plane q of the chunk: value q % 2 of input z-step sfirst + q / 2
  ZCursor zc_ld(p.d2, sfirst);
  int ld_q = 0;
  auto load_next = [&](R* dst) {  // the next plane in order (uniform)
    const int s2 = ld_q < np ? zc_ld.src() : -1;
    const int c2 = ld_q & 1;
    const long pl = s2 >= 0 ? in_elem<K>(s2, c2, p.d2.nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = span_rsrc(xa + ps_in * pl, s2 >= 0 ? ps_in : 0);
#pragma unroll
    for (int j = 0; j < NAP; ++j) dst[j] = plane_ld<R>(rs, aoff[j]);
    if (c2) zc_ld.next();
    ++ld_q;
  };

  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
  const bool xwave = wv < G::NWX;
  // x-task: row xr of column group xg; consecutive lanes take consecutive rows
  const int xt = wv * 64 + ln;
  const int xr = xt % AR, xg = xt / AR;
  const int xo = xr * TWP + xg * QX;  // run 0 of the task in A (16-byte aligned)
  const int bo = (xg * QX) * BHT + xr;
  // y-task: output column (band c0, step i0l) and output step yi along y;
  // consecutive lanes take consecutive output values in memory
  const int yt = (wv - G::NWX) * 64 + ln;
  const int yl = yt % (2 * TX), yi = yt / (2 * TX);
  const int c0 = K == KIDWT ? (yl & 1) : yl / TX;
  const int i0l = K == KIDWT ? (yl >> 1) : yl % TX;
  const int ycol = c0 * TX + i0l;     // column of B
  const int yo = ycol * BHT + yi;     // run 0 of the task in B
  const bool out_ok = !xwave && (X0 + i0l < p.d0.olo + p.d0.nout) && (Y0 + yi < p.d1.olo + p.d1.nout);
  const long oxy = out_elem<K>(X0 + i0l - p.d0.oy, c0, p.d0.ny) + rs_out * out_elem<K>(Y0 + yi - p.d1.oy, 0, p.d1.ny);
  const long oyb = rs_out * (out_elem<K>(0, 1, p.d1.ny) - out_elem<K>(0, 0, p.d1.ny));  // y-band 1 - y-band 0
  const long ozb = ps_out * (out_elem<K>(0, 1, p.d2.ny) - out_elem<K>(0, 0, p.d2.ny));  // z-band 1 - z-band 0
  long zo = ps_out * out_elem<K>(zb - p.d2.oy, 0, p.d2.ny);
  const long zstep = ps_out * (out_elem<K>(1, 0, p.d2.ny) - out_elem<K>(0, 0, p.d2.ny));

  R ring[2][NP];  // [y-band][plane mod 16]
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int k = 0; k < NP; ++k) ring[b][k] = R(0);
  R buf[PD][NAP];
#pragma unroll
  for (int u = 0; u < PD; ++u) load_next(buf[u]);
#pragma unroll
  for (int j = 0; j < NAP; ++j) A[aslot[j]] = buf[0][j];
  load_next(buf[0]);
  __syncthreads();

  const int nsteps = np + 1;
  for (int sb = 0;; sb += NP) {
#pragma unroll
    for (int u = 0; u < NP; ++u) {
      const int s = sb + u;
      // ---- tile of plane s + 1, then the load of plane s + 1 + PD
      {
        R* a1 = A + ((u + 1) & 1) * HA;
        R* bq = buf[(u + 1) % PD];
#pragma unroll
        for (int j = 0; j < NAP; ++j) a1[aslot[j]] = bq[j];
        load_next(bq);
      }
      R outv[2][2];  // ME false: [y-band][z-band] of the step this plane completes
#pragma unroll
      for (int b = 0; b < 2; ++b) outv[b][0] = outv[b][1] = R(0);
      if (xwave) {
        // ---- x-task on plane s: two runs of 16-byte reads along the row
        const R* row = A + (u & 1) * HA + xo;
        R* bd = B + (u & 1) * HB + bo;
        if constexpr (KB == 0) {
          R v[2][RX * V];
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int m = 0; m < RX; ++m) lds_read16(row + c * TWSP + m * V, &v[c][m * V]);
#pragma unroll
          for (int q = 0; q < QX; ++q) {
            R o0, o1;
            wav_step<R, Op>([&](int ss, int c) { return v[c][q + ss]; }, o0, o1);
            bd[q * BHT] = o0;
            bd[(TX + q) * BHT] = o1;
          }
        } else {
          // piece m of both runs feeds every output whose window holds it;
          // each accumulator still takes its terms step by step, value 0
          // before value 1 (the order of wav_step)
          using C = WavCoef<Op>;
          R t0[QX], t1[QX];
#pragma unroll
          for (int q = 0; q < QX; ++q) t0[q] = t1[q] = R(0);
#pragma unroll
          for (int m = 0; m < RX; ++m) {
            if (m > 0 && m % KB == 0) __builtin_amdgcn_sched_barrier(0);
            R v0[V], v1[V];
            lds_read16(row + m * V, v0);
            lds_read16(row + TWSP + m * V, v1);
#pragma unroll
            for (int e = 0; e < V; ++e) {
              const int k = m * V + e;
#pragma unroll
              for (int q = 0; q < QX; ++q) {
                const int ss = k - q;
                if (ss >= 0 && ss < W) {
                  t0[q] = fmar(v0[e], C::template c<R>(0, 0, ss), t0[q]);
                  t1[q] = fmar(v0[e], C::template c<R>(1, 0, ss), t1[q]);
                  t0[q] = fmar(v1[e], C::template c<R>(0, 1, ss), t0[q]);
                  t1[q] = fmar(v1[e], C::template c<R>(1, 1, ss), t1[q]);
                }
              }
            }
          }
#pragma unroll
          for (int q = 0; q < QX; ++q) {
            bd[q * BHT] = t0[q];
            bd[(TX + q) * BHT] = t1[q];
          }
        }
      } else {
        // ---- y-task on plane s - 1: two runs along the column, into the ring
        const R* colp = B + ((u + 1) & 1) * HB + yo;
        R y0, y1;
        if constexpr (KB == 0) {
          wav_step<R, Op>([&](int ss, int c) { return colp[c * TH + ss]; }, y0, y1);
        } else {
          using C = WavCoef<Op>;
          y0 = y1 = R(0);
#pragma unroll
          for (int ss = 0; ss < W; ++ss) {
            if (ss > 0 && ss % KB == 0) __builtin_amdgcn_sched_barrier(0);
            const R a = colp[ss], b = colp[TH + ss];
            y0 = fmar(a, C::template c<R>(0, 0, ss), y0);
            y1 = fmar(a, C::template c<R>(1, 0, ss), y1);
            y0 = fmar(b, C::template c<R>(0, 1, ss), y0);
            y1 = fmar(b, C::template c<R>(1, 1, ss), y1);
          }
        }
        ring[0][(u + NP - 1) % NP] = y0;
        ring[1][(u + NP - 1) % NP] = y1;
        // ---- z: after the second value of an input step, the output step
        // whose window (planes s - 16 .. s - 1) is complete
        if (u % 2 == 0) {  // static after unrolling
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            R o0, o1;
            wav_step<R, Op>([&](int ss, int c) { return ring[b][(u + 2 * ss + c) % NP]; }, o0, o1);
            if (p.flags & F_A) {
              o0 = o0 * p.a0;
              o1 = o1 * p.a0;
            }
            outv[b][0] = o0;
            outv[b][1] = o1;
          }
          if constexpr (ME) {
            if (s >= NP && out_ok) {
#pragma unroll
              for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int zbd = 0; zbd < 2; ++zbd) {
                  R* yp = ya + zo + zbd * ozb + oxy + b * oyb;
                  R v = outv[b][zbd];
                  if (p.flags & F_AY) v = fmar(p.a_y, *yp, v);
                  *yp = v;
                }
            }
          }
        }
      }
      if (u % 2 == 0) {  // static after unrolling
        const bool em = s >= NP;  // uniform: an output step is complete
        if constexpr (!ME) {
          // stored by every wave; the x-waves', the invalid lanes' and (before
          // plane 16) every value dropped
          const __amdgpu_buffer_rsrc_t r0 = span_rsrc(ya + zo, em ? ps_out : 0);
          const __amdgpu_buffer_rsrc_t r1 = span_rsrc(ya + zo + ozb, em ? ps_out : 0);
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const int boff = out_ok ? (int)((oxy + b * oyb) * (long)sizeof(R)) : kOffNone;
            plane_st<R, NTS>(r0, boff, outv[b][0]);
            plane_st<R, NTS>(r1, boff, outv[b][1]);
          }
        }
        zo += em ? zstep : 0;
      }
      __syncthreads();
      if (s + 1 == nsteps) return;  // the last output step is stored
    }
  }
}

}  // namespace lc

// lc_body_wav.hpp -- device body of the single-pass 3D sym8 wavelet
// transform (internal; lc_fused_wav.hip describes the algorithm).  The body
// takes its LDS from the caller, so the single-call, batched and persistent
// kernels share it.
#pragma once

#include "lc_fused.hpp"

namespace lc {

// C[c'][c](s): weight of input value c of window step s in output value c'
template <class Op>
struct WavCoef;
template <class S>
struct WavCoef<Dwt<S>> {
  static constexpr int K = KDWT;
  template <typename R>
  __device__ __forceinline__ static constexpr R c(int co, int ci, int s) {
    return co == 0 ? (ci == 0 ? S::template le<R>(s) : S::template lo<R>(s))
                   : (ci == 0 ? S::template he<R>(s) : S::template ho<R>(s));
  }
};
template <class S>
struct WavCoef<Idwt<S>> {
  static constexpr int K = KIDWT;
  template <typename R>
  __device__ __forceinline__ static constexpr R c(int co, int ci, int s) {
    return co == 0 ? (ci == 0 ? S::template lo<R>(s) : S::template ho<R>(s))
                   : (ci == 0 ? S::template le<R>(s) : S::template he<R>(s));
  }
};

// both output values of one step from a window v(s, c), s-major order (the
// order of Dwt::compute / Idwt::compute)
template <typename R, class Op, class V>
__device__ __forceinline__ void wav_step(const V& v, R& o0, R& o1) {
  using C = WavCoef<Op>;
  R t0 = R(0), t1 = R(0);
#pragma unroll
  for (int s = 0; s < Op::W; ++s) {
    t0 = fmar(v(s, 0), C::template c<R>(0, 0, s), t0);
    t1 = fmar(v(s, 0), C::template c<R>(1, 0, s), t1);
    t0 = fmar(v(s, 1), C::template c<R>(0, 1, s), t0);
    t1 = fmar(v(s, 1), C::template c<R>(1, 1, s), t1);
  }
  o0 = t0;
  o1 = t1;
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
constexpr size_t sepw3d_lds_bytes() {
  return lds_align(sizeof(R) * 2 * 2 * (TY + Op::W - 1) * ((2 * (TX + Op::W - 1)) | 1)) +
         lds_align(sizeof(R) * 2 * 2 * (TY + Op::W - 1) * ((2 * TX) | 1));
}

template <typename R, class Op, int TX, int TY, int QX, int BZ, int AD>
__device__ __forceinline__ void sepw3d_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                            unsigned char* lds) {
  constexpr int K = WavCoef<Op>::K;
  constexpr int W = Op::W, LO = Op::LO, UP = Op::UP;
  constexpr int NT = 2 * TX * TY;
  constexpr int TWS = TX + W - 1;          // input steps per tile row
  constexpr int AROW = (2 * TWS) | 1;      // LDS row stride (odd)
  constexpr int TH = TY + W - 1;           // input steps per tile column
  constexpr int NR = 2 * TH;               // tile rows (both values of a step)
  constexpr int BW = (2 * TX) | 1;         // x-filtered row stride (odd)
  constexpr int NA = NR * 2 * TWS;         // tile values per z-plane
  constexpr int NAT = (NA + NT - 1) / NT;
  constexpr int NSEG = TX / QX;
  constexpr int NXT = 2 * NR * NSEG;       // x-tasks of both z-planes
  constexpr int WS = W - 1 + BZ;
  static_assert(TX % QX == 0 && BZ % AD == 0, "tile shape");
  // LDS: the tile of both z-planes, then their x-filtered rows
  R (*A)[NR * AROW] = reinterpret_cast<R (*)[NR * AROW]>(lds);
  R (*B)[NR * BW] = reinterpret_cast<R (*)[NR * BW]>(lds + lds_align(sizeof(R) * 2 * NR * AROW));

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long rs_in = 2L * p.d0.nx, ps_in = 4L * p.d0.nx * p.d1.nx;
  const long rs_out = 2L * p.d0.ny, ps_out = 4L * p.d0.ny * p.d1.ny;

  // ---- this lane's output column: value c0 of x-step i0l, y-step i1l
  const int col = t % (2 * TX), i1l = t / (2 * TX);
  const int c0 = K == KIDWT ? (col & 1) : col / TX;
  const int i0l = K == KIDWT ? (col >> 1) : col % TX;
  const int bcol = c0 * TX + i0l;
  const bool out_ok = (X0 + i0l < p.d0.olo + p.d0.nout) && (Y0 + i1l < p.d1.olo + p.d1.nout);
  const long yoff_xy = out_elem<K>(X0 + i0l - p.d0.oy, c0, p.d0.ny) +
                       rs_out * out_elem<K>(Y0 + i1l - p.d1.oy, 0, p.d1.ny);
  const long ystep1 = rs_out * (out_elem<K>(0, 1, p.d1.ny) - out_elem<K>(0, 0, p.d1.ny));

  // ---- tile values of this lane (same xy positions in both z-planes)
  int aoff[NAT], aslot[NAT];
#pragma unroll
  for (int k = 0; k < NAT; ++k) {
    const int e = t + k * NT;
    const int rr = e / (2 * TWS), cc = e % (2 * TWS);
    // memory order along x: DWT interleaves the pair, IDWT stores bands
    const int s0 = K == KDWT ? (cc >> 1) : cc % TWS;
    const int v0 = K == KDWT ? (cc & 1) : cc / TWS;
    const int v1 = rr / TH, s1 = rr % TH;
    const int j0 = fsrc(p.d0, X0 + LO + s0), j1 = fsrc(p.d1, Y0 + LO + s1);
    const bool ok = e < NA && j0 >= 0 && j1 >= 0;
    aoff[k] = plane_boff<R>(ok ? (int)(in_elem<K>(j0, v0, p.d0.nx) + rs_in * in_elem<K>(j1, v1, p.d1.nx)) : -1);
    aslot[k] = e < NA ? rr * AROW + v0 * TWS + s0 : 2 * TWS;  // else row 0's padding: never read
  }
  // the two z-planes of input step s2 as buffer resources: zero padding is
  // read as out-of-range (no branch, no select on the loaded values)
  auto load_src = [&](int s2, R (*dst)[NAT]) {  // s2 uniform
    const long np2 = s2 >= 0 ? ps_in : 0;
    const __amdgpu_buffer_rsrc_t r0 = span_rsrc(xa + (s2 >= 0 ? ps_in * in_elem<K>(s2, 0, p.d2.nx) : 0), np2);
    const __amdgpu_buffer_rsrc_t r1 = span_rsrc(xa + (s2 >= 0 ? ps_in * in_elem<K>(s2, 1, p.d2.nx) : 0), np2);
#pragma unroll
    for (int k = 0; k < NAT; ++k) {
      dst[0][k] = plane_ld<R>(r0, aoff[k]);
      dst[1][k] = plane_ld<R>(r1, aoff[k]);
    }
  };

  R areg[AD][2][NAT];
  const int plast = ze - 1 + UP;
  const int pfirst = zb + LO;
#pragma unroll
  for (int h = 0; h < AD; ++h)
    if (pfirst + h <= plast) load_src(fsrc(p.d2, pfirst + h), areg[h]);
  // source of the input step loaded next (every filter step loads the step AD
  // ahead; past the last one it loads a step nobody uses)
  ZCursor zc_ld(p.d2, pfirst + AD);

  // x and y passes of the next input z-step (tile in ar) -> pv[c2][c1']
  auto filter_step = [&](R (*ar)[NAT], R (*pv)[2]) {
#pragma unroll
    for (int k = 0; k < NAT; ++k) {
      A[0][aslot[k]] = ar[0][k];
      A[1][aslot[k]] = ar[1][k];
    }
    load_src(zc_ld.src(), ar);
    zc_ld.next();
    __syncthreads();
    for (int task = t; task < NXT; task += NT) {
      // consecutive lanes on consecutive rows (odd stride: no bank conflicts)
      const int c2 = task / (NR * NSEG), rem = task % (NR * NSEG);
      const int rr = rem % NR, seg = rem / NR;
      const R* row = &A[c2][rr * AROW + seg * QX];
      R* brow = &B[c2][rr * BW + seg * QX];
#pragma unroll
      for (int q = 0; q < QX; ++q) {
        R o0, o1;
        wav_step<R, Op>([&](int s, int c) { return row[c * TWS + q + s]; }, o0, o1);
        brow[q] = o0;
        brow[TX + q] = o1;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2)
      wav_step<R, Op>([&](int s, int c) { return B[c2][(c * TH + i1l + s) * BW + bcol]; },
                      pv[c2][0], pv[c2][1]);
  };

  R win[2][WS][2];  // [c1'][window step][c2]
  for (int q0 = 0; q0 < W - 1; q0 += AD) {
#pragma unroll
    for (int j = 0; j < AD; ++j) {
      const int q = q0 + j;
      if (q < W - 1) {
        R pv[2][2];
        filter_step(areg[j], pv);  // input step pfirst + q
#pragma unroll
        for (int c1 = 0; c1 < 2; ++c1) {
#pragma unroll
          for (int s = 0; s < W - 2; ++s) {
            win[c1][s][0] = win[c1][s + 1][0];
            win[c1][s][1] = win[c1][s + 1][1];
          }
          win[c1][W - 2][0] = pv[0][c1];
          win[c1][W - 2][1] = pv[1][c1];
        }
      }
    }
  }
  for (int z0 = zb; z0 < ze; z0 += BZ) {
#pragma unroll
    for (int u = 0; u < BZ; ++u) {
      const int z = z0 + u;
      if (z < ze) {
        R pv[2][2];
        filter_step(areg[((W - 1) % AD + u) % AD], pv);  // input step z + UP
        const long zo0 = ps_out * out_elem<K>(z - p.d2.oy, 0, p.d2.ny);
        const long zo1 = ps_out * out_elem<K>(z - p.d2.oy, 1, p.d2.ny);
#pragma unroll
        for (int c1 = 0; c1 < 2; ++c1) {
          win[c1][W - 1 + u][0] = pv[0][c1];
          win[c1][W - 1 + u][1] = pv[1][c1];
          R o0, o1;
          wav_step<R, Op>([&](int s, int c) { return win[c1][u + s][c]; }, o0, o1);
          if (p.flags & F_A) {
            o0 = o0 * p.a0;
            o1 = o1 * p.a0;
          }
          if (out_ok) {
            R* y0 = ya + yoff_xy + c1 * ystep1 + zo0;
            R* y1 = ya + yoff_xy + c1 * ystep1 + zo1;
            if (p.flags & F_AY) {
              o0 = fmar(p.a_y, *y0, o0);
              o1 = fmar(p.a_y, *y1, o1);
            }
            *y0 = o0;
            *y1 = o1;
          }
        }
      }
    }
#pragma unroll
    for (int c1 = 0; c1 < 2; ++c1)
#pragma unroll
      for (int s = 0; s < W - 1; ++s) {
        win[c1][s][0] = win[c1][s + BZ][0];
        win[c1][s][1] = win[c1][s + BZ][1];
      }
  }
}

}  // namespace lc

// lc_body_sump.hpp -- device body of the pipelined kinetic-form kernel
// (internal).  Same arithmetic and summation order as sum3d2_body (y = a0 Fx
// x~ + a1 Fy x~ + a2 Fz x~ + a_x x~ + a_y y, three FP64 energies), so the two
// give identical outputs; organised like the pipelined separable kernel, with
// ONE barrier per plane:
//
// A block owns a TX x TY output tile and marches along z one plane per step;
// every lane owns one output column and keeps the column's input values
// along z in a register window.  In step s (output plane s of the chunk)
//
//   * every lane writes plane s + 1 of the cross-shaped tile to T[(s+1)&1]:
//     its own centre value from the window, its share of the halo (14
//     columns left and right, 14 rows above and below) from registers loaded
//     PD steps ahead, and starts the halo load of plane s + 1 + PD;
//   * the task lanes filter plane s (T[s&1]) along x (QX outputs of a row)
//     or along y (QY outputs of a column) into S[s&1];
//   * every lane combines plane s - 1 of its column: the x- and y-terms from
//     S[(s-1)&1], the z-term from its window, the epilogue and the energies,
//     and stores the output.
//
// Every LDS half is written and read in different steps, so the barrier at
// the end of a step orders everything.  The step loop is unrolled U times:
// the window holds 28 + U planes and slides by U planes per iteration, the
// halo registers are static slots, and the column values of the next
// iteration are loaded one iteration ahead.
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int U>
struct SumpGeo {
  static constexpr int W = F::N, HL = -F::LO, HR = F::UP;
  static constexpr int NT = TX * TY;                 // lanes = output columns
  static constexpr int TW = (TX + HL + HR) | 1;      // LDS tile row stride (odd)
  static constexpr int TH = TY + HL + HR;            // tile rows
  static constexpr int SW = TX | 1;                  // term row stride (odd)
  static constexpr int NHP = TY * (HL + HR) + TX * (HL + HR);  // halo values per plane
  static constexpr int NHL = (NHP + NT - 1) / NT;    // per lane per plane
  static constexpr int NXT = TY * (TX / QX);         // x-tasks per plane
  static constexpr int NYT = TX * (TY / QY);         // y-tasks per plane
  static constexpr int WS = W - 1 + U;               // z window (planes)
  static constexpr size_t ST = lds_align(sizeof(R) * TH * TW);   // one tile half
  static constexpr size_t SS = lds_align(sizeof(R) * TY * SW);   // one term half
  static constexpr size_t lds_bytes() {
    return 2 * ST + 4 * SS + lds_align(sizeof(double) * (NT / 64 + 1));
  }
  static_assert(TX % QX == 0 && TY % QY == 0 && NT % 64 == 0, "tile shape");
  static_assert(NXT % 64 == 0 && NYT % 64 == 0 && NXT + NYT <= NT, "whole waves per task kind");
  static_assert(U % PD == 0 && U % 2 == 0, "static halo slots and LDS halves");
};

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int U>
__device__ __forceinline__ void sump_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                          long bid, unsigned char* lds) {
  using G = SumpGeo<R, F, TX, TY, QX, QY, PD, U>;
  constexpr int W = G::W, HL = G::HL, HR = G::HR, TW = G::TW, TH = G::TH, SW = G::SW;
  constexpr int NT = G::NT, NHP = G::NHP, NHL = G::NHL, WS = G::WS;
  constexpr int HT = G::ST / sizeof(R), HS = G::SS / sizeof(R);
  R* const T = reinterpret_cast<R*>(lds);                   // [2][TH][TW]
  R* const SX = reinterpret_cast<R*>(lds + 2 * G::ST);      // [2][TY][SW]
  R* const SY = SX + 2 * HS;                                // [2][TY][SW]
  double* red = reinterpret_cast<double*>(lds + 2 * G::ST + 4 * G::SS);

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int nz = min(p.zchunk, p.d2.olo + p.d2.nout - zb);  // output planes of this chunk
  const long sxy = (long)p.d0.nx * p.d1.nx;

  // ---- the lane's output column
  const int tx = t % TX, ty = t / TX;
  const int s0c = fsrc(p.d0, X0 + tx), s1c = fsrc(p.d1, Y0 + ty);
  const int coff = (s0c >= 0 && s1c >= 0) ? s0c + p.d0.nx * s1c : -1;
  const int cboff = plane_boff<R>(coff);
  const bool out_ok = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);
  const int tcen = (ty + HL) * TW + tx + HL;   // the column's slot in a tile plane
  const int scen = ty * SW + tx;               // and in a term plane

  // ---- halo values of this lane: element t + j NT of every plane
  int hslot[NHL], hoff[NHL];
#pragma unroll
  for (int j = 0; j < NHL; ++j) {
    const int f = t + j * NT;
    const bool valid = f < NHP;
    int r = 0, c = 0;
    if (f < TY * (HL + HR)) {  // x-halo: rows of the tile, columns left / right
      r = f / (HL + HR);
      const int u = f % (HL + HR);
      c = u < HL ? u - HL : TX + (u - HL);
    } else if (valid) {        // y-halo: columns of the tile, rows above / below
      const int g = f - TY * (HL + HR);
      c = g % TX;
      const int u = g / TX;
      r = u < HL ? u - HL : TY + (u - HL);
    }
    hslot[j] = valid ? (r + HL) * TW + (c + HL) : 0;  // else the corner (0, 0): no task reads it
    const int s0 = fsrc(p.d0, X0 + c), s1 = fsrc(p.d1, Y0 + r);
    hoff[j] = plane_boff<R>((valid && s0 >= 0 && s1 >= 0) ? s0 + p.d0.nx * s1 : -1);
  }
  // plane cursors (zero padding and planes past the input read as zero):
  // the column values run 14 + U planes ahead of the halo
  ZCursor zc_col(p.d2, zb - HL), zc_halo(p.d2, zb);
  auto col_next = [&]() -> R {
    const R v = plane_ld<R>(plane_rsrc(xa, sxy, zc_col.src()), cboff);
    zc_col.next();
    return v;
  };
  auto halo_next = [&](R* dst) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, zc_halo.src());
    zc_halo.next();
#pragma unroll
    for (int j = 0; j < NHL; ++j) dst[j] = plane_ld<R>(rs, hoff[j]);
  };
  // output plane z has an input value (x~ exists) on every periodic plane,
  // else on the stored / logical range
  const bool zper = p.d2.bc == PER;
  const int zoff = p.d2.bc == SHRINK ? p.d2.ox : 0, zlim = p.d2.bc == SHRINK ? p.d2.nx : p.d2.n;
  const long ocol = (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - 1 - p.d2.oy) * syo;   // output offset of plane s - 1

  // ---- task of this lane (uniform per wave)
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int kind = wv * 64 < G::NXT ? 0 : (wv * 64 < G::NXT + G::NYT ? 1 : 2);
  int tsrc = 0, tdst = 0;
  if (kind == 0) {         // consecutive lanes on consecutive rows (odd stride)
    const int r = t % TY, c0 = (t / TY) * QX;
    tsrc = (r + HL) * TW + c0;
    tdst = r * SW + c0;
  } else if (kind == 1) {  // consecutive lanes on consecutive columns
    const int u = t - G::NXT;
    const int c = u % TX, r0 = (u / TX) * QY;
    tsrc = r0 * TW + c + HL;
    tdst = r0 * SW + c;
  }

  // ---- window: win[k] = input plane zb - 1 - HL + k (plane s - 1 - HL + k in step s of the iteration)
  R win[WS];
  win[0] = R(0);                             // plane zb - 1 - HL: only the skipped combine of step 0
#pragma unroll
  for (int k = 1; k < WS; ++k) win[k] = col_next();
  R nxt[U];
#pragma unroll
  for (int k = 0; k < U; ++k) nxt[k] = col_next();
  R hreg[PD][NHL];
  // the tile of plane 0: centre from the window, halo loaded now
  {
    R h0[NHL];
    halo_next(h0);
    T[tcen] = win[HL + 1];
#pragma unroll
    for (int j = 0; j < NHL; ++j) T[hslot[j]] = h0[j];
  }
#pragma unroll
  for (int d = 0; d < PD; ++d) halo_next(hreg[(d + 1) % PD]);  // planes 1 .. PD
  __syncthreads();

  const bool want_dot = (p.flags & F_DOT) && arr == p.narr - 1;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;

  // steps 0 .. nz: the task of step s covers plane s, the combine plane s - 1
  for (int sb = 0; sb <= nz; sb += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = sb + u;
      // ---- tile of plane s + 1 (centre: input plane s + 1 = win[u + HL + 2])
      {
        R* t1 = T + ((u + 1) & 1) * HT;
        R* hq = hreg[(u + 1) % PD];
        t1[tcen] = win[u + HL + 2];
#pragma unroll
        for (int j = 0; j < NHL; ++j) t1[hslot[j]] = hq[j];
        halo_next(hq);  // plane s + 1 + PD
      }
      // ---- task on plane s
      if (kind == 0) {
        const R* row = T + (u & 1) * HT + tsrc;
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
        R* sd = SX + (u & 1) * HS + tdst;
#pragma unroll
        for (int i = 0; i < QX; ++i) sd[i] = acc[i];
      } else if (kind == 1) {
        const R* colp = T + (u & 1) * HT + tsrc;
        R acc[QY];
#pragma unroll
        for (int i = 0; i < QY; ++i) acc[i] = R(0);
#pragma unroll
        for (int k = 0; k < QY + W - 1; ++k) {
          const R v = colp[k * TW];
#pragma unroll
          for (int i = 0; i < QY; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
        R* sd = SY + (u & 1) * HS + tdst;
#pragma unroll
        for (int i = 0; i < QY; ++i) sd[i * SW] = acc[i];
      }
      // ---- combine plane s - 1 (input planes s - 1 - HL .. s - 1 + HR = win[u .. u + W - 1])
      if (s >= 1 && s <= nz) {
        const int z = zb + s - 1;
        R zt = R(0);
#pragma unroll
        for (int k = 0; k < W; ++k) zt = fmar(win[u + k], tap<R, F>(k), zt);
        const R xc = win[u + HL];
        const bool has_x = out_ok && coff >= 0 && (zper || (unsigned)(z - zoff) < (unsigned)zlim);
        R t0 = SX[((u + 1) & 1) * HS + scen] * p.a0;
        const R t1 = SY[((u + 1) & 1) * HS + scen] * p.a1;
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
          R* yp = ya + (ocol + zo);
          if (p.flags & F_AY) tot = fmar(p.a_y, *yp, tot);
          *yp = tot;
        }
      }
      zo += syo;
      __syncthreads();
    }
    // slide the window by U planes; the next U come from the prefetch
#pragma unroll
    for (int k = 0; k < W - 1; ++k) win[k] = win[k + U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      win[W - 1 + k] = nxt[k];
      nxt[k] = col_next();
    }
  }
  if (p.part) {
    double sm = wave_block_sum<R>(d0, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[bid] = sm;
    sm = wave_block_sum<R>(d1, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[p.nparts + bid] = sm;
    sm = wave_block_sum<R>(d2, red, NT);
    if (t == 0 && arr == p.narr - 1) p.part[2 * p.nparts + bid] = sm;
  }
  (void)HR;
  (void)TH;
}

}  // namespace lc

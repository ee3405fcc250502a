// lc_body_sum2.hpp -- device body of the two-plane kinetic-form kernel
// (internal).  Same arithmetic as sum3d_body (y = a0 Fx x~ + a1 Fy x~ +
// a2 Fz x~ + a_x x~ + a_y y with the three FP64 energies), organised for LDS
// bandwidth and occupancy:
//
//   * a block owns a TX x TY output tile and marches along z QZ planes per
//     step, with one barrier pair per step;
//   * every lane owns one output column (TX TY lanes) and keeps the column's
//     values along z in a register window of W - 1 + QZ steps; it writes
//     the centre of the step's planes into the LDS tile from that window,
//     while the cross-shaped filter halo comes from registers prefetched one
//     step ahead;
//   * after the first barrier every lane runs ONE task of QX x-outputs of a
//     tile row or QY y-outputs of a tile column (QX + W - 1 LDS reads per
//     QX outputs, so LDS traffic stays below the FMA rate);
//   * after the second barrier the column owner adds its z-term from the
//     window, applies the epilogue and stores the QZ outputs.
// Index rules per dim as in lc_fused.hpp (x~ wrapped on periodic dims,
// zero outside the stored input on free dims).
#pragma once

#include "lc_body_conv.hpp"

namespace lc {

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ>
struct Sum2Geo {
  static constexpr int W = F::N, HL = -F::LO, HR = F::UP;
  static constexpr int NT = TX * TY;                 // lanes = output columns
  static constexpr int TW = (TX + HL + HR) | 1;      // LDS tile row stride (odd)
  static constexpr int TH = TY + HL + HR;            // tile rows
  static constexpr int SW = TX | 1;                  // x-/y-term row stride (odd)
  static constexpr int NHP = TY * (HL + HR) + TX * (HL + HR);  // halo values per plane
  static constexpr int NHL = (NHP + NT - 1) / NT;    // per lane per plane
  static constexpr int NHT = QZ * NHL;               // per lane per step
  static constexpr int NXT = QZ * TY * (TX / QX);    // x-tasks per step
  static constexpr int NYT = QZ * TX * (TY / QY);    // y-tasks per step
  static constexpr int WS = W - 1 + QZ;              // z window
  static constexpr size_t lds_bytes() {
    return lds_align(sizeof(R) * QZ * TH * TW) + 2 * lds_align(sizeof(R) * QZ * TY * SW) +
           lds_align(sizeof(double) * (NT / 64 + 1));
  }
  static_assert(TX % QX == 0 && TY % QY == 0 && NT % 64 == 0, "tile shape");
  static_assert(NXT + NYT <= NT, "one task per lane at most");
};

// KC > 0: a scheduling fence every KC tile reads of a task bounds the values
// in flight (fewer VGPRs, more waves per SIMD).  PW: steps of column planes in
// flight; the column loads are the ones that miss L2 (each plane's column
// value is first needed 14 planes ahead of its tile), so PW > 1 issues them
// PW steps ahead, with the step loop unrolled PW times so that every
// in-flight register has a static slot.  ME: the epilogue reads y (F_AY).
// Without (ME false) every step issues the same memory operations: the
// outputs go out as buffer stores that drop the lanes outside the output
// and the planes past the chunk (out of range), and the set-up ends with as
// many dropped stores as a step, so the compiler's wait counts are exact and
// a step waits for the loads it uses, not for the previous step's stores.
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int KC, int PW = 1, bool ME = true>
__device__ __forceinline__ void sum3d2_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr,
                                            long bid, unsigned char* lds) {
  using G = Sum2Geo<R, F, TX, TY, QX, QY, QZ>;
  constexpr int W = G::W, HL = G::HL, HR = G::HR, TW = G::TW, TH = G::TH, SW = G::SW;
  constexpr int NT = G::NT, NHP = G::NHP, NHL = G::NHL, NHT = G::NHT, WS = G::WS;
  R* T = reinterpret_cast<R*>(lds);                                   // [QZ][TH][TW]
  R* SX = reinterpret_cast<R*>(lds + lds_align(sizeof(R) * QZ * TH * TW));  // [QZ][TY][SW]
  R* SY = SX + lds_align(sizeof(R) * QZ * TY * SW) / sizeof(R);
  double* red = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(SY) +
                                          lds_align(sizeof(R) * QZ * TY * SW));

  const int t = threadIdx.x;
  const R* __restrict__ xa = p.x + arr * p.nti;
  R* __restrict__ ya = p.y + arr * p.nto;
  const int X0 = p.d0.olo + bx * TX, Y0 = p.d1.olo + by * TY;
  const int zb = p.d2.olo + zc * p.zchunk;
  const int ze = min(zb + p.zchunk, p.d2.olo + p.d2.nout);
  const long sxy = (long)p.d0.nx * p.d1.nx;

  // ---- the lane's output column
  const int tx = t % TX, ty = t / TX;
  const int s0c = fsrc(p.d0, X0 + tx), s1c = fsrc(p.d1, Y0 + ty);
  const int coff = (s0c >= 0 && s1c >= 0) ? s0c + p.d0.nx * s1c : -1;
  const int cboff = plane_boff<R>(coff);
  const bool out_ok = (X0 + tx < p.d0.olo + p.d0.nout) && (Y0 + ty < p.d1.olo + p.d1.nout);

  // ---- halo values of this lane: element t + j NT of every plane of a step
  // (register k = q NHL + j holds plane q: the plane is known statically)
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
  // planes of the column values and halos loaded next, one cursor each
  // (zero padding read as out-of-range)
  ZCursor zc_col(p.d2, zb + F::LO), zc_halo(p.d2, zb);
  auto col_next = [&]() -> R {
    const R v = plane_ld<R>(plane_rsrc(xa, sxy, zc_col.src()), cboff);
    zc_col.next();
    return v;
  };
  auto load_halo = [&](R* dst) {
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, zc_halo.src());
      zc_halo.next();
#pragma unroll
      for (int j = 0; j < NHL; ++j) dst[q * NHL + j] = plane_ld<R>(rs, hoff[j]);
    }
  };
  // output plane z has an input value (x~ exists) on every periodic plane,
  // else on the stored / logical range
  const bool zper = p.d2.bc == PER;
  const int zoff = p.d2.bc == SHRINK ? p.d2.ox : 0, zlim = p.d2.bc == SHRINK ? p.d2.nx : p.d2.n;
  // output column of this lane and the offset of the next output plane
  const long ocol = (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - p.d2.oy) * syo;

  // ---- window: steps zb + LO .. zb + LO + WS - 1 of the column
  R win[WS];
#pragma unroll
  for (int s = 0; s < W - 1; ++s) win[s] = col_next();  // planes zb + LO .. zb + UP - 1
  // the window's next planes, PW steps of them, and the first step's halo,
  // issued in the order of a step's loads (halo, then the far column planes)
  R nxt[PW][QZ];
  R hreg[NHT];
#pragma unroll
  for (int w = 0; w < PW - 1; ++w)
#pragma unroll
    for (int q = 0; q < QZ; ++q) nxt[w][q] = col_next();
  load_halo(hreg);
#pragma unroll
  for (int q = 0; q < QZ; ++q) nxt[PW - 1][q] = col_next();
  if constexpr (!ME) {
    const __amdgpu_buffer_rsrc_t none = span_rsrc(ya, 0);
#pragma unroll
    for (int q = 0; q < QZ; ++q) plane_st<R, true>(none, kOffNone + q * (int)sizeof(R), R(0));
  }

  const bool want_dot = (p.flags & F_DOT) && arr == p.narr - 1;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;

  for (int zu = zb; zu < ze; zu += PW * QZ) {
#pragma unroll
    for (int u = 0; u < PW; ++u) {
      const int z0 = zu + u * QZ;
      if (PW > 1 && z0 >= ze) break;  // uniform
      // window gains the planes z0 + UP .. z0 + UP + QZ - 1
#pragma unroll
      for (int q = 0; q < QZ; ++q) win[W - 1 + q] = nxt[u][q];
      // tile of planes z0 .. z0+QZ-1: centre from the window, halo from registers
#pragma unroll
      for (int q = 0; q < QZ; ++q) T[(q * TH + ty + HL) * TW + tx + HL] = win[HL + q];
#pragma unroll
      for (int q = 0; q < QZ; ++q)
#pragma unroll
        for (int j = 0; j < NHL; ++j)
          T[q * TH * TW + hslot[j]] = hreg[q * NHL + j];
      // prefetch: halo planes one step ahead, column planes PW steps ahead
      // (after the last steps: planes nobody uses).  The wave issues them at
      // a raised priority, so that they enter the memory pipeline ahead of
      // the other waves' filter work on the SIMD (FP64 kernel -1.5%)
      __builtin_amdgcn_s_setprio(1);
      load_halo(hreg);
#pragma unroll
      for (int q = 0; q < QZ; ++q) nxt[u][q] = col_next();
      __builtin_amdgcn_s_setprio(0);
      __syncthreads();
      // ---- one task per lane: x-outputs of a row or y-outputs of a column
      if (t < G::NXT) {
        // consecutive lanes on consecutive rows (odd stride: no bank conflicts)
        const int r = t % (QZ * TY), c0 = (t / (QZ * TY)) * QX;
        const int q = r / TY, rr = r % TY;
        const R* row = T + (q * TH + rr + HL) * TW + c0;
        R acc[QX];
#pragma unroll
        for (int i = 0; i < QX; ++i) acc[i] = R(0);
#pragma unroll
        for (int k = 0; k < QX + W - 1; ++k) {
          if (KC > 0 && k > 0 && k % KC == 0) __builtin_amdgcn_sched_barrier(0);
          const R v = row[k];
#pragma unroll
          for (int i = 0; i < QX; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
#pragma unroll
        for (int i = 0; i < QX; ++i) SX[(q * TY + rr) * SW + c0 + i] = acc[i];
      } else if (t < G::NXT + G::NYT) {
        const int u = t - G::NXT;
        const int c = u % TX, rest = u / TX;
        const int q = rest % QZ, r0 = (rest / QZ) * QY;
        const R* colp = T + (q * TH + r0) * TW + c + HL;
        R acc[QY];
#pragma unroll
        for (int i = 0; i < QY; ++i) acc[i] = R(0);
#pragma unroll
        for (int k = 0; k < QY + W - 1; ++k) {
          if (KC > 0 && k > 0 && k % KC == 0) __builtin_amdgcn_sched_barrier(0);
          const R v = colp[k * TW];
#pragma unroll
          for (int i = 0; i < QY; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
#pragma unroll
        for (int i = 0; i < QY; ++i) SY[(q * TY + r0 + i) * SW + c] = acc[i];
      }
      __syncthreads();
      // ---- combine, epilogue, store
#pragma unroll
      for (int q = 0; q < QZ; ++q) {
        const int z = z0 + q;
        if (!ME || z < ze) {
          R zt = R(0);
#pragma unroll
          for (int k = 0; k < W; ++k) zt = fmar(win[q + k], tap<R, F>(k), zt);
          const R xc = win[HL + q];
          const bool has_x =
              out_ok && coff >= 0 && (zper || (unsigned)(z - zoff) < (unsigned)zlim) && (ME || z < ze);
          R t0 = SX[(q * TY + ty) * SW + tx] * p.a0;
          const R t1 = SY[(q * TY + ty) * SW + tx] * p.a1;
          const R t2 = zt * p.a2;
          if (want_dot && has_x) {
            d0 = fmar((double)xc, (double)t0, d0);
            d1 = fmar((double)xc, (double)t1, d1);
            d2 = fmar((double)xc, (double)t2, d2);
          }
          if ((p.flags & F_AX) && has_x) t0 = fmar(xc, p.a_x, t0);
          R tot = t1 + t0;
          tot = t2 + tot;
          // streaming store: the output is not read again by this kernel, and
          // keeping it out of L2 leaves room for the halo values that
          // neighbouring tiles re-read (5-8% faster at 256^3)
          if constexpr (ME) {
            if (out_ok) {
              R* yp = ya + (ocol + zo);
              if (p.flags & F_AY) tot = fmar(p.a_y, *yp, tot);
              __builtin_nontemporal_store(tot, yp);
            }
          } else {
            plane_st<R, true>(span_rsrc(ya + zo, z < ze ? syo : 0),
                              out_ok ? (int)(ocol * (long)sizeof(R)) : kOffNone, tot);
          }
          zo += syo;
        }
      }
      // slide the window by QZ steps
#pragma unroll
      for (int s = 0; s < W - 1; ++s) win[s] = win[s + QZ];
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
  (void)HR;
}

}  // namespace lc

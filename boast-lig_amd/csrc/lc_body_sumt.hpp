// lc_body_sumt.hpp -- device body of the trailing-window kinetic-form kernel
// (internal).  Same arithmetic and summation order as sum3d2_body, so both
// give identical bits; the geometry is Sum2Geo's.
//
// In the two-plane kernel a lane loads plane z's column value when its z
// window first needs it, 14 planes before the tile of plane z is built, and
// the halo values of plane z one step before the tile.  The neighbouring
// tiles' halo reads of the lines a column load fetched then come ~14 steps
// later, after the XCD's L2 has turned over, and miss again: the kernel reads
// about twice its input.  Here a plane's column values and halo values are
// loaded in the same step (one step ahead of its tile), so the tiles of an
// XCD fetch each line about once, and the z window trails the tile:
//
//   * step k builds the tile of planes t .. t + QZ - 1 (t = zb + k QZ) and
//     appends their column values to the window, which then holds planes
//     t - 28 .. t + QZ - 1;
//   * the x/y-term of plane t + q (t1 + t0, with the a_x term and the
//     energies d0, d1) waits 14 planes in a register queue `pend`;
//   * the outputs of the step are planes t - 14 .. t - 14 + QZ - 1: z-term
//     from the window, plus the queued x/y-term, epilogue, store;
//   * after the chunk's last tile, 14 more planes of column values (no tile,
//     no barrier) complete the chunk's last outputs.
// The queue costs 14 registers per lane (28 VGPRs in FP64), so the variants
// use 16 x 16 tiles of 4 waves, three blocks per CU.
#pragma once

#include "lc_body_sum2.hpp"

namespace lc {

// PH: steps of planes (column and halo values) in flight; the step loop is
// unrolled PH times so that every in-flight register has a static slot.
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int PH = 1>
__device__ __forceinline__ void sumt_body(const FusedArgs<R>& p, int bx, int by, int zc, int arr, long bid,
                                          unsigned char* lds) {
  using G = Sum2Geo<R, F, TX, TY, QX, QY, QZ>;
  constexpr int W = G::W, HL = G::HL, HR = G::HR, TW = G::TW, TH = G::TH, SW = G::SW;
  constexpr int NT = G::NT, NHP = G::NHP, NHL = G::NHL, NHT = G::NHT, WS = G::WS;
  constexpr int NQ = HR;  // x/y-terms waiting for their z window: 14 planes
  static_assert(HL == HR && NQ % QZ == 0, "symmetric filter, whole steps in the queue");
  R* T = reinterpret_cast<R*>(lds);                                         // [QZ][TH][TW]
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
  // column and halo values of the same planes come from one cursor
  ZCursor zcur(p.d2, zb - HL);
  auto col_next = [&]() -> R {
    const R v = plane_ld<R>(plane_rsrc(xa, sxy, zcur.src()), cboff);
    zcur.next();
    return v;
  };
  // planes t .. t + QZ - 1 (column values and halos; the cursor is at t)
  auto load_step = [&](R* col, R* halo) {
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
      const __amdgpu_buffer_rsrc_t rs = plane_rsrc(xa, sxy, zcur.src());
      zcur.next();
      col[q] = plane_ld<R>(rs, cboff);
#pragma unroll
      for (int j = 0; j < NHL; ++j) halo[q * NHL + j] = plane_ld<R>(rs, hoff[j]);
    }
  };
  const bool zper = p.d2.bc == PER;
  const int zoff = p.d2.bc == SHRINK ? p.d2.ox : 0, zlim = p.d2.bc == SHRINK ? p.d2.nx : p.d2.n;
  auto has_x = [&](int z) { return out_ok && coff >= 0 && (zper || (unsigned)(z - zoff) < (unsigned)zlim); };
  const long ocol = (X0 + tx - p.d0.oy) + (long)p.d0.ny * (Y0 + ty - p.d1.oy);
  const long syo = (long)p.d0.ny * p.d1.ny;
  long zo = (long)(zb - p.d2.oy) * syo;  // offset of the next output plane

  // ---- window: planes t - 28 .. t + QZ - 1 once step t's columns are in;
  // before the first step it holds planes zb - 14 .. zb - 1 in its top 14
  // slots (the lower slots leave the window before any output reads them)
  R win[WS];
#pragma unroll
  for (int s = 0; s < HL; ++s) win[s] = R(0);
#pragma unroll
  for (int s = 0; s < HL; ++s) win[HL + s] = col_next();  // planes zb - 14 .. zb - 1
  R pend[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) pend[i] = R(0);
  R ncol[PH][QZ], hreg[PH][NHT];
#pragma unroll
  for (int h = 0; h < PH; ++h) load_step(ncol[h], hreg[h]);  // planes zb .. zb + PH QZ - 1

  const bool want_dot = (p.flags & F_DOT) && arr == p.narr - 1;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;

  // the outputs of the step whose newest plane is tn (planes tn - 14 + q)
  // (no contraction of the epilogue's products and sums: the same
  // operations as the two-plane kernel, so both give the same bits)
  auto emit = [&](int tn) {
#pragma clang fp contract(off)
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
      const int z = tn - HR + q;
      if (z >= zb && z < ze) {
        R zt = R(0);
#pragma unroll
        for (int k = 0; k < W; ++k) zt = fmar(win[q + k], tap<R, F>(k), zt);
        const R t2 = zt * p.a2;
        if (want_dot && has_x(z)) d2 = fmar((double)win[HL + q], (double)t2, d2);
        R tot = t2 + pend[q];
        if (out_ok) {
          R* yp = ya + (ocol + zo);
          if (p.flags & F_AY) tot = fmar(p.a_y, *yp, tot);
          // streaming store: the output is not read again by this kernel
          __builtin_nontemporal_store(tot, yp);
        }
        zo += syo;
      }
    }
  };

  for (int tu = zb; tu < ze; tu += PH * QZ) {
#pragma unroll
    for (int u = 0; u < PH; ++u) {
      const int tb = tu + u * QZ;  // first plane of the step
      if (PH > 1 && tb >= ze) break;  // uniform
      // window gains planes tb .. tb + QZ - 1; their tile goes to LDS
#pragma unroll
      for (int q = 0; q < QZ; ++q) win[W - 1 + q] = ncol[u][q];
#pragma unroll
      for (int q = 0; q < QZ; ++q) T[(q * TH + ty + HL) * TW + tx + HL] = ncol[u][q];
#pragma unroll
      for (int q = 0; q < QZ; ++q)
#pragma unroll
        for (int j = 0; j < NHL; ++j) T[q * TH * TW + hslot[j]] = hreg[u][q * NHL + j];
      load_step(ncol[u], hreg[u]);  // PH steps ahead (past the chunk: only columns are used)
      __syncthreads();
      // ---- one task per lane: x-outputs of a row or y-outputs of a column
      if (t < G::NXT) {
        const int r = t % (QZ * TY), c0 = (t / (QZ * TY)) * QX;
        const int q = r / TY, rr = r % TY;
        const R* row = T + (q * TH + rr + HL) * TW + c0;
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
          const R v = colp[k * TW];
#pragma unroll
          for (int i = 0; i < QY; ++i)
            if (k - i >= 0 && k - i < W) acc[i] = fmar(v, tap<R, F>(k - i), acc[i]);
        }
#pragma unroll
        for (int i = 0; i < QY; ++i) SY[(q * TY + r0 + i) * SW + c] = acc[i];
      }
      __syncthreads();
      // ---- x/y-terms of planes tb + q (energies d0, d1 and a_x now: the
      // centre value is the window's newest), then the outputs of planes
      // tb - 14 + q, then the queue and the window slide by QZ
      R xy[QZ];
#pragma unroll
      for (int q = 0; q < QZ; ++q) {
#pragma clang fp contract(off)
        const int z = tb + q;
        const R xc = win[W - 1 + q];
        R t0 = SX[(q * TY + ty) * SW + tx] * p.a0;
        const R t1 = SY[(q * TY + ty) * SW + tx] * p.a1;
        const bool hx = z < ze && has_x(z);
        if (want_dot && hx) {
          d0 = fmar((double)xc, (double)t0, d0);
          d1 = fmar((double)xc, (double)t1, d1);
        }
        if ((p.flags & F_AX) && hx) t0 = fmar(xc, p.a_x, t0);
        xy[q] = t1 + t0;
      }
      emit(tb);
#pragma unroll
      for (int i = 0; i < NQ - QZ; ++i) pend[i] = pend[i + QZ];
#pragma unroll
      for (int q = 0; q < QZ; ++q) pend[NQ - QZ + q] = xy[q];
#pragma unroll
      for (int s = 0; s < W - 1; ++s) win[s] = win[s + QZ];
    }
  }
  // ---- drain: 14 more planes of column values, no tile; the planes in
  // flight are dropped and the drain loads its own, from the first plane
  // after the last tile
  const int tend = zb + (ze - zb + QZ - 1) / QZ * QZ;
  zcur = ZCursor(p.d2, tend);
  R dcol[QZ];
#pragma unroll
  for (int q = 0; q < QZ; ++q) dcol[q] = col_next();
  for (int td = tend; td < ze + HR; td += QZ) {
#pragma unroll
    for (int q = 0; q < QZ; ++q) win[W - 1 + q] = dcol[q];
#pragma unroll
    for (int q = 0; q < QZ; ++q) dcol[q] = col_next();
    emit(td);
#pragma unroll
    for (int i = 0; i < NQ - QZ; ++i) pend[i] = pend[i + QZ];
#pragma unroll
    for (int s = 0; s < W - 1; ++s) win[s] = win[s + QZ];
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

}  // namespace lc

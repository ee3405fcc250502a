// lc_fused_impl.hpp -- kernels, launchers and variant dispatch of the fused
// 3D convolution operators (internal).  Included once per filter by
// lc_fused_f*.hip, which instantiate it, so the variants compile in parallel;
// lc_fused.hip holds the filter-independent host code.
#pragma once

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lc_body_conv.hpp"
#include "lc_body_sum2.hpp"
#include "lc_body_sep2.hpp"
#include "lc_body_sepp.hpp"
#include "lc_body_sepr.hpp"
#include "lc_body_sepq.hpp"
#include "lc_body_sumt.hpp"
#include "lc_body_sump.hpp"

namespace lc {

hipError_t launch_batch_dots(const double* part, const int* first, double* const* dots, int count,
                             hipStream_t st);

// ====================================================================== kernels
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD, int WPE>
__global__ __launch_bounds__(TX * TY / CPT) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[sum3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, HD>()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sum3d_body<R, F, TX, TY, CPT, QX, BZ, HD>(p, bx, by, zc, arr,
                                           (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD, int WPE>
__global__ __launch_bounds__(TX * TY / CPT) __attribute__((amdgpu_waves_per_eu(WPE))) void sep3d_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[sep3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, AD>()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sep3d_body<R, F, TX, TY, CPT, QX, BZ, AD>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD, int WPE>
__global__ __launch_bounds__(TX * TY / CPT) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  const int bx = tile % p.tiles_x, by = tile / p.tiles_x;
  __shared__ __align__(16) unsigned char lds[sum3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, HD>()];
  sum3d_body<R, F, TX, TY, CPT, QX, BZ, HD>(p, bx, by, zc, 0, (long)local, lds);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD, int WPE>
__global__ __launch_bounds__(TX * TY / CPT) __attribute__((amdgpu_waves_per_eu(WPE))) void sep3d_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[sep3d_lds_bytes<R, F, TX, TY, CPT, QX, BZ, AD>()];
  sep3d_body<R, F, TX, TY, CPT, QX, BZ, AD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// ====================================================================== launch
template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int HD, int WPE>
hipError_t launch_sum3d(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                               R a_y, double* dot, hipStream_t st, double* scratch,
                               size_t scratch_bytes) {
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
  p.a_x = a_x; p.a_y = a_y;
  p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
  p.nparts = tiles * zch;
  if (dot) {
    if (scratch_bytes < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
    p.part = scratch;
  }
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sum3d_fused<R, F, TX, TY, CPT, QX, BZ, HD, WPE><<<grid, TX * TY / CPT, 0, st>>>(p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
}

template <typename R, class F, int TX, int TY, int CPT, int QX, int BZ, int AD, int WPE>
hipError_t launch_sep3d(const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                               hipStream_t st, const R* pot, R* rho) {
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = a; p.a_y = a_y;
  p.flags = (a != R(1) ? F_A : 0) | (a_y != R(0) ? F_AY : 0) | (pot ? F_POT : 0);
  p.pot = pot; p.rho = rho;
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sep3d_fused<R, F, TX, TY, CPT, QX, BZ, AD, WPE><<<grid, TX * TY / CPT, 0, st>>>(p);
  return hipGetLastError();
}

// ---- two-plane kinetic form (lc_body_sum2.hpp)
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int KC>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d2_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  const long bid = (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc;
  if (p.flags & F_AY)
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, 1, true>(p, bx, by, zc, arr, bid, lds);
  else
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, 1, false>(p, bx, by, zc, arr, bid, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int KC>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d2_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  if (p.flags & F_AY)
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, 1, true>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local, lds);
  else
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, 1, false>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local,
                                                         lds);
}

// the same with the column planes PW > 1 steps in flight
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int KC, int PW>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d2w_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  const long bid = (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc;
  if (p.flags & F_AY)
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, PW, true>(p, bx, by, zc, arr, bid, lds);
  else
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, PW, false>(p, bx, by, zc, arr, bid, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int KC, int PW>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sum3d2w_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  if (p.flags & F_AY)
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, PW, true>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local,
                                                        lds);
  else
    sum3d2_body<R, F, TX, TY, QX, QY, QZ, KC, PW, false>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local,
                                                         lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int KC, int ZC, int PW = 1>
hipError_t launch_sum3d2(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                         double* dot, hipStream_t st, double* scratch, size_t scratch_bytes) {
  if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  if (ZC == 0) {
    p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  } else if (ZC < 0) {
    // at most -ZC blocks on any CU, at least one where the grid allows
    const long per = tiles * std::max(narr, 1);
    long chunks = (long)(-ZC) * num_cu() / per;
    if (chunks * per < num_cu()) chunks = (num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / 16)));
    p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
    p.zchunk = (p.zchunk + QZ - 1) / QZ * QZ;
  } else {
    // a z chunk only costs the window prefill (W - 1 column loads per lane):
    // chunks down to ZC planes put ~8 blocks on every CU, and down to 16
    // planes when that would leave CUs without a block
    const long per = tiles * std::max(narr, 1);
    long chunks = (8L * num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / ZC)));
    if (chunks * per < num_cu())
      chunks = std::max(chunks, std::min<long>((num_cu() + per - 1) / per, std::max(1, p.d2.nout / 16)));
    p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
    p.zchunk = (p.zchunk + QZ - 1) / QZ * QZ;
  }
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
  p.a_x = a_x; p.a_y = a_y;
  p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
  p.nparts = tiles * zch;
  if (dot) {
    if (scratch_bytes < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
    p.part = scratch;
  }
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  if constexpr (PW == 1)
    sum3d2_fused<R, F, TX, TY, QX, QY, QZ, WPE, KC><<<grid, TX * TY, 0, st>>>(p);
  else
    sum3d2w_fused<R, F, TX, TY, QX, QY, QZ, WPE, KC, PW><<<grid, TX * TY, 0, st>>>(p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
}

// ---- trailing-window kinetic form (lc_body_sumt.hpp)
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int PH>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sumt_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sumt_body<R, F, TX, TY, QX, QY, QZ, PH>(p, bx, by, zc, arr,
                                          (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int PH>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sumt_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[Sum2Geo<R, F, TX, TY, QX, QY, QZ>::lds_bytes()];
  sumt_body<R, F, TX, TY, QX, QY, QZ, PH>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local, lds);
}

template <class F, int ZC, int Q>
int sum_zchunk(long tiles, int nz, int narr);

// the trailing window needs a symmetric filter (the queue holds UP planes)
template <class F>
constexpr bool sumt_filter() { return -F::LO == F::UP && F::UP % 2 == 0; }

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int ZC, int PH>
hipError_t launch_sumt(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                       double* dot, hipStream_t st, double* scratch, size_t scratch_bytes) {
  if constexpr (!sumt_filter<F>()) {
    return hipErrorInvalidValue;
  } else {
    if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
    FusedArgs<R> p{};
    p.x = x; p.y = y;
    p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
    p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
    p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
    p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
    p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
    p.narr = narr;
    p.tiles_x = (p.d0.nout + TX - 1) / TX;
    p.tiles_y = (p.d1.nout + TY - 1) / TY;
    const long tiles = (long)p.tiles_x * p.tiles_y;
    p.zchunk = sum_zchunk<F, ZC, QZ>(tiles, p.d2.nout, narr);
    const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
    p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
    p.a_x = a_x; p.a_y = a_y;
    p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
    p.nparts = tiles * zch;
    if (dot) {
      if (scratch_bytes < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
      p.part = scratch;
    }
    if (narr == 0) return hipSuccess;
    dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
    sumt_fused<R, F, TX, TY, QX, QY, QZ, WPE, PH><<<grid, TX * TY, 0, st>>>(p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !dot) return e;
    return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
  }
}

// ---- pipelined kinetic form (lc_body_sump.hpp)
template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int U, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sump_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[SumpGeo<R, F, TX, TY, QX, QY, PD, U>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sump_body<R, F, TX, TY, QX, QY, PD, U>(p, bx, by, zc, arr,
                                          (bx + (long)p.tiles_x * by) + (long)gridDim.x * zc, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int U, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sump_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[SumpGeo<R, F, TX, TY, QX, QY, PD, U>::lds_bytes()];
  sump_body<R, F, TX, TY, QX, QY, PD, U>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, (long)local, lds);
}

// z chunk of the kinetic forms that read planes through buffer records; ZC as
// in launch_sum3d2 (0: pick_zchunk, < 0: at most -ZC blocks on any CU, > 0:
// chunks down to ZC planes for ~8 blocks per CU), rounded to a multiple of Q
template <class F, int ZC, int Q>
int sum_zchunk(long tiles, int nz, int narr) {
  if (ZC == 0) return pick_zchunk(tiles, nz, std::max(narr, 1), F::N, num_cu());
  const long per = tiles * std::max(narr, 1);
  long chunks;
  if (ZC < 0) {
    chunks = (long)(-ZC) * num_cu() / per;
    if (chunks * per < num_cu()) chunks = (num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, nz / 16)));
  } else {
    chunks = (8L * num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, nz / ZC)));
    if (chunks * per < num_cu())
      chunks = std::max(chunks, std::min<long>((num_cu() + per - 1) / per, std::max(1, nz / 16)));
  }
  const int zc = (int)((nz + chunks - 1) / chunks);
  return (zc + Q - 1) / Q * Q;
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int U, int WPE, int ZC>
hipError_t launch_sump(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                       double* dot, hipStream_t st, double* scratch, size_t scratch_bytes) {
  if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
  FusedArgs<R> p{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  p.zchunk = sum_zchunk<F, ZC, 1>(tiles, p.d2.nout, narr);
  const int zch = (p.d2.nout + p.zchunk - 1) / p.zchunk;
  p.a0 = (R)a[0]; p.a1 = (R)a[1]; p.a2 = (R)a[2];
  p.a_x = a_x; p.a_y = a_y;
  p.flags = (a_x != R(0) ? F_AX : 0) | (a_y != R(0) ? F_AY : 0) | (dot ? F_DOT : 0);
  p.nparts = tiles * zch;
  if (dot) {
    if (scratch_bytes < sizeof(double) * 3 * (size_t)p.nparts) return hipErrorInvalidValue;
    p.part = scratch;
  }
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)tiles, (unsigned)zch, (unsigned)narr);
  sump_fused<R, F, TX, TY, QX, QY, PD, U, WPE><<<grid, TX * TY, 0, st>>>(p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  return sum_partial_sets(scratch, p.nparts, dot, 3, p.nparts, st);
}

// ---- two-plane separable form (lc_body_sep2.hpp)
template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int PD>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sep3d2_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[Sep2Geo<R, F, TX, TY, QX, QY, QZ, PD>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sep3d2_body<R, F, TX, TY, QX, QY, QZ, PD>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int PD>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sep3d2_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[Sep2Geo<R, F, TX, TY, QX, QY, QZ, PD>::lds_bytes()];
  sep3d2_body<R, F, TX, TY, QX, QY, QZ, PD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// arguments and z chunk of the separable forms that read planes through
// buffer records (two-plane and pipelined); returns the number of z chunks
template <typename R, class F, int TX, int TY, int ZC>
int sep_plane_args(const Box3& b, int narr, const R* x, R* y, R a, R a_y, const R* pot, R* rho,
                   FusedArgs<R>& p) {
  p = FusedArgs<R>{};
  p.x = x; p.y = y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = narr;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  const long tiles = (long)p.tiles_x * p.tiles_y;
  if (ZC == 0) {
    p.zchunk = pick_zchunk(tiles, p.d2.nout, std::max(narr, 1), F::N, num_cu());
  } else if (ZC < 0) {
    // at most -ZC blocks on any CU: chunks = floor(-ZC CUs / tiles), so no
    // CU holds one block more than the others (balance beats occupancy: 162
    // tiles of 284 planes run 486 blocks of 95 planes 6% faster than 648 of
    // 71), yet at least one block per CU where the grid allows
    const long per = tiles * std::max(narr, 1);
    long chunks = (long)(-ZC) * num_cu() / per;
    if (chunks * per < num_cu()) chunks = (num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / (2 * F::N))));
    p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
  } else {
    // a chunk recomputes W - 1 filtered planes: shorter chunks buy occupancy
    const long per = tiles * std::max(narr, 1);
    long chunks = (4L * num_cu() + per - 1) / per;
    chunks = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / ZC)));
    if (chunks * per < num_cu())
      chunks = std::max(chunks, std::min<long>((num_cu() + per - 1) / per, std::max(1, p.d2.nout / (2 * F::N))));
    p.zchunk = (int)((p.d2.nout + chunks - 1) / chunks);
  }
  p.a0 = a; p.a_y = a_y;
  p.flags = (a != R(1) ? F_A : 0) | (a_y != R(0) ? F_AY : 0) | (pot ? F_POT : 0);
  p.pot = pot; p.rho = rho;
  return (p.d2.nout + p.zchunk - 1) / p.zchunk;
}

template <typename R, class F, int TX, int TY, int QX, int QY, int QZ, int WPE, int ZC, int PD>
hipError_t launch_sep3d2(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                         const R* pot, R* rho) {
  if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
  FusedArgs<R> p;
  const int zch = sep_plane_args<R, F, TX, TY, ZC>(b, narr, x, y, a, a_y, pot, rho, p);
  if (narr == 0) return hipSuccess;
  dim3 grid((unsigned)(p.tiles_x * p.tiles_y), (unsigned)zch, (unsigned)narr);
  sep3d2_fused<R, F, TX, TY, QX, QY, QZ, WPE, PD><<<grid, TX * TY, 0, st>>>(p);
  return hipGetLastError();
}

// ---- pipelined separable form (lc_body_sepp.hpp)
template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sepp_fused(
    FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[SeppGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sepp_body<R, F, TX, TY, QX, QY, PD>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sepp_batch(
    const FusedArgs<R>* __restrict__ tab, const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[SeppGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  sepp_body<R, F, TX, TY, QX, QY, PD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

// the pipelined form is built for the 16-tap filters (MagicFilter and its
// transpose) only; other filters report an invalid variant
template <class F>
constexpr bool sepp_filter() { return F::N == 16; }

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE, int ZC>
hipError_t launch_sepp(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                       const R* pot, R* rho) {
  if constexpr (!sepp_filter<F>()) {
    return hipErrorInvalidValue;
  } else {
    if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
    FusedArgs<R> p;
    const int zch = sep_plane_args<R, F, TX, TY, ZC>(b, narr, x, y, a, a_y, pot, rho, p);
    if (narr == 0) return hipSuccess;
    dim3 grid((unsigned)(p.tiles_x * p.tiles_y), (unsigned)zch, (unsigned)narr);
    sepp_fused<R, F, TX, TY, QX, QY, PD, WPE><<<grid, TX * TY, 0, st>>>(p);
    return hipGetLastError();
  }
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
hipError_t launch_sepp_batch(const FusedArgs<R>* tab, const BatchStage* s, hipStream_t st) {
  if constexpr (!sepp_filter<F>()) {
    return hipErrorInvalidValue;
  } else {
    sepp_batch<R, F, TX, TY, QX, QY, PD, WPE><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count);
    return hipSuccess;
  }
}

// ---- ring-owning separable form (lc_body_sepr.hpp), 16-tap filters only
template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__((SeprGeo<R, F, TX, TY, QX, QY, PD>::NT))
__attribute__((amdgpu_waves_per_eu(WPE))) void sepr_fused(FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[SeprGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  sepr_body<R, F, TX, TY, QX, QY, PD>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__((SeprGeo<R, F, TX, TY, QX, QY, PD>::NT))
__attribute__((amdgpu_waves_per_eu(WPE))) void sepr_batch(const FusedArgs<R>* __restrict__ tab,
                                                           const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[SeprGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  sepr_body<R, F, TX, TY, QX, QY, PD>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE, int ZC>
hipError_t launch_sepr(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                       const R* pot, R* rho) {
  if constexpr (!sepp_filter<F>()) {
    return hipErrorInvalidValue;
  } else {
    if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
    FusedArgs<R> p;
    const int zch = sep_plane_args<R, F, TX, TY, ZC>(b, narr, x, y, a, a_y, pot, rho, p);
    if (narr == 0) return hipSuccess;
    dim3 grid((unsigned)(p.tiles_x * p.tiles_y), (unsigned)zch, (unsigned)narr);
    sepr_fused<R, F, TX, TY, QX, QY, PD, WPE>
        <<<grid, SeprGeo<R, F, TX, TY, QX, QY, PD>::NT, 0, st>>>(p);
    return hipGetLastError();
  }
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
hipError_t launch_sepr_batch(const FusedArgs<R>* tab, const BatchStage* s, hipStream_t st) {
  if constexpr (!sepp_filter<F>()) {
    return hipErrorInvalidValue;
  } else {
    sepr_batch<R, F, TX, TY, QX, QY, PD, WPE>
        <<<s->blocks, SeprGeo<R, F, TX, TY, QX, QY, PD>::NT, 0, st>>>(tab, s->first, s->count);
    return hipSuccess;
  }
}

// ---- 128-bit-LDS separable form (lc_body_sepq.hpp), 16-tap filters only
template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__((SepqGeo<R, F, TX, TY, QX, QY, PD>::NT))
__attribute__((amdgpu_waves_per_eu(WPE))) void sepq_fused(FusedArgs<R> p) {
  __shared__ __align__(16) unsigned char lds[SepqGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  int bx, by, zc, arr;
  block_coords(p.tiles_x, bx, by, zc, arr);
  if (p.flags & (F_AY | F_POT))
    sepq_body<R, F, TX, TY, QX, QY, PD, true>(p, bx, by, zc, arr, lds);
  else
    sepq_body<R, F, TX, TY, QX, QY, PD, false>(p, bx, by, zc, arr, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
__global__ __launch_bounds__((SepqGeo<R, F, TX, TY, QX, QY, PD>::NT))
__attribute__((amdgpu_waves_per_eu(WPE))) void sepq_batch(const FusedArgs<R>* __restrict__ tab,
                                                           const int* __restrict__ first, int count) {
  const int lin = xcd_linear();
  const int o = batch_item(first, count, lin);
  const FusedArgs<R> p = tab[o];
  const int local = lin - first[o];
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = local % tiles, zc = local / tiles;
  __shared__ __align__(16) unsigned char lds[SepqGeo<R, F, TX, TY, QX, QY, PD>::lds_bytes()];
  if (p.flags & (F_AY | F_POT))
    sepq_body<R, F, TX, TY, QX, QY, PD, true>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
  else
    sepq_body<R, F, TX, TY, QX, QY, PD, false>(p, tile % p.tiles_x, tile / p.tiles_x, zc, 0, lds);
}

template <typename R, class F, int TX, int TY, int QX, int QY>
constexpr bool sepq_ok() { return sepp_filter<F>() && sepq_shape_ok<R, TX, TY, QX, QY>(); }

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE, int ZC>
hipError_t launch_sepq(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                       const R* pot, R* rho) {
  if constexpr (!sepq_ok<R, F, TX, TY, QX, QY>()) {
    return hipErrorInvalidValue;
  } else {
    if (!plane_fits(b, sizeof(R))) return hipErrorInvalidValue;
    FusedArgs<R> p;
    const int zch = sep_plane_args<R, F, TX, TY, ZC>(b, narr, x, y, a, a_y, pot, rho, p);
    if (narr == 0) return hipSuccess;
    dim3 grid((unsigned)(p.tiles_x * p.tiles_y), (unsigned)zch, (unsigned)narr);
    sepq_fused<R, F, TX, TY, QX, QY, PD, WPE>
        <<<grid, SepqGeo<R, F, TX, TY, QX, QY, PD>::NT, 0, st>>>(p);
    return hipGetLastError();
  }
}

template <typename R, class F, int TX, int TY, int QX, int QY, int PD, int WPE>
hipError_t launch_sepq_batch(const FusedArgs<R>* tab, const BatchStage* s, hipStream_t st) {
  if constexpr (!sepq_ok<R, F, TX, TY, QX, QY>()) {
    return hipErrorInvalidValue;
  } else {
    sepq_batch<R, F, TX, TY, QX, QY, PD, WPE>
        <<<s->blocks, SepqGeo<R, F, TX, TY, QX, QY, PD>::NT, 0, st>>>(tab, s->first, s->count);
    return hipSuccess;
  }
}

// ====================================================================== variants
// Compiled variants of the two fused forms: tile TX x TY, CPT rows per lane,
// QX outputs per x-task, BZ planes per window slide, HD / AD prefetch depth.
// Every entry of these tables also has a batched form; indices 5..7 are the
// wave-specialised kernels of lc_fused_ws.hip (single calls only).  The
// variant of a call comes from the tuner's table for its exact shape, else
// LIBCONV_{SUM3D,SEP3D}_VARIANT, else a per-form default (lc_fused.hip).
#define LC_SUM_VARIANTS(X)                                                          \
  X(0, 32, 16, 2, 2, 4, 2, 1) X(1, 32, 16, 2, 2, 4, 1, 1) X(2, 32, 8, 2, 2, 4, 2, 1)    \
  X(3, 64, 8, 2, 4, 4, 2, 1) X(4, 32, 16, 2, 2, 2, 2, 1) X(8, 32, 16, 1, 2, 2, 1, 1)    \
  X(9, 32, 8, 1, 2, 2, 1, 1) X(10, 64, 8, 1, 4, 2, 1, 1) X(11, 32, 16, 1, 4, 4, 1, 1)  \
  X(12, 32, 16, 2, 2, 4, 1, 2) X(13, 32, 16, 1, 2, 2, 1, 4) X(14, 32, 16, 1, 4, 2, 1, 3) \
  X(15, 32, 8, 1, 2, 2, 1, 4)
#define LC_SEP_VARIANTS(X)                                                          \
  X(0, 32, 16, 2, 4, 4, 4, 1) X(1, 32, 16, 2, 4, 4, 1, 1) X(2, 32, 8, 2, 4, 4, 2, 1)    \
  X(3, 64, 8, 2, 4, 4, 2, 1) X(4, 32, 16, 2, 4, 4, 2, 1) X(8, 32, 16, 1, 4, 2, 1, 1)    \
  X(9, 32, 8, 1, 4, 2, 1, 1) X(10, 64, 8, 1, 4, 2, 1, 1) X(11, 32, 16, 1, 4, 4, 1, 1)  \
  X(12, 32, 16, 2, 4, 4, 1, 2) X(13, 64, 16, 1, 4, 2, 1, 2) X(14, 32, 32, 2, 4, 2, 1, 2) \
  X(15, 64, 32, 2, 4, 2, 1, 1)
#define LC_WS_VARIANTS(X) X(5, 32, 8) X(6, 64, 8) X(7, 32, 16)
// two-plane kinetic form (sum only): tile, QX / QY outputs per task, QZ planes
// per step, waves per EU
#define LC_SUM2_VARIANTS(X)                                                                      \
  X(16, 32, 16, 4, 4, 2, 4, 0, 0) X(17, 32, 16, 4, 4, 2, 2, 0, 0) X(18, 64, 8, 4, 4, 2, 4, 0, 0)       \
  X(19, 32, 16, 8, 8, 4, 2, 0, 0) X(20, 32, 16, 4, 4, 2, 4, 8, 16) X(21, 32, 16, 8, 8, 4, 2, 8, 32)    \
  X(22, 16, 16, 4, 4, 2, 4, 8, 16) X(23, 16, 16, 4, 4, 2, 2, 0, 16) X(24, 32, 8, 4, 4, 2, 3, 0, 16)    \
  X(25, 32, 8, 4, 4, 2, 2, 0, 16) X(26, 16, 16, 4, 4, 2, 2, 0, 64) X(27, 16, 16, 4, 4, 2, 4, 8, 64)    \
  X(28, 32, 16, 8, 8, 2, 2, 0, 0) X(29, 32, 16, 8, 8, 2, 2, 0, 32) X(30, 32, 16, 8, 8, 2, 2, 8, 32)   \
  X(31, 32, 16, 8, 8, 2, 2, 0, -2) X(32, 32, 16, 8, 8, 4, 2, 8, -2) X(33, 32, 16, 8, 8, 2, 1, 0, -1)    \
  X(34, 32, 16, 8, 8, 4, 1, 0, -1) X(35, 16, 16, 4, 4, 2, 1, 0, -1) X(36, 32, 16, 8, 8, 2, 2, 0, -3)
// two-plane separable form (sep only): tile, QX / QY outputs per task, QZ
// planes per step, waves per EU, shortest z chunk (0: 4 W, ~2 blocks per CU;
// else chunks down to ZC planes for ~4 blocks per CU), prefetch depth in steps
// (+4: P shares A's LDS, see Sep2Geo)
#define LC_SEP2_VARIANTS(X)                                                                        \
  X(16, 32, 16, 4, 2, 2, 2, 0, 1) X(17, 32, 16, 4, 4, 2, 2, 0, 1) X(18, 32, 16, 4, 2, 2, 4, 0, 1)        \
  X(19, 32, 8, 8, 2, 2, 3, 0, 1) X(20, 64, 8, 8, 2, 2, 2, 0, 1) X(21, 32, 32, 4, 4, 2, 2, 0, 1)          \
  X(22, 32, 16, 4, 4, 2, 2, 32, 1) X(23, 32, 16, 4, 4, 2, 3, 32, 1) X(24, 32, 16, 4, 4, 2, 2, 48, 1)     \
  X(25, 32, 16, 4, 4, 2, 2, 0, 2) X(26, 32, 16, 4, 2, 2, 4, 0, 2) X(27, 32, 16, 4, 4, 2, 2, 48, 2)          \
  X(28, 32, 16, 4, 4, 4, 1, 0, 1) X(29, 32, 16, 4, 4, 4, 2, 0, 1) X(30, 32, 16, 4, 4, 4, 2, 48, 1)          \
  X(31, 32, 16, 4, 4, 4, 2, 0, 5) X(32, 32, 16, 4, 4, 4, 2, 32, 5) X(33, 32, 16, 4, 4, 2, 2, 0, 5)           \
  X(34, 32, 16, 4, 4, 2, 2, 32, 5) X(35, 32, 16, 4, 2, 2, 4, 0, 5) X(36, 32, 16, 4, 4, 2, 3, 0, 6)
// pipelined separable form (sep only, 16-tap filters): tile, QX / QY outputs
// per task, prefetch depth in planes, waves per EU, shortest z chunk as above
// (ZC < 0: at most -ZC blocks on any CU)
#define LC_SEPP_VARIANTS(X)                                                                       \
  X(37, 32, 16, 4, 2, 2, 4, 0) X(38, 32, 16, 4, 2, 2, 6, 0) X(39, 32, 16, 4, 4, 2, 4, 0)            \
  X(40, 32, 16, 4, 2, 1, 6, 0) X(41, 32, 16, 4, 2, 4, 4, 0) X(42, 32, 16, 8, 2, 2, 4, 0)            \
  X(43, 32, 32, 4, 2, 2, 4, 0) X(44, 32, 16, 4, 2, 2, 4, 32) X(45, 32, 16, 4, 2, 2, 4, -2)          \
  X(46, 32, 16, 4, 2, 2, 6, -2) X(47, 32, 16, 4, 2, 2, 4, -3) X(48, 32, 16, 4, 2, 2, 4, -4)            \
  X(60, 32, 16, 4, 2, 4, 4, -2) X(61, 32, 16, 4, 2, 8, 2, -2)
// ring-owning separable form (sep only, 16-tap filters): tile, QX / QY
// outputs per x- / y-task, prefetch depth in planes, waves per EU, z chunk
// policy as above
#define LC_SEPR_VARIANTS(X)                                                                       \
  X(49, 32, 16, 4, 2, 2, 2, -2) X(50, 32, 16, 4, 2, 2, 4, -2) X(51, 32, 16, 4, 2, 2, 2, -3)          \
  X(52, 32, 16, 8, 4, 2, 2, -2) X(53, 32, 16, 8, 4, 2, 4, -4) X(54, 64, 16, 4, 2, 2, 1, -1)          \
  X(55, 32, 16, 4, 2, 2, 3, 0) X(56, 32, 16, 4, 2, 4, 2, -2) X(57, 32, 16, 4, 2, 8, 2, -2)          \
  X(58, 32, 16, 8, 4, 4, 2, -2) X(59, 32, 16, 8, 4, 8, 2, -2)
// pipelined kinetic form (sum only): tile, QX / QY outputs per task, halo
// prefetch depth in planes, steps per window slide, waves per EU, z chunk
// policy as above
#define LC_SUMP_VARIANTS(X)                                                                           \
  X(37, 32, 16, 2, 2, 2, 4, 4, -2) X(38, 32, 16, 4, 4, 2, 4, 4, -2) X(39, 32, 16, 2, 2, 2, 4, 2, -1)     \
  X(40, 32, 16, 4, 4, 2, 4, 2, -1) X(41, 32, 16, 2, 2, 2, 4, 3, -2) X(42, 32, 16, 2, 2, 2, 8, 4, -2)     \
  X(43, 32, 16, 4, 4, 2, 4, 4, -3) X(44, 32, 8, 2, 2, 2, 4, 4, -4) X(45, 32, 16, 4, 4, 2, 4, 1, -1)
// 128-bit-LDS separable form (sep only, 16-tap filters; QX and QY whole
// 16-byte pieces, so some entries take one precision only): tile, QX / QY,
// prefetch depth, waves per EU, z chunk policy
#define LC_SEPQ_VARIANTS(X)                                                                       \
  X(62, 32, 16, 4, 2, 2, 2, -2) X(63, 32, 16, 4, 2, 2, 4, -2) X(64, 32, 16, 8, 4, 2, 2, -2)          \
  X(65, 32, 16, 8, 4, 2, 4, -4) X(66, 64, 16, 8, 4, 2, 2, -2) X(67, 32, 16, 4, 2, 2, 2, -3)          \
  X(68, 32, 32, 8, 4, 2, 2, -2) X(69, 32, 32, 4, 2, 2, 1, -1) X(70, 32, 24, 4, 2, 2, 1, -1)          \
  X(71, 64, 32, 8, 4, 2, 1, -1) X(72, 32, 32, 8, 4, 2, 2, -3)
// two-plane kinetic form with the column planes PW steps in flight: as
// LC_SUM2_VARIANTS, then PW
#define LC_SUM2W_VARIANTS(X) \
  X(46, 32, 16, 8, 8, 2, 4, 0, -2, 2) X(47, 32, 16, 8, 8, 2, 2, 0, -2, 2) X(48, 32, 16, 8, 8, 4, 2, 8, -2, 2) \
  X(49, 32, 32, 8, 8, 2, 4, 0, -1, 2) X(50, 32, 32, 8, 8, 2, 4, 0, -1, 1)
// trailing-window kinetic form (sum only, symmetric filters): tile, QX / QY
// outputs per task, QZ planes per step, waves per EU, z chunk policy, steps
// of planes in flight
#define LC_SUMT_VARIANTS(X) \
  X(51, 16, 16, 8, 8, 2, 3, -3, 1) X(52, 16, 16, 4, 4, 2, 3, -3, 1) X(53, 32, 16, 8, 8, 2, 2, -2, 1) \
  X(54, 16, 16, 4, 4, 2, 3, -3, 2) X(55, 32, 16, 8, 8, 2, 2, -2, 2) X(56, 16, 16, 4, 4, 2, 2, -3, 2)
constexpr int kFusedVariantCount = 57;
constexpr int kSepVariantCount = 73;

template <typename R, class F>
hipError_t sum3d_variant(int v, const Box3& b, int narr, const R* x, R* y, const double* a,
                                R a_x, R a_y, double* dot, hipStream_t st, double* sc, size_t sb) {
  switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, HD, WPE) \
  case I: return launch_sum3d<R, F, TX, TY, CPT, QX, BZ, HD, WPE>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY) \
  case I: return launch_sum3d_ws<R, F, TX, TY, 4, 4, 2>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb, 4, F::N);
    LC_WS_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC) \
  case I: return launch_sum3d2<R, F, TX, TY, QX, QY, QZ, WPE, KC, ZC>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUM2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, U, WPE, ZC) \
  case I: return launch_sump<R, F, TX, TY, QX, QY, PD, U, WPE, ZC>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUMP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC, PW) \
  case I: return launch_sum3d2<R, F, TX, TY, QX, QY, QZ, WPE, KC, ZC, PW>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUM2W_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PH) \
  case I: return launch_sumt<R, F, TX, TY, QX, QY, QZ, WPE, ZC, PH>(b, narr, x, y, a, a_x, a_y, dot, st, sc, sb);
    LC_SUMT_VARIANTS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename R, class F>
hipError_t sep3d_variant(int v, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                                hipStream_t st, const R* pot, R* rho) {
  switch (v) {
#define X(I, TX, TY, CPT, QX, BZ, AD, WPE) \
  case I: return launch_sep3d<R, F, TX, TY, CPT, QX, BZ, AD, WPE>(b, narr, x, y, a, a_y, st, pot, rho);
    LC_SEP_VARIANTS(X)
#undef X
#define X(I, TX, TY) \
  case I:            \
    return pot ? hipErrorInvalidValue : launch_sep3d_ws<R, F, TX, TY, 4, 4, 2>(b, narr, x, y, a, a_y, st, 4, F::N);
    LC_WS_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PD) \
  case I: return launch_sep3d2<R, F, TX, TY, QX, QY, QZ, WPE, ZC, PD>(b, narr, x, y, a, a_y, st, pot, rho);
    LC_SEP2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC) \
  case I: return launch_sepp<R, F, TX, TY, QX, QY, PD, WPE, ZC>(b, narr, x, y, a, a_y, st, pot, rho);
    LC_SEPP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC) \
  case I: return launch_sepr<R, F, TX, TY, QX, QY, PD, WPE, ZC>(b, narr, x, y, a, a_y, st, pot, rho);
    LC_SEPR_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC) \
  case I: return launch_sepq<R, F, TX, TY, QX, QY, PD, WPE, ZC>(b, narr, x, y, a, a_y, st, pot, rho);
    LC_SEPQ_VARIANTS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename R>
bool items_fit(const BatchItem<R>* it, int count) {  // planes within the buffer-load limit
  for (int o = 0; o < count; ++o)
    if (!plane_fits(it[o].b, sizeof(R))) return false;
  return true;
}

template <typename R, class F, int TX, int TY>
hipError_t fused_stage_build(int kind, const BatchItem<R>* it, int count, BatchStage* s) {
  std::vector<FusedArgs<R>> tab(count);
  std::vector<int> first(count + 1, 0);
  std::vector<double*> dots(count, nullptr);
  long tiles_all = 0;
  for (int o = 0; o < count; ++o) {
    const Box3& b = it[o].b;
    FusedArgs<R>& p = tab[o];
    p = FusedArgs<R>{};
    p.x = it[o].x; p.y = it[o].y;
    p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
    p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
    p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
    p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
    p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
    p.narr = 1;
    p.tiles_x = (p.d0.nout + TX - 1) / TX;
    p.tiles_y = (p.d1.nout + TY - 1) / TY;
    p.a0 = (R)it[o].a[0]; p.a1 = (R)it[o].a[1]; p.a2 = (R)it[o].a[2];
    p.a_x = (R)it[o].a_x; p.a_y = (R)it[o].a_y;
    if (kind == BK_SUM)
      p.flags = (p.a_x != R(0) ? F_AX : 0) | (p.a_y != R(0) ? F_AY : 0) | (it[o].dot ? F_DOT : 0);
    else
      p.flags = (p.a0 != R(1) ? F_A : 0) | (p.a_y != R(0) ? F_AY : 0);
    dots[o] = it[o].dot;
    if (it[o].dot) s->any_dot = true;
    tiles_all += (long)p.tiles_x * p.tiles_y;
  }
  // the same number of z chunks for every item: enough blocks for ~2 per CU
  const long chunks = std::max(1L, (2L * num_cu() + tiles_all - 1) / std::max(1L, tiles_all));
  long total = 0;
  for (int o = 0; o < count; ++o) {
    FusedArgs<R>& p = tab[o];
    const long c = std::max(1L, std::min<long>(chunks, std::max(1, p.d2.nout / (4 * F::N))));
    p.zchunk = std::max(1, (int)((p.d2.nout + c - 1) / c));
    const long nb = (long)p.tiles_x * p.tiles_y * ((p.d2.nout + p.zchunk - 1) / p.zchunk);
    first[o] = (int)total;
    p.nparts = nb;
    total += nb;
  }
  first[count] = (int)total;
  if (total > 0x7fffffffL) return hipErrorInvalidValue;
  s->blocks = (int)total;
  if (kind == BK_SUM && s->any_dot) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->part), sizeof(double) * 3 * std::max(total, 1L));
    if (e != hipSuccess) return e;
    for (int o = 0; o < count; ++o)
      if (it[o].dot) tab[o].part = s->part + 3L * first[o];
    e = hipMalloc(reinterpret_cast<void**>(&s->dots), sizeof(double*) * count);
    if (e != hipSuccess) return e;
    e = hipMemcpy(s->dots, dots.data(), sizeof(double*) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  return batch_upload(tab.data(), sizeof(FusedArgs<R>) * count, first.data(), count, s);
}

template <typename R, class F>
hipError_t fused_stage_t(bool build, int kind, const BatchItem<R>* it, int count, BatchStage* s,
                                hipStream_t st) {
  if (build) {
    if (kind == BK_SUM) {
      switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, HD, WPE) \
  case I: return fused_stage_build<R, F, TX, TY>(kind, it, count, s);
        LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC) \
  case I: return items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) : hipErrorInvalidValue;
        LC_SUM2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, U, WPE, ZC) \
  case I: return items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) : hipErrorInvalidValue;
        LC_SUMP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC, PW) \
  case I: return items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) : hipErrorInvalidValue;
        LC_SUM2W_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PH) \
  case I: return sumt_filter<F>() && items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) : hipErrorInvalidValue;
        LC_SUMT_VARIANTS(X)
#undef X
      }
    } else {
      switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, AD, WPE) \
  case I: return fused_stage_build<R, F, TX, TY>(kind, it, count, s);
        LC_SEP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PD) \
  case I: return items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) : hipErrorInvalidValue;
        LC_SEP2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                                                   \
  case I:                                                                                                  \
    return sepp_filter<F>() && items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) \
                                                    : hipErrorInvalidValue;
        LC_SEPP_VARIANTS(X)
        LC_SEPR_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                                                  \
  case I:                                                                                                 \
    return sepq_ok<R, F, TX, TY, QX, QY>() && items_fit(it, count) ? fused_stage_build<R, F, TX, TY>(kind, it, count, s) \
                                                                 : hipErrorInvalidValue;
        LC_SEPQ_VARIANTS(X)
#undef X
      }
    }
    return hipErrorInvalidValue;
  }
  if (s->blocks == 0) return hipSuccess;
  const FusedArgs<R>* tab = static_cast<const FusedArgs<R>*>(s->tab);
  if (kind == BK_SUM) {
    switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, HD, WPE)                                                                   \
  case I:                                                                                              \
    sum3d_batch<R, F, TX, TY, CPT, QX, BZ, HD, WPE><<<s->blocks, TX * TY / CPT, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SUM_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC)                                                         \
  case I:                                                                                            \
    sum3d2_batch<R, F, TX, TY, QX, QY, QZ, WPE, KC><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SUM2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, U, WPE, ZC)                                                         \
  case I:                                                                                            \
    sump_batch<R, F, TX, TY, QX, QY, PD, U, WPE><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SUMP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, KC, ZC, PW)                                                     \
  case I:                                                                                            \
    sum3d2w_batch<R, F, TX, TY, QX, QY, QZ, WPE, KC, PW><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SUM2W_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PH)                                                        \
  case I:                                                                                            \
    if constexpr (sumt_filter<F>()) {                                                                \
      sumt_batch<R, F, TX, TY, QX, QY, QZ, WPE, PH><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count); \
      break;                                                                                         \
    } else {                                                                                         \
      return hipErrorInvalidValue;                                                                   \
    }
      LC_SUMT_VARIANTS(X)
#undef X
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !s->any_dot) return e;
    return launch_batch_dots(s->part, s->first, s->dots, s->count, st);
  } else {
    switch (s->variant) {
#define X(I, TX, TY, CPT, QX, BZ, AD, WPE)                                                                   \
  case I:                                                                                              \
    sep3d_batch<R, F, TX, TY, CPT, QX, BZ, AD, WPE><<<s->blocks, TX * TY / CPT, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SEP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, QZ, WPE, ZC, PD)                                                                 \
  case I:                                                                                            \
    sep3d2_batch<R, F, TX, TY, QX, QY, QZ, WPE, PD><<<s->blocks, TX * TY, 0, st>>>(tab, s->first, s->count); \
    break;
      LC_SEP2_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                       \
  case I: {                                                                       \
    const hipError_t eb = launch_sepp_batch<R, F, TX, TY, QX, QY, PD, WPE>(tab, s, st); \
    if (eb != hipSuccess) return eb;                                              \
    break;                                                                        \
  }
      LC_SEPP_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                       \
  case I: {                                                                       \
    const hipError_t eb = launch_sepr_batch<R, F, TX, TY, QX, QY, PD, WPE>(tab, s, st); \
    if (eb != hipSuccess) return eb;                                              \
    break;                                                                        \
  }
      LC_SEPR_VARIANTS(X)
#undef X
#define X(I, TX, TY, QX, QY, PD, WPE, ZC)                                       \
  case I: {                                                                       \
    const hipError_t eb = launch_sepq_batch<R, F, TX, TY, QX, QY, PD, WPE>(tab, s, st); \
    if (eb != hipSuccess) return eb;                                              \
    break;                                                                        \
  }
      LC_SEPQ_VARIANTS(X)
#undef X
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

#define LC_FUSED_INSTANTIATE(R, F)                                                                      \
  template hipError_t sum3d_variant<R, F>(int, const Box3&, int, const R*, R*, const double*, R, R,    \
                                          double*, hipStream_t, double*, size_t);                      \
  template hipError_t sep3d_variant<R, F>(int, const Box3&, int, const R*, R*, R, R, hipStream_t, const R*, R*);     \
  template hipError_t fused_stage_t<R, F>(bool, int, const BatchItem<R>*, int, BatchStage*, hipStream_t);

}  // namespace lc

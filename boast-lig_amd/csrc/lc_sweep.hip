// lc_sweep.hip -- one-dimensional libconv sweeps on gfx950.
//
// A sweep applies one operator (convolution filter, wavelet analysis or
// synthesis) along one dimension of an array seen as (L, line, R) with L
// the product of the faster dims and R the product of the slower dims
// (the libconv 1D dispatcher's ndat_left / ndat_right,
// BigDFT/GenericConvolution2.rb:1701-1766).  Two kernels:
//
//   sweep_strided  L > 1: one lane per (l, r) column, consecutive lanes on
//                  consecutive l so every load and store is coalesced; the
//                  lane walks a chunk of the line with a register ring of W
//                  input steps, so each input value is loaded once per chunk.
//   sweep_contig   L == 1 (the processed dim is the contiguous one): a group
//                  of lanes stages a line segment plus its filter halo in LDS
//                  with coalesced loads, then every lane computes Q
//                  consecutive outputs from a register window.
//
// Epilogue per output (ConvolutionFilter#post_process_and_store_values,
// GenericConvolution2.rb:173-203): t *= a; dot += x_i t; t += a_x x_i;
// t += a_y y_i; y_i = t -- the dot and a_x steps only where the output
// carries an input value (grow borders skip them).  dot is accumulated in
// FP64 whatever the storage precision, per block, then summed in a fixed
// order by a one-block kernel (deterministic).
#include "lc_kernels.hpp"

namespace lc {

// ---------------------------------------------------------------- helpers
template <typename R>
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
  }
  return s;
}

template <int K, int BC, typename R>
__device__ __forceinline__ R load_step(const R* __restrict__ xb, long L, int j, int c, int n, int nx,
                                       int ox) {
  if (BC == PER) {
    return xb[in_elem<K>(pmod(j, n), c, nx) * L];
  } else if (BC == GROW) {
    const bool ok = (j >= 0) & (j < n);
    const R v = xb[in_elem<K>(ok ? j : 0, c, nx) * L];
    return ok ? v : R(0);
  } else {
    return xb[in_elem<K>(j - ox, c, nx) * L];
  }
}

// ---------------------------------------------------------------- strided
template <typename R, class Op, int BC>
__global__ __launch_bounds__(256) void sweep_strided(SweepArgs<R> p) {
  constexpr int K = KindOf<Op>::v;
  constexpr int W = Op::W, IN = Op::IN, OUT = Op::OUT, LO = Op::LO, UP = Op::UP;
  __shared__ double red[4];
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = col < p.ncols;
  double dot = 0.0;
  if (active) {
    const long l = col % p.L;
    const long rest = col / p.L;
    const long r = rest % p.nr;
    const int arr = (int)(rest / p.nr);
    const R* __restrict__ xb = p.x + arr * p.nti + r * p.line_in + l;
    R* __restrict__ yb = p.y + arr * p.nto + r * p.line_out + l;
    const bool want_dot = Op::DOT && (p.flags & F_DOT) && arr == p.narr - 1;
    const int i0 = p.i_first + (int)blockIdx.y * p.chunk;
    const int i1 = min(i0 + p.chunk, p.i_first + p.nout);
    R win[W][IN];
    // ring slot of step j is (j - (i0 + LO)) mod W; prefill W-1 steps
    int jp = BC == PER ? pmod(i0 + LO, p.n) : 0;
#pragma unroll
    for (int s = 0; s < W - 1; ++s) {
      const int j = i0 + LO + s;
#pragma unroll
      for (int c = 0; c < IN; ++c) {
        if (BC == PER)
          win[s][c] = xb[in_elem<K>(jp, c, p.nx) * p.L];
        else
          win[s][c] = load_step<K, BC>(xb, p.L, j, c, p.n, p.nx, p.ox);
      }
      if (BC == PER) { ++jp; if (jp == p.n) jp = 0; }
    }
    for (int base = i0; base < i1; base += W) {
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const int i = base + u;
        if (i < i1) {
          const int j = i + UP;
          const int sl = (u + W - 1) % W;
#pragma unroll
          for (int c = 0; c < IN; ++c) {
            if (BC == PER)
              win[sl][c] = xb[in_elem<K>(jp, c, p.nx) * p.L];
            else
              win[sl][c] = load_step<K, BC>(xb, p.L, j, c, p.n, p.nx, p.ox);
          }
          if (BC == PER) { ++jp; if (jp == p.n) jp = 0; }
          R w2[W][IN];
#pragma unroll
          for (int s = 0; s < W; ++s)
#pragma unroll
            for (int c = 0; c < IN; ++c) w2[s][c] = win[(u + s) % W][c];
          R o[OUT];
          Op::compute(w2, o);
          if (p.flags & F_A) {
#pragma unroll
            for (int c = 0; c < OUT; ++c) o[c] = o[c] * p.a;
          }
          if (Op::DOT) {
            const bool has_x = BC != GROW || (i >= 0 && i < p.n);
            if (has_x) {
              const R xc = w2[-LO][0];
              if (want_dot) dot = fmar((double)xc, (double)o[0], dot);
              if (p.flags & F_AX) o[0] = fmar(xc, p.a_x, o[0]);
            }
          }
#pragma unroll
          for (int c = 0; c < OUT; ++c) {
            R* yp = yb + out_elem<K>(i - p.oy, c, p.ny) * p.L;
            if (p.flags & F_AY) o[c] = fmar(p.a_y, *yp, o[c]);
            *yp = o[c];
          }
        }
      }
    }
  }
  if (p.dot_part) {
    const double s = block_sum<R>(dot, red);
    if (threadIdx.x == 0) p.dot_part[blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------- contiguous
// Q outputs per lane, TPL = 2^tpl_log2 lanes per line segment, 256/TPL lines
// per block; LDS holds, per line, the NS = TPL*Q + W - 1 input steps of the
// segment, one array per input component.
template <typename R, class Op, int BC, int Q>
__global__ __launch_bounds__(256) void sweep_contig(SweepArgs<R> p, int tpl_log2) {
  constexpr int K = KindOf<Op>::v;
  constexpr int W = Op::W, IN = Op::IN, OUT = Op::OUT, LO = Op::LO;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R* sm = reinterpret_cast<R*>(smem_raw);
  __shared__ double red[4];
  const int TPL = 1 << tpl_log2;
  const int LPB = 256 >> tpl_log2;
  const int seg = TPL * Q;
  const int NS = seg + W - 1;
  const int tl = threadIdx.x & (TPL - 1);
  const int lnb = threadIdx.x >> tpl_log2;
  const long line = (long)blockIdx.x * LPB + lnb;
  const bool active = line < p.ncols;
  const int i0 = p.i_first + (int)blockIdx.y * seg;
  const int iend = p.i_first + p.nout;
  const long r = active ? line % p.nr : 0;
  const int arr = active ? (int)(line / p.nr) : 0;
  const R* __restrict__ xb = p.x + arr * p.nti + r * p.line_in;
  R* __restrict__ yb = p.y + arr * p.nto + r * p.line_out;
  if (active) {
    // the last segment may extend past the last output: its steps beyond
    // iend - 1 + UP are never used and may lie outside x (shrink), so they
    // are not loaded
    const int jmax = iend - 1 + Op::UP;
    for (int e = tl; e < NS * IN; e += TPL) {
      int s, c;
      if (K == KDWT) { s = e >> 1; c = e & 1; }      // interleaved sample pairs
      else if (K == KIDWT) { c = e / NS; s = e - c * NS; }  // L block then H block
      else { s = e; c = 0; }
      const int j = i0 + LO + s;
      sm[(c * LPB + lnb) * NS + s] = j <= jmax ? load_step<K, BC>(xb, 1, j, c, p.n, p.nx, p.ox) : R(0);
    }
  }
  __syncthreads();
  double dot = 0.0;
  if (active) {
    const bool want_dot = Op::DOT && (p.flags & F_DOT) && arr == p.narr - 1;
    const int ib = i0 + tl * Q;
    R win[Q + W - 1][IN];
#pragma unroll
    for (int s = 0; s < Q + W - 1; ++s)
#pragma unroll
      for (int c = 0; c < IN; ++c) win[s][c] = sm[(c * LPB + lnb) * NS + tl * Q + s];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = ib + q;
      if (i < iend) {
        R w2[W][IN];
#pragma unroll
        for (int s = 0; s < W; ++s)
#pragma unroll
          for (int c = 0; c < IN; ++c) w2[s][c] = win[q + s][c];
        R o[OUT];
        Op::compute(w2, o);
        if (p.flags & F_A) {
#pragma unroll
          for (int c = 0; c < OUT; ++c) o[c] = o[c] * p.a;
        }
        if (Op::DOT) {
          const bool has_x = BC != GROW || (i >= 0 && i < p.n);
          if (has_x) {
            const R xc = w2[-LO][0];
            if (want_dot) dot = fmar((double)xc, (double)o[0], dot);
            if (p.flags & F_AX) o[0] = fmar(xc, p.a_x, o[0]);
          }
        }
#pragma unroll
        for (int c = 0; c < OUT; ++c) {
          R* yp = yb + out_elem<K>(i - p.oy, c, p.ny);
          if (p.flags & F_AY) o[c] = fmar(p.a_y, *yp, o[c]);
          *yp = o[c];
        }
      }
    }
  }
  if (p.dot_part) {
    const double s = block_sum<R>(dot, red);
    if (threadIdx.x == 0) p.dot_part[blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// Sums of nsets (<= 4) sets of nparts partials, in a fixed order (one block):
// out[k] = sum of set k.  Every thread keeps one accumulator per set and
// loads four strided partials of each set before adding them (in index
// order), so the loads of a thread are in flight together.
constexpr int kPartThreads = 1024;
__global__ __launch_bounds__(kPartThreads) void sum_partials(const double* __restrict__ part, long nparts,
                                                              double* __restrict__ out, int nsets, long stride) {
  __shared__ double red[kPartThreads / 64][4];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  constexpr int U = 4;
  long i = threadIdx.x;
  for (; i + (U - 1) * kPartThreads < nparts; i += U * kPartThreads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < nsets) {
        double a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = part[k * stride + i + u * kPartThreads];
#pragma unroll
        for (int u = 0; u < U; ++u) v[k] += a[u];
      }
    }
  }
  for (; i < nparts; i += kPartThreads)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nsets) v[k] += part[k * stride + i];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    if (lane == 0) red[wid][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < nsets) {
    double s = 0.0;
    for (int w = 0; w < kPartThreads / 64; ++w) s += red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------- launcher
template <typename R, class Op, int BC>
static hipError_t run_sweep(const SweepArgs<R>& a0, hipStream_t st, double* dot_out) {
  SweepArgs<R> p = a0;
  constexpr int Q = 4;
  if (p.ncols == 0 || p.nout == 0) {
    if (dot_out) return hipMemsetAsync(dot_out, 0, sizeof(double), st);
    return hipSuccess;
  }
  dim3 grid;
  hipError_t e;
  if (p.L == 1) {
    int lg = 2;
    while (lg < 6 && (1 << lg) * Q < p.nout) ++lg;
    const int TPL = 1 << lg, LPB = 256 >> lg, seg = TPL * Q;
    const int NS = seg + Op::W - 1;
    grid = dim3((unsigned)((p.ncols + LPB - 1) / LPB), (unsigned)((p.nout + seg - 1) / seg));
    const size_t sh = (size_t)LPB * NS * Op::IN * sizeof(R);
    if (!p.dot_part) {
      sweep_contig<R, Op, BC, Q><<<grid, 256, sh, st>>>(p, lg);
    } else {
      sweep_contig<R, Op, BC, Q><<<grid, 256, sh, st>>>(p, lg);
    }
  } else {
    // chunk so that the grid holds ~8 waves per CU while chunks stay >= 2W
    const long waves_cols = (p.ncols + 63) / 64;
    int nchunk = (int)((8L * 256 + waves_cols - 1) / waves_cols);
    int maxchunk = p.nout / (2 * Op::W);
    if (nchunk > maxchunk) nchunk = maxchunk;
    if (nchunk < 1) nchunk = 1;
    p.chunk = (p.nout + nchunk - 1) / nchunk;
    grid = dim3((unsigned)((p.ncols + 255) / 256), (unsigned)((p.nout + p.chunk - 1) / p.chunk));
    sweep_strided<R, Op, BC><<<grid, 256, 0, st>>>(p);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (p.dot_part && dot_out) {
    sum_partials<<<1, kPartThreads, 0, st>>>(p.dot_part, (long)grid.x * grid.y, dot_out, 1, 0);
    e = hipGetLastError();
  }
  return e;
}

template <typename R>
long sweep_partials_needed(const SweepArgs<R>& p, int W) {
  if (p.L == 1) {
    int lg = 2;
    while (lg < 6 && (1 << lg) * 4 < p.nout) ++lg;
    const int LPB = 256 >> lg, seg = (1 << lg) * 4;
    return ((p.ncols + LPB - 1) / LPB) * (long)((p.nout + seg - 1) / seg);
  }
  return ((p.ncols + 255) / 256) * (long)p.nout;  // upper bound (chunk >= 1)
}

template <typename R, class Op>
static hipError_t run_bc(int bc, const SweepArgs<R>& p, hipStream_t st, double* dot_out) {
  switch (bc) {
    case PER: return run_sweep<R, Op, PER>(p, st, dot_out);
    case GROW: return run_sweep<R, Op, GROW>(p, st, dot_out);
    case SHRINK: return run_sweep<R, Op, SHRINK>(p, st, dot_out);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t sweep(int kind, int id, int bc, const SweepArgs<R>& p, hipStream_t st, double* dot_out) {
  if (kind == KCONV) {
    switch (id) {
      case OP_MD: return run_bc<R, Conv<FMD>>(bc, p, st, dot_out);
      case OP_IMD: return run_bc<R, Conv<FIMD>>(bc, p, st, dot_out);
      case OP_D2: return run_bc<R, Conv<FD2>>(bc, p, st, dot_out);
      case OP_D1: return run_bc<R, Conv<FD1>>(bc, p, st, dot_out);
    }
  } else if (kind == KDWT) {
    if (id == WAV_SYM8) return run_bc<R, Dwt<Sym8Dec>>(bc, p, st, nullptr);
    if (id == WAV_SYMICOMB8)
      return bc == PER ? run_bc<R, Dwt<CombPDec>>(bc, p, st, nullptr)
                       : run_bc<R, Dwt<CombGDec>>(bc, p, st, nullptr);
  } else if (kind == KIDWT) {
    if (id == WAV_SYM8) return run_bc<R, Idwt<Sym8Rec>>(bc, p, st, nullptr);
    if (id == WAV_SYMICOMB8)
      return bc == PER ? run_bc<R, Idwt<CombPRec>>(bc, p, st, nullptr)
                       : run_bc<R, Idwt<CombGRec>>(bc, p, st, nullptr);
  }
  return hipErrorInvalidValue;
}

template hipError_t sweep<double>(int, int, int, const SweepArgs<double>&, hipStream_t, double*);
template hipError_t sweep<float>(int, int, int, const SweepArgs<float>&, hipStream_t, double*);
template long sweep_partials_needed<double>(const SweepArgs<double>&, int);
template long sweep_partials_needed<float>(const SweepArgs<float>&, int);

hipError_t sum_partial_sets(const double* part, long nparts, double* out, int nsets, long stride,
                            hipStream_t st) {
  if (nsets < 1 || nsets > 4) return hipErrorInvalidValue;
  sum_partials<<<1, kPartThreads, 0, st>>>(part, nparts, out, nsets, stride);
  return hipGetLastError();
}

}  // namespace lc

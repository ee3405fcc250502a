// lc_cpu.cpp -- the CPU baseline timed beside the GPU: a C++/OpenMP
// restatement of the reference's generated CPU sweeps for the 3D
// MagicFilter -> Kinetic chain.  TEST INFRASTRUCTURE ONLY (like lc_oracle.c):
// bench.py's cpu_baseline leg and tests/ load it, the product never does.
//
// Structure of the generated code it keeps (BigDFT/GenericConvolution2.rb):
//   * one full-array pass per 1D sweep (GenericConvolutionOperator#procedure,
//     :2160-2240): MF3D is three sweeps through two work arrays, the Kinetic
//     sum form three passes accumulating into y, as kinetic_per_ref_optim
//     (BigDFT/Kinetic.rb:395-958: conv_kin_x writes, conv_kin_y/z add);
//   * per line, three regions (ConvolutionOperator1d#conv_lines :1363-1381):
//     the centre, where every tap reads x[i + l] directly, and the begin /
//     end borders, where a mod_arr-style table wraps (periodic) or drops
//     (free) the taps (Filter#get_loop_start_end :91-121);
//   * unrolled lines (#for_conv :1425-1440, the U accumulators tt): the
//     strided sweeps filter a block of consecutive lines of the contiguous
//     dim at once, the contiguous sweep vectorises along its line, each with
//     the taps fully unrolled (compile-time filter tables);
//   * OpenMP over the outer non-processed dims, energies as reductions.
// Work arrays are the caller's: nothing is allocated in a timed call.
// Arithmetic: FP64 FMA per tap in increasing tap order; results agree with
// the oracle to rounding (the tests check 1e-13).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../boast-lig_amd/csrc/lc_taps.h"

namespace {

#define AS_D(v) v,
constexpr double kMD[] = {LC_TAPS_MD(AS_D)};
constexpr double kD2[] = {LC_TAPS_D2(AS_D)};
constexpr int kMDLo = LC_MD_LO, kMDN = LC_MD_N;
constexpr int kD2Lo = LC_D2_LO, kD2N = LC_D2_N;

enum { PER = 0, GROW = 1, SHRINK = -1 };

// One processed dim: outputs i in [i0, i0 + nout), logical input j = i + l,
// storage index of j (or -1: zero) from the BC.
struct Axis {
  int n, bc, nx, ny, lo, up, i0, nout, ox;
  // storage index of logical input j, -1 where x~ is zero
  int src(int j) const {
    if (bc == PER) { int r = j % n; return r < 0 ? r + n : r; }
    if (bc == GROW) return (j >= 0 && j < n) ? j : -1;
    return j - ox;  // shrink: always stored
  }
  // first / last output whose taps all read consecutive stored inputs
  int c0() const { return bc == SHRINK ? i0 : std::max(i0, -lo); }
  int c1() const { return bc == SHRINK ? i0 + nout : std::min(i0 + nout, n - up); }
};

Axis axis(int n, int bc, int nx, int ny, int lo, int N) {
  Axis a;
  a.n = n; a.bc = bc; a.nx = nx; a.ny = ny; a.lo = lo; a.up = lo + N - 1;
  a.i0 = bc == GROW ? -a.up : 0;
  a.nout = bc == GROW ? n + N - 1 : n;
  a.ox = bc == SHRINK ? lo : 0;
  return a;
}

// Non-processed dim: output index o in [0, nout) of the box read at storage
// row src (or zero when the row is outside the input under grow).
struct Side {
  int nout, first_out;  // output rows, logical index of row 0
  int bc, n, ox;
  int row(int o) const {
    const int j = first_out + o;
    if (bc == GROW) return (j >= 0 && j < n) ? j : -1;
    return j - ox;
  }
};

Side side(int n, int bc, int lo, int N) {
  Side s;
  s.bc = bc; s.n = n;
  s.nout = bc == GROW ? n + N - 1 : n;
  s.first_out = bc == GROW ? -(lo + N - 1) : 0;
  s.ox = bc == SHRINK ? lo : 0;
  return s;
}

constexpr int BI = 16;  // unrolled lines of the strided sweeps (contiguous dim block)

// ---- sweep along the contiguous dim 0: y(i, r) = sum_l f_l x~(i + l, r)
template <int N>
void line_x(const double* f, const Axis& a, const double* __restrict__ xr, double* __restrict__ yr,
            double scale, bool add) {
  const int c0 = a.c0(), c1 = a.c1();
  auto border = [&](int i) {
    double t = 0.0;
    for (int l = 0; l < N; ++l) {
      const int s = a.src(i + a.lo + l);
      if (s >= 0) t = fma(xr[s], f[l], t);
    }
    return t;
  };
  for (int i = a.i0; i < c0; ++i) {
    const double t = border(i) * scale;
    yr[i - a.i0] = add ? yr[i - a.i0] + t : t;
  }
  const double* xb = xr + (a.bc == SHRINK ? -a.ox : 0) + a.lo;
#pragma omp simd
  for (int i = c0; i < c1; ++i) {
    double t = 0.0;
#pragma GCC unroll 32
    for (int l = 0; l < N; ++l) t = fma(xb[i + l], f[l], t);
    t *= scale;
    yr[i - a.i0] = add ? yr[i - a.i0] + t : t;
  }
  for (int i = std::max(c1, c0); i < a.i0 + a.nout; ++i) {
    const double t = border(i) * scale;
    yr[i - a.i0] = add ? yr[i - a.i0] + t : t;
  }
}

// Per-thread scratch that survives the calls: grown on first use, so a timed
// call (after a warm-up) allocates nothing.
double* thread_buf(std::vector<double>& v, size_t n) {
  if (v.size() < n) v.resize(n);
  return v.data();
}

// ---- sweep along a strided dim: y(q, i) = sum_l f_l x~(q, i + l) for a
// block of BI consecutive q of the contiguous dim (the unrolled lines).  The
// block's input column is first copied to a contiguous per-thread cache
// (stored rows x BI values), as kinetic_per_ref_optim's zcache does
// (BigDFT/Kinetic.rb:524-700): the strided rows of a power-of-two grid would
// otherwise fall on the same cache sets.
template <int N>
void lines_strided(const double* f, const Axis& a, const double* __restrict__ x, long xs,
                   double* __restrict__ y, long ys, int nq, double scale, bool add) {
  static thread_local std::vector<double> cache;
  const int c0 = a.c0(), c1 = a.c1();
  const int rows = a.bc == SHRINK ? a.nx : a.n;  // stored rows the taps read
  double* __restrict__ zc = thread_buf(cache, (size_t)rows * BI);
  for (int q0 = 0; q0 < nq; q0 += BI) {
    const int qn = std::min(BI, nq - q0);
    for (int r = 0; r < rows; ++r) {
      const double* xr = x + (long)r * xs + q0;
      double* zr = zc + (long)r * BI;
      if (qn == BI) {
#pragma omp simd
        for (int q = 0; q < BI; ++q) zr[q] = xr[q];
      } else {
        for (int q = 0; q < BI; ++q) zr[q] = q < qn ? xr[q] : 0.0;
      }
    }
    for (int i = a.i0; i < a.i0 + a.nout; ++i) {
      double acc[BI];
      for (int q = 0; q < BI; ++q) acc[q] = 0.0;
      if (i >= c0 && i < c1) {
        const double* xb = zc + (long)((a.bc == SHRINK ? -a.ox : 0) + a.lo + i) * BI;
#pragma GCC unroll 32
        for (int l = 0; l < N; ++l) {
          const double fl = f[l];
          const double* xl = xb + (long)l * BI;
#pragma omp simd
          for (int q = 0; q < BI; ++q) acc[q] = fma(xl[q], fl, acc[q]);
        }
      } else {
        for (int l = 0; l < N; ++l) {
          const int sidx = a.src(i + a.lo + l);
          if (sidx < 0) continue;
          const double* xl = zc + (long)sidx * BI;
#pragma omp simd
          for (int q = 0; q < BI; ++q) acc[q] = fma(xl[q], f[l], acc[q]);
        }
      }
      double* yl = y + (long)(i - a.i0) * ys + q0;
      if (add)
        for (int q = 0; q < qn; ++q) yl[q] += acc[q] * scale;
      else
        for (int q = 0; q < qn; ++q) yl[q] = acc[q] * scale;
    }
  }
}

}  // namespace

extern "C" {

int lcc_num_threads(void);

// 3D MagicFilter (separable product, filter sym8_md) as three sweeps:
// x (nx) -> w1 -> w2 -> y (ny).  w1 holds ny0 nx1 nx2, w2 ny0 ny1 nx2 values.
int lcc_magicfilter3d(const int* n, const int* bc, const int* nx, const int* ny, const double* x,
                      double* y, double* w1, double* w2) {
  Axis a[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = axis(n[d], bc[d], nx[d], ny[d], kMDLo, kMDN);
    if (nx[d] < (bc[d] == SHRINK ? n[d] + kMDN - 1 : n[d])) return -1;
    if (ny[d] < a[d].nout) return -1;
  }
  // sweep 0 along x: rows (j1, j2) of x, all of them (the other dims are
  // still input), into w1 (ny0, nx1, nx2)
#pragma omp parallel for schedule(static) collapse(2)
  for (int j2 = 0; j2 < nx[2]; ++j2)
    for (int j1 = 0; j1 < nx[1]; ++j1)
      line_x<kMDN>(kMD, a[0], x + ((long)j2 * nx[1] + j1) * nx[0], w1 + ((long)j2 * nx[1] + j1) * ny[0],
                   1.0, false);
  // sweep 1 along y: w1 (ny0, nx1, nx2) -> w2 (ny0, ny1, nx2); lines of dim 0 unrolled
#pragma omp parallel for schedule(static)
  for (int j2 = 0; j2 < nx[2]; ++j2)
    lines_strided<kMDN>(kMD, a[1], w1 + (long)j2 * nx[1] * ny[0], ny[0], w2 + (long)j2 * ny[1] * ny[0], ny[0],
                        a[0].nout, 1.0, false);
  // sweep 2 along z: w2 (ny0, ny1, nx2) -> y; blocks of the contiguous
  // (dim 0, dim 1) plane, one OpenMP chunk per dim-1 row
  const long sp = (long)ny[0] * ny[1];
#pragma omp parallel for schedule(static)
  for (int j1 = 0; j1 < a[1].nout; ++j1)
    lines_strided<kMDN>(kMD, a[2], w2 + (long)j1 * ny[0], sp, y + (long)j1 * ny[0], sp, a[0].nout, 1.0, false);
  return 0;
}

// 3D Kinetic sum form y = sum_d a_d D2_d x~ (filter sym8_d2) with the three
// energies dot[d] = <x~, a_d D2_d x~> over the output box, as three passes
// (x writes, y and z accumulate) in the manner of kinetic_per_ref_optim.
int lcc_kinetic3d(const int* n, const int* bc, const int* nx, const int* ny, const double* a,
                  const double* x, double* y, double* dot) {
  Axis ax[3];
  Side sd[3];
  for (int d = 0; d < 3; ++d) {
    ax[d] = axis(n[d], bc[d], nx[d], ny[d], kD2Lo, kD2N);
    sd[d] = side(n[d], bc[d], kD2Lo, kD2N);
    if (nx[d] < (bc[d] == SHRINK ? n[d] + kD2N - 1 : n[d])) return -1;
    if (ny[d] < ax[d].nout) return -1;
  }
  const long syo = (long)ny[0] * ny[1], sxi = (long)nx[0] * nx[1];
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  // centre value of output point (o0, o1, o2), zero where there is no input
  auto centre = [&](int o0, int o1, int o2) -> double {
    const int r0 = sd[0].row(o0), r1 = sd[1].row(o1), r2 = sd[2].row(o2);
    if (r0 < 0 || r1 < 0 || r2 < 0) return 0.0;
    return x[(long)r2 * sxi + (long)r1 * nx[0] + r0];
  };
  // pass x: rows (o1, o2) of the output box
#pragma omp parallel for schedule(static) collapse(2) reduction(+ : e0)
  for (int o2 = 0; o2 < sd[2].nout; ++o2)
    for (int o1 = 0; o1 < sd[1].nout; ++o1) {
      double* yr = y + (long)o2 * syo + (long)o1 * ny[0];
      const int r1 = sd[1].row(o1), r2 = sd[2].row(o2);
      if (r1 < 0 || r2 < 0) {
        for (int o0 = 0; o0 < ax[0].nout; ++o0) yr[o0] = 0.0;
        continue;
      }
      line_x<kD2N>(kD2, ax[0], x + (long)r2 * sxi + (long)r1 * nx[0], yr, a[0], false);
      if (dot)
        for (int o0 = 0; o0 < ax[0].nout; ++o0) e0 = fma(centre(o0, o1, o2), yr[o0], e0);
    }
  // pass y: planes o2, lines of dim 0 unrolled; the energy needs the pass's
  // own term, so each plane filters into a row buffer first
#pragma omp parallel reduction(+ : e1)
  {
    static thread_local std::vector<double> term_y;
    double* term = thread_buf(term_y, (size_t)ax[1].nout * ax[0].nout);
#pragma omp for schedule(static)
    for (int o2 = 0; o2 < sd[2].nout; ++o2) {
      const int r2 = sd[2].row(o2);
      if (r2 < 0) continue;
      // x rows of this plane restricted to the output columns of dim 0
      const double* xp = x + (long)r2 * sxi;
      // dim 0 of the output box maps to stored columns o0 - ox (shrink) or o0
      // (periodic); grow columns outside the input contribute zero
      const int c_off = sd[0].first_out - sd[0].ox;  // stored column of o0 = 0
      if (sd[0].bc == GROW) {
        // general path: gather the columns with input, filter, scatter
        for (int o1 = 0; o1 < ax[1].nout; ++o1)
          for (int o0 = 0; o0 < ax[0].nout; ++o0) {
            const int r0 = sd[0].row(o0);
            double t = 0.0;
            if (r0 >= 0)
              for (int l = 0; l < kD2N; ++l) {
                const int s = ax[1].src(ax[1].i0 + o1 + ax[1].lo + l);
                if (s >= 0) t = fma(xp[(long)s * nx[0] + r0], kD2[l], t);
              }
            term[(size_t)o1 * ax[0].nout + o0] = t * a[1];
          }
      } else {
        lines_strided<kD2N>(kD2, ax[1], xp + c_off, nx[0], term, ax[0].nout, ax[0].nout, a[1], false);
      }
      double* yp = y + (long)o2 * syo;
      for (int o1 = 0; o1 < ax[1].nout; ++o1) {
        const double* tr = term + (size_t)o1 * ax[0].nout;
        double* yr = yp + (long)o1 * ny[0];
        for (int o0 = 0; o0 < ax[0].nout; ++o0) {
          yr[o0] += tr[o0];
          if (dot) e1 = fma(centre(o0, o1, o2), tr[o0], e1);
        }
      }
    }
  }
  // pass z: rows o1, lines of dim 0 unrolled, into a row buffer per (o1)
#pragma omp parallel reduction(+ : e2)
  {
    static thread_local std::vector<double> term_z;
    double* term = thread_buf(term_z, (size_t)ax[2].nout * ax[0].nout);
#pragma omp for schedule(static)
    for (int o1 = 0; o1 < sd[1].nout; ++o1) {
      const int r1 = sd[1].row(o1);
      if (r1 < 0) continue;
      const double* xr = x + (long)r1 * nx[0];
      const int c_off = sd[0].first_out - sd[0].ox;
      if (sd[0].bc == GROW) {
        for (int o2 = 0; o2 < ax[2].nout; ++o2)
          for (int o0 = 0; o0 < ax[0].nout; ++o0) {
            const int r0 = sd[0].row(o0);
            double t = 0.0;
            if (r0 >= 0)
              for (int l = 0; l < kD2N; ++l) {
                const int s = ax[2].src(ax[2].i0 + o2 + ax[2].lo + l);
                if (s >= 0) t = fma(xr[(long)s * sxi + r0], kD2[l], t);
              }
            term[(size_t)o2 * ax[0].nout + o0] = t * a[2];
          }
      } else {
        lines_strided<kD2N>(kD2, ax[2], xr + c_off, sxi, term, ax[0].nout, ax[0].nout, a[2], false);
      }
      for (int o2 = 0; o2 < ax[2].nout; ++o2) {
        const double* tr = term + (size_t)o2 * ax[0].nout;
        double* yr = y + (long)o2 * syo + (long)o1 * ny[0];
        for (int o0 = 0; o0 < ax[0].nout; ++o0) {
          yr[o0] += tr[o0];
          if (dot) e2 = fma(centre(o0, o1, o2), tr[o0], e2);
        }
      }
    }
  }
  if (dot) { dot[0] = e0; dot[1] = e1; dot[2] = e2; }
  return 0;
}

}  // extern "C"

#ifdef _OPENMP
#include <omp.h>
extern "C" int lcc_num_threads(void) { return omp_get_max_threads(); }
#else
extern "C" int lcc_num_threads(void) { return 1; }
#endif

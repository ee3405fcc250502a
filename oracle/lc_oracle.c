/*
 * lc_oracle.c -- FP64 CPU oracle of the libconv separable convolutions.
 * TEST INFRASTRUCTURE ONLY (see lc_oracle.h): the checker for the HIP
 * library and the CPU baseline timed beside it, never the product.
 *
 * Every routine restates a piece of the BOAST generator; the file:line
 * citations are relative to the reference checkout.
 */
#include "lc_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../boast-lig_amd/csrc/lc_taps.h"

#define AS_D(v) v,
static const double TAP_MD[] = {LC_TAPS_MD(AS_D)};
static const double TAP_IMD[] = {LC_TAPS_IMD(AS_D)};
static const double TAP_D2[] = {LC_TAPS_D2(AS_D)};
static const double TAP_D1[] = {LC_TAPS_D1(AS_D)};
static const double TAP_LP[] = {LC_TAPS_LP(AS_D)};
static const double TAP_HP[] = {LC_TAPS_HP(AS_D)};
static const double TAP_CLPP[] = {LC_TAPS_CLPP(AS_D)};
static const double TAP_CHPP[] = {LC_TAPS_CHPP(AS_D)};
static const double TAP_CLPG[] = {LC_TAPS_CLPG(AS_D)};
static const double TAP_CHPG[] = {LC_TAPS_CHPG(AS_D)};

/* BC_FREE (2) is BoundaryConditions::FREE ("fd", discard) of
 * GenericConvolution2.rb:684-740: input [0, n) zero outside, output [0, n)
 * (the grow result restricted to the input box).  The generator lists it but
 * does not emit it (CONDITIONS = [PERIODIC, GROW, SHRINK]); libconv.f90:245
 * asks for it ('F', "free, nogrow") for the Kinetic step of the local
 * Hamiltonian.  The oracle takes it for convolutions only. */
enum { BC_PER = 0, BC_GROW = 1, BC_SHRINK = -1, BC_FREE = 2 };

int lco_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ConvolutionFilter: taps f[l], l in [lo, up] (GenericConvolution2.rb:124-171) */
typedef struct {
    int lo, up;
    const double *f; /* f[0] is tap lo */
} cfilter;

static int get_cfilter(int fil, cfilter *cf) {
    switch (fil) {
    case LCO_MD: cf->lo = LC_MD_LO; cf->up = LC_MD_LO + LC_MD_N - 1; cf->f = TAP_MD; return 0;
    case LCO_IMD: cf->lo = LC_IMD_LO; cf->up = LC_IMD_LO + LC_IMD_N - 1; cf->f = TAP_IMD; return 0;
    case LCO_D2: cf->lo = LC_D2_LO; cf->up = LC_D2_LO + LC_D2_N - 1; cf->f = TAP_D2; return 0;
    case LCO_D1: cf->lo = LC_D1_LO; cf->up = LC_D1_LO + LC_D1_N - 1; cf->f = TAP_D1; return 0;
    default: return -1;
    }
}

int lco_filter_taps(int fil, int *lo, double *out, int cap) {
    cfilter cf;
    if (get_cfilter(fil, &cf)) return -1;
    int nt = cf.up - cf.lo + 1;
    *lo = cf.lo;
    for (int i = 0; i < nt && i < cap; ++i) out[i] = cf.f[i];
    return nt;
}

/* true modulo (Fortran modulo); the generated C helper (a+b)%b is wrong for
 * a < -b (GenericConvolution2.rb:2314-2330), which the oracle avoids. */
static inline long pmod(long a, long b) {
    long r = a % b;
    return r < 0 ? r + b : r;
}

/* Input origin (first logical index stored) and output origin along the
 * processed dim: shrink input starts at lo, grow output at -up
 * (ConvolutionFilter#get_input_dim_ld_shrink / get_output_dim_ld_grow,
 * GenericConvolution2.rb:219-225). */
static inline int in_origin(int bc, int lo) { return bc == BC_SHRINK ? lo : 0; }
static inline int out_origin(int bc, int up) { return bc == BC_GROW ? -up : 0; }

/* One line of a convolution sweep (ConvolutionOperator1d#conv_lines and
 * #for_conv, GenericConvolution2.rb:1363-1440, epilogue of
 * ConvolutionFilter#post_process_and_store_values :173-203):
 *   t = sum_l f[l] x[i+l]; t *= a; dot += x[i] t; t += a_x x[i];
 *   t += a_y y[i] (or y[i] when accumulating); y[i] = t.
 * Grow outputs in [-up, n-1-lo]; the dot and a_x steps are skipped where
 * no input exists (i outside [0, n)). */
static double conv_line(const cfilter *cf, int bc, int n,
                        const double *x, long sx, int ox,
                        double *y, long sy, int oy,
                        int use_a, double a, int use_ax, double a_x,
                        int use_ay, double a_y, int want_dot) {
    const int lo = cf->lo, up = cf->up;
    int i0 = 0, i1 = n - 1;
    if (bc == BC_GROW) { i0 = -up; i1 = n - 1 - lo; }
    double dot = 0.0;
    for (int i = i0; i <= i1; ++i) {
        double t = 0.0;
        for (int l = lo; l <= up; ++l) {
            long j = (long)i + l;
            double v;
            if (bc == BC_PER) {
                v = x[pmod(j, n) * sx];
            } else if (bc == BC_GROW || bc == BC_FREE) {
                if (j < 0 || j >= n) continue; /* both sides clamped */
                v = x[j * sx];
            } else {
                v = x[(j - ox) * sx];
            }
            t = v * cf->f[l - lo] + t;
        }
        if (use_a) t = t * a;
        int has_x = !(bc == BC_GROW && (i < 0 || i >= n));
        if (has_x) {
            double xc = x[((long)i - ox) * sx];
            if (want_dot) dot = xc * t + dot;
            if (use_ax) t = xc * a_x + t;
        }
        double *yp = &y[((long)i - oy) * sy];
        if (use_ay) t = a_y * (*yp) + t;
        *yp = t;
    }
    return dot;
}

/* Wavelet filter split into even/odd halves (WaveletFilter#split_filters,
 * GenericConvolution2.rb:386-410; centres :303-330, :411-430). */
typedef struct {
    int lo, up;               /* pair offsets of the split filters */
    double le[32], lo_[32], he[32], ho[32];
} wfilter;

static int get_wfilter(int wav, int recompose, int bc, wfilter *wf) {
    const double *low, *high;
    int len, center;
    if (wav == LCO_SYM8) {
        low = TAP_LP; high = TAP_HP; len = LC_LP_N; center = 8;
    } else if (wav == LCO_SYMICOMB8) {
        if (bc == BC_PER) { low = TAP_CLPP; high = TAP_CHPP; len = LC_CLPP_N; center = -LC_CLPP_LO; }
        else { low = TAP_CLPG; high = TAP_CHPG; len = LC_CLPG_N; center = -LC_CLPG_LO; }
    } else {
        return -1;
    }
    int cntr = recompose ? (len - center - 1) : center;
    int half = cntr / 2;
    int m = len / 2;
    wf->lo = -half;
    wf->up = m - half - 1;
    for (int k = 0; k < m; ++k) {
        if (recompose) { /* reversed filters */
            wf->le[k] = low[len - 1 - 2 * k];
            wf->lo_[k] = low[len - 2 - 2 * k];
            wf->he[k] = high[len - 1 - 2 * k];
            wf->ho[k] = high[len - 2 - 2 * k];
        } else {
            wf->le[k] = low[2 * k];
            wf->lo_[k] = low[2 * k + 1];
            wf->he[k] = high[2 * k];
            wf->ho[k] = high[2 * k + 1];
        }
    }
    return 0;
}

/* fetch pair j of a wavelet input line with BC (zero outside for grow) */
static inline int pair_index(int bc, long j, int n, int ox, long *pos) {
    if (bc == BC_PER) { *pos = pmod(j, n); return 1; }
    if (bc == BC_GROW) { if (j < 0 || j >= n) return 0; *pos = j; return 1; }
    *pos = j - ox;
    return 1;
}

/* One line of wavelet analysis (WaveletFilterDecompose,
 * GenericConvolution2.rb:505-587 with the 4-FMA body :445-462):
 * input pairs (s[2j], s[2j+1]) at offsets c + 2(j-ox); output L_i at i-oy,
 * H_i at ny + i-oy.  y = a*t (+ a_y y). */
static void dwt_line(const wfilter *wf, int bc, int n,
                     const double *x, long sx, int ox, int nxl,
                     double *y, long sy, int oy, int nyl,
                     int use_a, double a, int use_ay, double a_y) {
    (void)nxl;
    int i0 = 0, i1 = n - 1;
    if (bc == BC_GROW) { i0 = -wf->up; i1 = n - 1 - wf->lo; }
    for (int i = i0; i <= i1; ++i) {
        double lt = 0.0, ht = 0.0;
        for (int l = wf->lo; l <= wf->up; ++l) {
            long pos;
            if (!pair_index(bc, (long)i + l, n, ox, &pos)) continue;
            double se = x[(2 * pos) * sx], so = x[(2 * pos + 1) * sx];
            int k = l - wf->lo;
            lt = se * wf->le[k] + lt;
            ht = se * wf->he[k] + ht;
            lt = so * wf->lo_[k] + lt;
            ht = so * wf->ho[k] + ht;
        }
        if (use_a) { lt = lt * a; ht = ht * a; }
        double *yl = &y[((long)i - oy) * sy];
        double *yh = &y[((long)i - oy + nyl) * sy];
        if (use_ay) { lt = (*yl) * a_y + lt; ht = (*yh) * a_y + ht; }
        *yl = lt;
        *yh = ht;
    }
}

/* One line of wavelet synthesis (WaveletFilterRecompose,
 * GenericConvolution2.rb:589-671): input L_j at j-ox, H_j at nx + j-ox;
 * output samples s[2i] at 2(i-oy), s[2i+1] at 2(i-oy)+1. */
static void idwt_line(const wfilter *wf, int bc, int n,
                      const double *x, long sx, int ox, int nxl,
                      double *y, long sy, int oy, int nyl,
                      int use_a, double a, int use_ay, double a_y) {
    (void)nyl;
    int i0 = 0, i1 = n - 1;
    if (bc == BC_GROW) { i0 = -wf->up; i1 = n - 1 - wf->lo; }
    for (int i = i0; i <= i1; ++i) {
        double et = 0.0, ot = 0.0;
        for (int l = wf->lo; l <= wf->up; ++l) {
            long pos;
            if (!pair_index(bc, (long)i + l, n, ox, &pos)) continue;
            double lv = x[pos * sx], hv = x[(pos + nxl) * sx];
            int k = l - wf->lo;
            et = lv * wf->lo_[k] + et;
            ot = lv * wf->le[k] + ot;
            et = hv * wf->ho[k] + et;
            ot = hv * wf->he[k] + ot;
        }
        if (use_a) { et = et * a; ot = ot * a; }
        double *ye = &y[(2 * ((long)i - oy)) * sy];
        double *yo = &y[(2 * ((long)i - oy) + 1) * sy];
        if (use_ay) { et = (*ye) * a_y + et; ot = (*yo) * a_y + ot; }
        *ye = et;
        *yo = ot;
    }
}

static int check_common(int d, int idim, const int *n, const int *nx, const int *ny, int narr) {
    if (d < 1 || d > 16 || idim < 0 || idim >= d || narr < 0) return -1;
    for (int i = 0; i < d; ++i)
        if (n[i] < 1 || nx[i] < 1 || ny[i] < 1) return -1;
    return 0;
}

/* libconv 1D dispatcher (GenericConvolutionOperator1d#procedure,
 * GenericConvolution2.rb:1560-1766): ndat_left/right are products of nx[]
 * of the other dims (x2 per dim for wavelets); array j of the batch starts
 * at j*nti (input) and j*nto (output); a == 1 skips the scaling, a_x == 0
 * the a_x term and a_y == 0 the read of y (:1650-1699). */
int lco_s0s0_1d(int fil, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y,
                double a, double a_x, double a_y, double *dot) {
    cfilter cf;
    if (get_cfilter(fil, &cf) || check_common(d, idim, n, nx, ny, narr)) return -1;
    if (bc != BC_PER && bc != BC_GROW && bc != BC_SHRINK && bc != BC_FREE) return -1;
    int L = cf.up - cf.lo + 1;
    int nn = n[idim];
    if ((bc == BC_PER || bc == BC_FREE) && (nx[idim] < nn || ny[idim] < nn)) return -1;
    if (bc == BC_GROW && (nx[idim] < nn || ny[idim] < nn + L - 1)) return -1;
    if (bc == BC_SHRINK && (nx[idim] < nn + L - 1 || ny[idim] < nn)) return -1;
    long left = 1, right = 1;
    for (int i = 0; i < idim; ++i) left *= nx[i];
    for (int i = idim + 1; i < d; ++i) right *= nx[i];
    long nti = (long)nx[idim] * left * right, nto = (long)ny[idim] * left * right;
    int ox = in_origin(bc, cf.lo), oy = out_origin(bc, cf.up);
    int use_a = a != 1.0, use_ax = a_x != 0.0, use_ay = a_y != 0.0;
    double last_dot = 0.0;
    for (int j = 0; j < narr; ++j) {
        const double *xa = x + j * nti;
        double *ya = y + j * nto;
        double acc = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : acc) schedule(static)
        for (long r = 0; r < right; ++r)
            for (long l = 0; l < left; ++l) {
                const double *xl = xa + l + r * left * nx[idim];
                double *yl = ya + l + r * left * ny[idim];
                acc += conv_line(&cf, bc, nn, xl, left, ox, yl, left, oy,
                                 use_a, a, use_ax, a_x, use_ay, a_y, dot != NULL);
            }
        last_dot = acc;
    }
    if (dot) *dot = last_dot;
    return 0;
}

static int wavelet_1d(int recompose, int wav, int d, int idim, const int *n, int bc,
                      const int *nx, const int *ny, int narr,
                      const double *x, double *y, double a, double a_y) {
    wfilter wf;
    if (check_common(d, idim, n, nx, ny, narr)) return -1;
    if (bc != BC_PER && bc != BC_GROW && bc != BC_SHRINK) return -1;
    if (get_wfilter(wav, recompose, bc, &wf)) return -1;
    int W = wf.up - wf.lo + 1; /* filter length in pairs */
    int nn = n[idim];
    if (bc == BC_PER && (nx[idim] < nn || ny[idim] < nn)) return -1;
    if (bc == BC_GROW && (nx[idim] < nn || ny[idim] < nn + W - 1)) return -1;
    if (bc == BC_SHRINK && (nx[idim] < nn + W - 1 || ny[idim] < nn)) return -1;
    long left = 1, right = 1;
    for (int i = 0; i < idim; ++i) left *= 2L * nx[i];
    for (int i = idim + 1; i < d; ++i) right *= 2L * nx[i];
    long nti = 2L * nx[idim] * left * right, nto = 2L * ny[idim] * left * right;
    int ox = in_origin(bc, wf.lo), oy = out_origin(bc, wf.up);
    int use_a = a != 1.0, use_ay = a_y != 0.0;
    for (int j = 0; j < narr; ++j) {
        const double *xa = x + j * nti;
        double *ya = y + j * nto;
#pragma omp parallel for collapse(2) schedule(static)
        for (long r = 0; r < right; ++r)
            for (long l = 0; l < left; ++l) {
                const double *xl = xa + l + r * left * 2L * nx[idim];
                double *yl = ya + l + r * left * 2L * ny[idim];
                if (recompose)
                    idwt_line(&wf, bc, nn, xl, left, ox, nx[idim], yl, left, oy, ny[idim],
                              use_a, a, use_ay, a_y);
                else
                    dwt_line(&wf, bc, nn, xl, left, ox, nx[idim], yl, left, oy, ny[idim],
                             use_a, a, use_ay, a_y);
            }
    }
    return 0;
}

int lco_s0s1_1d(int wav, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y) {
    return wavelet_1d(0, wav, d, idim, n, bc, nx, ny, narr, x, y, a, a_y);
}

int lco_s1s0_1d(int wav, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y) {
    return wavelet_1d(1, wav, d, idim, n, bc, nx, ny, narr, x, y, a, a_y);
}

static long prod(const int *v, int d, int mul) {
    long p = 1;
    for (int i = 0; i < d; ++i) p *= (long)mul * v[i];
    return p;
}

/* Composition over all dims (GenericConvolutionOperator#procedure,
 * GenericConvolution2.rb:2160-2240): sweep i reads the array whose dims
 * < i already carry the output extent ny[] and dims >= i the input extent
 * nx[], so the 1D dispatcher is called with cur[] = (ny[0..i-1], nx[i..]).
 * kind 0: convolution filter, 1: wavelet analysis, 2: wavelet synthesis. */
static int compose_md(int kind, int id, int d, const int *n, const int *bc,
                      const int *nx, const int *ny, int narr,
                      const double *x, double *y, double a, double a_x, double a_y) {
    if (d < 1 || d > 16 || narr < 0) return -1;
    int mul = kind == 0 ? 1 : 2;
    long ntot_in = prod(nx, d, mul), ntot_out = prod(ny, d, mul);
    if (a_x != 0.0 && ntot_in != ntot_out) return -1;
    /* largest intermediate */
    long wmax = 1;
    int cur[16], nxt[16];
    for (int i = 0; i < d; ++i) cur[i] = nx[i];
    for (int s = 0; s < d - 1; ++s) {
        for (int i = 0; i < d; ++i) nxt[i] = cur[i];
        nxt[s] = ny[s];
        long p = prod(nxt, d, mul);
        if (p > wmax) wmax = p;
        for (int i = 0; i < d; ++i) cur[i] = nxt[i];
    }
    /* intermediates start zeroed: their padding (storage beyond the logical
     * output) is never mixed into valid outputs, but stays finite */
    double *w1 = (double *)calloc((size_t)wmax, sizeof(double));
    double *w2 = (double *)calloc((size_t)wmax, sizeof(double));
    if (!w1 || !w2) { free(w1); free(w2); return -1; }
    int rc = 0;
    for (int j = 0; j < narr && rc == 0; ++j) {
        const double *src = x + j * ntot_in;
        double *ya = y + j * ntot_out;
        for (int i = 0; i < d; ++i) cur[i] = nx[i];
        for (int s = 0; s < d && rc == 0; ++s) {
            const int last = s == d - 1;
            double *dst = last ? ya : ((s % 2 == 0) ? w1 : w2);
            double sa = last ? a : 1.0, say = last ? a_y : 0.0;
            int nxa[16], nya[16];
            for (int i = 0; i < d; ++i) { nxa[i] = cur[i]; nya[i] = cur[i]; }
            nya[s] = ny[s];
            if (kind == 0)
                rc = lco_s0s0_1d(id, d, s, n, bc[s], nxa, nya, 1, src, dst, sa, 0.0, say, NULL);
            else if (kind == 1)
                rc = lco_s0s1_1d(id, d, s, n, bc[s], nxa, nya, 1, src, dst, sa, say);
            else
                rc = lco_s1s0_1d(id, d, s, n, bc[s], nxa, nya, 1, src, dst, sa, say);
            cur[s] = ny[s];
            src = dst;
        }
        if (rc) break;
        /* y += a_x x (same shape only) */
        if (a_x != 0.0) {
            const double *xa = x + j * ntot_in;
#pragma omp parallel for schedule(static)
            for (long p = 0; p < ntot_out; ++p) ya[p] = xa[p] * a_x + ya[p];
        }
    }
    free(w1); free(w2);
    return rc;
}

int lco_s0s0_md(int fil, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y,
                double a, double a_x, double a_y) {
    cfilter cf;
    if (get_cfilter(fil, &cf)) return -1;
    return compose_md(0, fil, d, n, bc, nx, ny, narr, x, y, a, a_x, a_y);
}

int lco_s0s1_md(int wav, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y) {
    return compose_md(1, wav, d, n, bc, nx, ny, narr, x, y, a, 0.0, a_y);
}

int lco_s1s0_md(int wav, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y) {
    return compose_md(2, wav, d, n, bc, nx, ny, narr, x, y, a, 0.0, a_y);
}

/* Kinetic-style sum over dims: y = sum_i a[i] F_i x~ + a_x x~ + a_y y on
 * the output box, where x~ is x wrapped along periodic dims and zero
 * outside its box along free dims.  For periodic BCs this is the chain of
 * GenericConvolutionOperator with :kinetic (w1 = a0 Fx x + a_x x,
 * w2 = a1 Fy x + w1, y = a2 Fz x + w2 + a_y y; GenericConvolution2.rb
 * :2160-2240) and kinetic_per_ref (BigDFT/Kinetic.rb:981-1061); the
 * generated free-BC chain reads x with the grown extents of w1, so free
 * BCs follow this definition instead.  dot[i] = <x, a[i] F_i x> over output
 * points that carry an input value (Kinetic.rb:21-112 kstrten). */
int lco_s0s0_1ds(int fil, int d, const int *n, const int *bc,
                 const int *nx, const int *ny, int narr,
                 const double *a, double a_x, double a_y,
                 const double *x, double *y, double *dot) {
    cfilter cf;
    if (get_cfilter(fil, &cf) || d < 1 || d > 16 || narr < 0) return -1;
    const int lo = cf.lo, up = cf.up, L = up - lo + 1;
    int o_lo[16], o_hi[16], ox[16], oy[16];
    long sx[16], sy[16];
    for (int i = 0; i < d; ++i) {
        if (n[i] < 1) return -1;
        if (bc[i] == BC_PER || bc[i] == BC_FREE) {
            if (nx[i] < n[i] || ny[i] < n[i]) return -1;
            o_lo[i] = 0; o_hi[i] = n[i] - 1;
        } else if (bc[i] == BC_GROW) {
            if (nx[i] < n[i] || ny[i] < n[i] + L - 1) return -1;
            o_lo[i] = -up; o_hi[i] = n[i] - 1 - lo;
        } else if (bc[i] == BC_SHRINK) {
            if (nx[i] < n[i] + L - 1 || ny[i] < n[i]) return -1;
            o_lo[i] = 0; o_hi[i] = n[i] - 1;
        } else {
            return -1;
        }
        ox[i] = in_origin(bc[i], lo);
        oy[i] = out_origin(bc[i], up);
    }
    sx[0] = sy[0] = 1;
    for (int i = 1; i < d; ++i) { sx[i] = sx[i - 1] * nx[i - 1]; sy[i] = sy[i - 1] * ny[i - 1]; }
    long ntot_in = prod(nx, d, 1), ntot_out = prod(ny, d, 1);
    long npts = 1;
    for (int i = 0; i < d; ++i) npts *= (o_hi[i] - o_lo[i] + 1);
    double last[16] = {0};
    for (int j = 0; j < narr; ++j) {
        const double *xa = x + j * ntot_in;
        double *ya = y + j * ntot_out;
        double dsum[16] = {0};
#pragma omp parallel
        {
            double dl[16] = {0};
#pragma omp for schedule(static)
            for (long p = 0; p < npts; ++p) {
                int idx[16];
                long q = p;
                for (int i = 0; i < d; ++i) {
                    long ext = o_hi[i] - o_lo[i] + 1;
                    idx[i] = o_lo[i] + (int)(q % ext);
                    q /= ext;
                }
                /* centre value (zero where a grown output has no input) */
                int has_x = 1;
                long cpos = 0;
                for (int i = 0; i < d; ++i) {
                    if (bc[i] == BC_GROW && (idx[i] < 0 || idx[i] >= n[i])) has_x = 0;
                    cpos += (long)(idx[i] - ox[i]) * sx[i];
                }
                double xc = has_x ? xa[cpos] : 0.0;
                double total = 0.0;
                for (int k = 0; k < d; ++k) {
                    /* base offset without dim k */
                    long base = 0;
                    int valid = 1;
                    for (int i = 0; i < d; ++i) {
                        if (i == k) continue;
                        if (bc[i] == BC_GROW && (idx[i] < 0 || idx[i] >= n[i])) valid = 0;
                        base += (long)(idx[i] - ox[i]) * sx[i];
                    }
                    double t = 0.0;
                    if (valid) {
                        for (int l = lo; l <= up; ++l) {
                            long jj = (long)idx[k] + l;
                            long pos;
                            if (bc[k] == BC_PER) pos = pmod(jj, n[k]);
                            else if (bc[k] == BC_GROW || bc[k] == BC_FREE) {
                                if (jj < 0 || jj >= n[k]) continue;
                                pos = jj;
                            }
                            else pos = jj - ox[k];
                            t = xa[base + pos * sx[k]] * cf.f[l - lo] + t;
                        }
                    }
                    if (a[k] != 1.0) t = t * a[k];
                    if (has_x) dl[k] = xc * t + dl[k];
                    if (k == 0) {
                        if (a_x != 0.0 && has_x) t = xc * a_x + t;
                        total = t;
                    } else {
                        total = t + total;
                    }
                }
                long opos = 0;
                for (int i = 0; i < d; ++i) opos += (long)(idx[i] - oy[i]) * sy[i];
                if (a_y != 0.0) total = a_y * ya[opos] + total;
                ya[opos] = total;
            }
#pragma omp critical
            for (int k = 0; k < d; ++k) dsum[k] += dl[k];
        }
        for (int k = 0; k < d; ++k) last[k] = dsum[k];
    }
    if (dot)
        for (int k = 0; k < d; ++k) dot[k] = last[k];
    return 0;
}


/* ------------------------------------------------------------------ Poisson
 * First-derivative filters of order nord (BigDFT/Poisson2.rb generate_filter;
 * GenericConvolution2.rb:237-294 PoissonFilter).  Row k of the table (k in
 * [-m, m], m = nord / 2) weighs the inputs of the (nord + 1)-point stencil on
 * the nodes j - k, j in [-m, m].  Along a line of n points:
 *   periodic:      y[i] = sum_j c(j, 0) x[(i + j) mod n];
 *   non-periodic:  i < m:       y[i] = sum_j c(j, i - m)         x[j + m],
 *                  i >= n - m:  y[i] = sum_j c(j, i - n + m + 1) x[n - 1 - m + j],
 *                  otherwise    y[i] = sum_j c(j, 0)             x[i + j]
 * (div_u_i_ref, Poisson2.rb:183-335, and PoissonFilter#set_filter_val /
 * #input_filter_index).  The weights are exact Lagrange weights (lc_taps.h,
 * generated by gen_taps.py); they equal the reference's decimal table except
 * c(0, -4) of nord 8, which the reference gives as +35/2 instead of -35/2. */
#define PQ(p, q) ((double)(p) / (double)(q)),
static const double PT2[] = {LC_POISSON_2(PQ)};
static const double PT4[] = {LC_POISSON_4(PQ)};
static const double PT6[] = {LC_POISSON_6(PQ)};
static const double PT8[] = {LC_POISSON_8(PQ)};
static const double PT16[] = {LC_POISSON_16(PQ)};
#undef PQ

static const double *poisson_table(int nord) {
    switch (nord) {
        case 2: return PT2;
        case 4: return PT4;
        case 6: return PT6;
        case 8: return PT8;
        case 16: return PT16;
    }
    return NULL;
}

int lco_poisson_taps(int nord, double *out) {
    const double *t = poisson_table(nord);
    if (!t) return -1;
    memcpy(out, t, sizeof(double) * (nord + 1) * (nord + 1));
    return 0;
}

/* derivative (unscaled) at point i of a line of n points, stride s */
static double pderiv(const double *tab, int m, const double *base, long s, int n, int i, int per) {
    const int w = 2 * m + 1;
    double t = 0.0;
    if (per) {
        const double *c = tab + (long)m * w;
        for (int j = -m; j <= m; ++j) t = t + c[j + m] * base[pmod((long)i + j, n) * s];
    } else if (i < m) {
        const double *c = tab + (long)i * w;              /* row i - m */
        for (int j = -m; j <= m; ++j) t = t + c[j + m] * base[(long)(j + m) * s];
    } else if (i >= n - m) {
        const double *c = tab + (long)(i - n + 2 * m + 1) * w;  /* row i - n + m + 1 */
        for (int j = -m; j <= m; ++j) t = t + c[j + m] * base[(long)(n - 1 - m + j) * s];
    } else {
        const double *c = tab + (long)m * w;
        for (int j = -m; j <= m; ++j) t = t + c[j + m] * base[((long)i + j) * s];
    }
    return t;
}

static int poisson_dims_ok(int nord, const int *n) {
    if (!poisson_table(nord) || !n) return 0;
    for (int k = 0; k < 3; ++k)
        if (n[k] < nord + 1) return 0;  /* div_u_i_ref: the grid exceeds the stencil */
    return 1;
}

/* Poisson_broker (Poisson2.rb:1951-2043): du = a D_idim u (+ a_y du), idim
 * 0-based, bc 0 or -2, on a 3D array n */
int lco_poisson_1d(int nord, int idim, const int *n, int bc, double a, double a_y,
                   const double *u, double *du) {
    if (!poisson_dims_ok(nord, n) || idim < 0 || idim > 2 || (bc != 0 && bc != -2)) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    const long s[3] = {1, n[0], (long)n[0] * n[1]};
    const long np = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        const int i[3] = {(int)(p % n[0]), (int)((p / n[0]) % n[1]), (int)(p / ((long)n[0] * n[1]))};
        const double *base = u + (p - (long)i[idim] * s[idim]);
        double t = a * pderiv(tab, m, base, s[idim], n[idim], i[idim], bc == 0);
        if (a_y != 0.0) t = a_y * du[p] + t;
        du[p] = t;
    }
    return 0;
}

static void geo_per(int geocode, int per[3]) {
    per[0] = geocode != 1;
    per[1] = geocode == 2;
    per[2] = geocode != 1;
}

/* dd[k] = D_k f / h[k] at point p of f (division after the sum, as
 * div_u_i_ref does) */
static void grad_at(const double *tab, int m, const int per[3], const int *n, const double *f, long p,
                    const double *h, int which, double dd[3]) {
    const long s[3] = {1, n[0], (long)n[0] * n[1]};
    const int i[3] = {(int)(p % n[0]), (int)((p / n[0]) % n[1]), (int)(p / ((long)n[0] * n[1]))};
    for (int k = 0; k < 3; ++k) {
        if (!(which & (1 << k))) continue;
        const double *base = f + (p - (long)i[k] * s[k]);
        dd[k] = pderiv(tab, m, base, s[k], n[k], i[k], per[k]) / h[k];
    }
}

/* div_u_i (Poisson2.rb:183-335 div_u_i_ref; cc as div_u_i_opt :884-1247 and
 * the BOAST composition :2067-2195):
 *   du = d_x u_x + d_y u_y + d_z u_z,
 *   cc = u_x^2 d_x u_x + u_y^2 d_y u_y + u_z^2 d_z u_z
 *        + 2 (u_x u_y d_x u_y + u_x u_z d_x u_z + u_y u_z d_y u_z)   (if cc) */
int lco_div_u_i(int geocode, const int *n, const double *u, double *du, int nord,
                const double *h, double *cc) {
    if (!poisson_dims_ok(nord, n) || !h) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    int per[3];
    geo_per(geocode, per);
    const long np = (long)n[0] * n[1] * n[2];
    const double *u0 = u, *u1 = u + np, *u2 = u + 2 * np;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double a[3], b[3], c[3];
        grad_at(tab, m, per, n, u0, p, h, 1, a);
        grad_at(tab, m, per, n, u1, p, h, cc ? 3 : 2, b);
        grad_at(tab, m, per, n, u2, p, h, cc ? 7 : 4, c);
        du[p] = a[0] + b[1] + c[2];
        if (cc) {
            const double x = u0[p], y = u1[p], z = u2[p];
            cc[p] = (x * x) * a[0] + (y * y) * b[1] + (z * z) * c[2] +
                    2.0 * (x * y * b[0] + x * z * c[0] + y * z * c[1]);
        }
    }
    return 0;
}

/* nabla_u (Poisson2.rb:722-874 nabla_u_ref): du(:, :, :, k) = D_k u / h_k */
int lco_nabla_u(int geocode, const int *n, const double *u, double *du, int nord,
                const double *h) {
    if (!poisson_dims_ok(nord, n) || !h) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    int per[3];
    geo_per(geocode, per);
    const long np = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double g[3];
        grad_at(tab, m, per, n, u, p, h, 7, g);
        du[p] = g[0];
        du[np + p] = g[1];
        du[2 * np + p] = g[2];
    }
    return 0;
}

/* nabla_u_and_square (Poisson2.rb:361-512): the gradient and
 * du2 = |grad u|^2 (components squared and summed in order x, y, z) */
int lco_nabla_u_and_square(int geocode, const int *n, const double *u, double *du,
                           double *du2, int nord, const double *h) {
    int rc = lco_nabla_u(geocode, n, u, du, nord, h);
    if (rc) return rc;
    const long np = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double t = du[p] * du[p];
        t = t + du[np + p] * du[np + p];
        t = t + du[2 * np + p] * du[2 * np + p];
        du2[p] = t;
    }
    return 0;
}

/* nabla_u_square (Poisson2.rb:2278-2359, nabla_u_square_ref :1690-1884):
 * ddu = (d_x^2 + d_y^2) + d_z^2, the gradient itself not kept */
int lco_nabla_u_square(int geocode, const int *n, const double *u, double *ddu, int nord,
                       const double *h) {
    if (!poisson_dims_ok(nord, n) || !h) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    int per[3];
    geo_per(geocode, per);
    const long np = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double g[3];
        grad_at(tab, m, per, n, u, p, h, 7, g);
        double t = g[0] * g[0];
        t = t + g[1] * g[1];
        ddu[p] = t + g[2] * g[2];
    }
    return 0;
}

/* nabla_u_epsilon (Poisson2.rb:2362-2443, nabla_u_epsilon_ref :1488-1680):
 * du(:, :, :, k) = (D_k u / h_k) eps */
int lco_nabla_u_epsilon(int geocode, const int *n, const double *u, double *du, int nord,
                        const double *h, const double *eps) {
    if (!poisson_dims_ok(nord, n) || !h || !eps) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    int per[3];
    geo_per(geocode, per);
    const long np = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double g[3];
        grad_at(tab, m, per, n, u, p, h, 7, g);
        for (int k = 0; k < 3; ++k) du[k * np + p] = g[k] * eps[p];
    }
    return 0;
}

/* update_rhopol (Poisson2.rb:2446-2550, update_rhopol_ref :520-699): per
 * point res = ((e1 d_x + e2 d_y) + e3 d_z) / (4 pi) with (e1, e2, e3) =
 * dlogeps(:, p); res = eta (res - rhopol); rhopol = res + rhopol; rhores2 =
 * the sum of res^2 over the points in storage order */
int lco_update_rhopol(int geocode, const int *n, const double *u, int nord, const double *h,
                      double eta, const double *dlogeps, double *rhopol, double *rhores2) {
    if (!poisson_dims_ok(nord, n) || !h || !dlogeps || !rhopol || !rhores2) return -1;
    const double *tab = poisson_table(nord);
    const int m = nord / 2;
    const double oneo4pi = 0.07957747154594766788444188168625718101722982287022822437383367203;
    int per[3];
    geo_per(geocode, per);
    const long np = (long)n[0] * n[1] * n[2];
    double *sq = (double *)malloc((size_t)(np > 0 ? np : 1) * sizeof(double));
    if (!sq) return -1;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < np; ++p) {
        double g[3];
        grad_at(tab, m, per, n, u, p, h, 7, g);
        const double *e = dlogeps + 3 * p;
        double res = e[0] * g[0] + e[1] * g[1];
        res = res + e[2] * g[2];
        res = res * oneo4pi;
        const double rho = rhopol[p];
        res = (res - rho) * eta;
        sq[p] = res * res;
        rhopol[p] = res + rho;
    }
    double s = 0.0;
    for (long p = 0; p < np; ++p) s += sq[p];
    free(sq);
    *rhores2 = s;
    return 0;
}

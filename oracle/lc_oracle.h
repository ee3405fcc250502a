/*
 * lc_oracle.h -- FP64 CPU oracle of the libconv separable convolutions.
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement of the reference algorithm is
 * the checker for the HIP library and the CPU baseline that bench.py times;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (boast-lig_amd/) never links or calls it.
 *
 * Semantics follow the BOAST generator (BigDFT/GenericConvolution2.rb):
 *   - 1D sweep with boundary conditions, ld (nx/ny) and the a / a_x / a_y /
 *     dot_in epilogue: ConvolutionOperator1d :1036-1442, filters :91-234;
 *   - sym8 wavelet analysis/synthesis: WaveletFilter :296-671;
 *   - libconv 1D dispatcher (ndat from nx[] of the other dims, narr
 *     batching): GenericConvolutionOperator1d :1444-1766;
 *   - multi-dimensional composition: GenericConvolutionOperator :1768-2245.
 * Boundary conditions: 0 periodic, 1 free grow, -1 free shrink
 * (BoundaryConditions :684-740).  Where the generated code reads out of
 * bounds (free BC with n below the filter width, C modulo of a < -b) the
 * oracle follows the mathematical definition: zero outside the input for
 * free BCs, true modulo for periodic.
 */
#ifndef LC_ORACLE_H
#define LC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* convolution filters */
enum { LCO_MD = 0, LCO_IMD = 1, LCO_D2 = 2, LCO_D1 = 3 };
/* wavelet families */
enum { LCO_SYM8 = 0, LCO_SYMICOMB8 = 1 };

/* 1D sweep of a convolution filter along dim idim (libconv s0s0_1d).
 * dot (may be NULL) receives <x, a*fil(x)> of the LAST array of the narr
 * batch, as the generated dispatcher leaves it.  Returns 0 or -1. */
int lco_s0s0_1d(int fil, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y,
                double a, double a_x, double a_y, double *dot);

/* 1D wavelet analysis (s0s1) / synthesis (s1s0) along dim idim; n, nx, ny
 * count pairs, every array extent is 2*nx[i] (2*ny[idim] for the output). */
int lco_s0s1_1d(int wav, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y);
int lco_s1s0_1d(int wav, int d, int idim, const int *n, int bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y);

/* separable product over all d dims: y = a * F_{d-1} ... F_0 x + a_x x
 * + a_y y, per-dim BC; x extents nx[], y extents ny[]. */
int lco_s0s0_md(int fil, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y,
                double a, double a_x, double a_y);

/* sum over dims (kinetic form): y = sum_i a[i] F_i x + a_x x + a_y y;
 * dot[i] (may be NULL) = <x, a[i] F_i x> over output points that carry an
 * input value, of the last array of the batch. */
int lco_s0s0_1ds(int fil, int d, const int *n, const int *bc,
                 const int *nx, const int *ny, int narr,
                 const double *a, double a_x, double a_y,
                 const double *x, double *y, double *dot);

/* multi-dimensional wavelet analysis / synthesis (all d dims, dim 0
 * first), y = a * W x + a_y y. */
int lco_s0s1_md(int wav, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y);
int lco_s1s0_md(int wav, int d, const int *n, const int *bc,
                const int *nx, const int *ny, int narr,
                const double *x, double *y, double a, double a_y);

/* tap access for tests: copies the taps of a convolution filter (returns
 * the count, writes the first tap index to *lo). */
int lco_filter_taps(int fil, int *lo, double *out, int cap);

/* number of OpenMP threads the oracle runs with (1 without OpenMP). */
int lco_num_threads(void);

/* Poisson nabla operators (BigDFT/Poisson2.rb; filters of
 * GenericConvolution2.rb:237-294).  nord in {2, 4, 6, 8, 16}; n[3] the
 * grid; bc 0 periodic or -2 non-periodic (one-sided rows at the borders);
 * geocode 1 free in x, y, z; 2 periodic in all; any other value periodic in
 * x and z, free in y (div_u_i_ref's perx / pery / perz).  Vector fields are
 * (n0, n1, n2, 3) arrays, component c at offset c n0 n1 n2. */
int lco_poisson_taps(int nord, double *out);  /* (nord+1)^2 values, row k outer */
int lco_poisson_1d(int nord, int idim, const int *n, int bc, double a, double a_y,
                   const double *u, double *du);
int lco_div_u_i(int geocode, const int *n, const double *u, double *du, int nord,
                const double *h, double *cc);
int lco_nabla_u(int geocode, const int *n, const double *u, double *du, int nord,
                const double *h);
int lco_nabla_u_and_square(int geocode, const int *n, const double *u, double *du,
                           double *du2, int nord, const double *h);
/* ddu = |grad u|^2 only */
int lco_nabla_u_square(int geocode, const int *n, const double *u, double *ddu, int nord,
                       const double *h);
/* du(:, :, :, k) = eps D_k u / h_k */
int lco_nabla_u_epsilon(int geocode, const int *n, const double *u, double *du, int nord,
                        const double *h, const double *eps);
/* res = (dlogeps . grad u) / (4 pi); res = eta (res - rhopol);
 * rhopol += res; *rhores2 = sum of res^2 (points in storage order);
 * dlogeps (3, n0, n1, n2), the component fastest */
int lco_update_rhopol(int geocode, const int *n, const double *u, int nord, const double *h,
                      double eta, const double *dlogeps, double *rhopol, double *rhores2);

#ifdef __cplusplus
}
#endif
#endif

/*
 * libconv.h -- C ABI of the MI355X (gfx950) libconv library.
 *
 * Drop-in replacement for the procedures that the BOAST generator of the
 * reference emits for BigDFT's separable convolutions
 * (BigDFT/GenericConvolution2.rb), plus the spellings of the intended
 * public interface (BigDFT/libconv.f90:303-456) and a stream-aware API for
 * device-resident data.  Arrays are Fortran-ordered (dim 0 fastest).
 *
 * Conventions of the drop-in names (BOAST C flavour, GenericConvolution2.rb
 * :1444-1766 and :1768-2245):
 *   - int and real scalars by value, arrays by pointer, out-scalars by
 *     pointer; d_ = FP64, s_ = FP32 storage and arithmetic;
 *   - n[] logical sizes (pairs for wavelets), nx[]/ny[] the input/output
 *     extents along each dim (the "ld" leading dims; the non-processed dims
 *     use nx[] for both arrays), narr arrays of the same shape back to back;
 *   - bc: 0 periodic, 1 free grow (output extended by the filter length),
 *     -1 free shrink (input extended);
 *   - a == 1 skips the scaling, a_x == 0 the a_x * x term and a_y == 0 the
 *     read of y, exactly as the generated dispatcher selects its variants;
 *   - pointers may be host or device memory (detected per call); results
 *     are complete on return; the functions are reentrant per host thread;
 *   - the reference procedures return nothing and check nothing: these
 *     validate their arguments, leave y untouched on error and record the
 *     error for lc_last_error() (LIBCONV_ABORT_ON_ERROR=1 aborts instead).
 * dot_in follows the generated dispatcher: <x, a fil(x)> of the last array
 * of the narr batch, accumulated in FP64 on the GPU.
 *
 * Free BCs with n below the filter width: the generated sweeps read out of
 * bounds there; this library treats x as zero outside its box (both sides
 * clamped) and wraps periodic indices with a true modulo.
 */
#ifndef LIBCONV_H
#define LIBCONV_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum {
  LC_OK = 0,
  LC_ERR_ARG = 1,      /* invalid dims / extents / bc / ids */
  LC_ERR_UNSUPPORTED = 2,
  LC_ERR_DEVICE = 3,   /* HIP runtime error (message holds the HIP text) */
  LC_ERR_NOMEM = 4
};
int lc_last_error(void);                 /* last error of this host thread, then cleared */
const char *lc_last_error_message(void); /* text of the last error of this host thread */

/* ------------------------------------------------- libconv 1D sweeps (C)
 * y := a * fil |X_idim| x + a_x * x + a_y * y     along dim idim
 * Replaces the generated dispatchers d_s0s0_1d_<filter>
 * (GenericConvolutionOperator1d#procedure, GenericConvolution2.rb:1560-1766;
 * built by MagicFilter1d, BigDFT/LibconvBase.rb:10-30, 238-248).        */
void d_s0s0_1d_sym8_md(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const double *x, double *y, double a, double a_x, double a_y);
void s_s0s0_1d_sym8_md(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const float *x, float *y, float a, float a_x, float a_y);
void d_s0s0_1d_sym8_imd(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                        int narr, const double *x, double *y, double a, double a_x, double a_y);
void s_s0s0_1d_sym8_imd(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                        int narr, const float *x, float *y, float a, float a_x, float a_y);
void d_s0s0_1d_sym8_d2(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const double *x, double *y, double a, double a_x, double a_y);
void s_s0s0_1d_sym8_d2(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const float *x, float *y, float a, float a_x, float a_y);
void d_s0s0_1d_sym8_d1(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const double *x, double *y, double a, double a_x, double a_y);
void s_s0s0_1d_sym8_d1(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                       int narr, const float *x, float *y, float a, float a_x, float a_y);
/* with dot_in = <x, a d2(x)> (Kinetic1d with :dot_in, BigDFT/Kinetic1d-nb.ipynb cell 1) */
void d_s0s0_1d_sym8_d2_dotin(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                             int narr, const double *x, double *y, double a, double a_x, double a_y,
                             double *dot_in);
void s_s0s0_1d_sym8_d2_dotin(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                             int narr, const float *x, float *y, float a, float a_x, float a_y,
                             float *dot_in);
/* FP32 storage with the FP64 energy (no counterpart in the reference) */
void s_s0s0_1d_sym8_d2_dotin64(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                               int narr, const float *x, float *y, float a, float a_x, float a_y,
                               double *dot_in);

/* ------------------------------------------------ libconv 1D wavelets (C)
 * s0s1: y := a * dwt |X_idim| x + a_y * y   (analysis, output [L | H])
 * s1s0: y := a * idwt |X_idim| x + a_y * y  (synthesis)
 * Replace d_s0s1_1d_sym8 / d_s1s0_1d_sym8 (Wavelet1d, BigDFT/LibconvBase.rb
 * :31-49, 226-236; generated listing BigDFT/Wavelet1d-nb.ipynb cell 4) and
 * the combined operators rtos1 = d_s0s1_1d_symicomb8 (DWT of the inverse
 * magic filter) and s1tor = d_s1s0_1d_symicomb8 (magic filter of the IDWT)
 * (BigDFT/LibconvBase.rb:250-262).                                         */
void d_s0s1_1d_sym8(int d, int idim, const int *n, int bc, const int *nx, const int *ny, int narr,
                    const double *x, double *y, double a, double a_y);
void s_s0s1_1d_sym8(int d, int idim, const int *n, int bc, const int *nx, const int *ny, int narr,
                    const float *x, float *y, float a, float a_y);
void d_s1s0_1d_sym8(int d, int idim, const int *n, int bc, const int *nx, const int *ny, int narr,
                    const double *x, double *y, double a, double a_y);
void s_s1s0_1d_sym8(int d, int idim, const int *n, int bc, const int *nx, const int *ny, int narr,
                    const float *x, float *y, float a, float a_y);
void d_s0s1_1d_symicomb8(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const double *x, double *y, double a, double a_y);
void s_s0s1_1d_symicomb8(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const float *x, float *y, float a, float a_y);
void d_s1s0_1d_symicomb8(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const double *x, double *y, double a, double a_y);
void s_s1s0_1d_symicomb8(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const float *x, float *y, float a, float a_y);

/* ------------------------------------------------------- cost companions
 * Flop count of a call (the generated *_cost procedures,
 * GenericConvolution2.rb:1269-1276, 1561-1568): narr * n[idim] * ndat *
 * (2 L for a filter of L taps, 64 per pair for sym8 wavelets).            */
void d_s0s0_1d_sym8_md_cost(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                            int narr, const double *x, double *y, double a, double a_x, double a_y,
                            int *cost);
void d_s0s0_1d_sym8_d2_dotin_cost(int d, int idim, const int *n, int bc, const int *nx,
                                  const int *ny, int narr, const double *x, double *y, double a,
                                  double a_x, double a_y, double *dot_in, int *cost);
void d_s0s1_1d_sym8_cost(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const double *x, double *y, double a, double a_y, int *cost);
void d_s1s0_1d_sym8_cost(int d, int idim, const int *n, int bc, const int *nx, const int *ny,
                         int narr, const double *x, double *y, double a, double a_y, int *cost);

/* ------------------------------------------------- multi-dimensional (C)
 * Replace the GenericConvolutionOperator procedures (GenericConvolution2.rb
 * :2131-2245) built by MagicFilter (BigDFT/MagicFilter.rb:46-70: transpose 1,
 * work, ld, narr), Kinetic (BigDFT/Kinetic.rb:1068-1090: transpose 0, a,
 * a_x, a_y, work, kinetic, zero_out_work) and Wavelet (BigDFT/Wavelet2.rb:
 * 2-21: transpose 1, work, wavelet).  The work arrays w1/w2 are accepted
 * for ABI compatibility; the library keeps its own device scratch.       */
void d_sym8_md_ld_narr(int d, const int *n, const int *bc, const int *nx, const int *ny, int narr,
                       const double *x, double *y, double *w1, double *w2);
void s_sym8_md_ld_narr(int d, const int *n, const int *bc, const int *nx, const int *ny, int narr,
                       const float *x, float *y, float *w1, float *w2);
void d_sym8_imd_ld_narr(int d, const int *n, const int *bc, const int *nx, const int *ny, int narr,
                        const double *x, double *y, double *w1, double *w2);
void s_sym8_imd_ld_narr(int d, const int *n, const int *bc, const int *nx, const int *ny, int narr,
                        const float *x, float *y, float *w1, float *w2);
/* y = sum_i a[i] d2_i(x) + a_x x + a_y y  (x extents per BC from n) */
void d_sym8_d2(int d, const int *n, const int *bc, const double *x, double *y, double *w1,
               double *w2, const double *a, double a_x, double a_y);
void s_sym8_d2(int d, const int *n, const int *bc, const float *x, float *y, float *w1, float *w2,
               const float *a, float a_x, float a_y);
/* same with the per-direction energies dot_in[i] = <x, a[i] d2_i(x)> */
void d_sym8_d2_dotin(int d, const int *n, const int *bc, const double *x, double *y, double *w1,
                     double *w2, const double *a, double a_x, double a_y, double *dot_in);
void s_sym8_d2_dotin64(int d, const int *n, const int *bc, const float *x, float *y, float *w1,
                       float *w2, const float *a, float a_x, float a_y, double *dot_in);
void d_dwt_sym8(int d, const int *n, const int *bc, const double *x, double *y, double *w1,
                double *w2);
void s_dwt_sym8(int d, const int *n, const int *bc, const float *x, float *y, float *w1, float *w2);
void d_idwt_sym8(int d, const int *n, const int *bc, const double *x, double *y, double *w1,
                 double *w2);
void s_idwt_sym8(int d, const int *n, const int *bc, const float *x, float *y, float *w1,
                 float *w2);

/* --------------------------------------- libconv.f90 spellings (Fortran)
 * BigDFT/libconv.f90:303-456 (alpha release names), every argument by
 * reference, lower case with a trailing underscore.  s1s0_1d_sym8_lp_d
 * takes idim (the spec omits it, BigDFT/libconv.f90:320-332), bc_id is a
 * scalar for the 1D names.                                                 */
void s0s1_1d_sym8_lp_d_(const int *d, const int *idim, const int *n, const int *bc_id,
                        const int *nx, const int *ny, const int *narr, const double *a,
                        const double *a_y, const double *x, double *y);
void s1s0_1d_sym8_lp_d_(const int *d, const int *idim, const int *n, const int *bc_id,
                        const int *nx, const int *ny, const int *narr, const double *a,
                        const double *a_y, const double *x, double *y);
void s0s1_sym8_lp_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                     const int *narr, const double *a, const double *a_y, const double *x,
                     double *y, double *work);
void s1s0_sym8_lp_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                     const int *narr, const double *a, const double *a_y, const double *x,
                     double *y, double *work);
void s0s0_1d_sym8_md_d_(const int *d, const int *idim, const int *n, const int *bc_id,
                        const int *nx, const int *ny, const int *narr, const double *a,
                        const double *a_x, const double *a_y, const double *x, double *y);
void s0s0_1d_sym8_mi_d_(const int *d, const int *idim, const int *n, const int *bc_id,
                        const int *nx, const int *ny, const int *narr, const double *a,
                        const double *a_x, const double *a_y, const double *x, double *y);
void s0s0_1ds_sym8_d2_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                         const int *narr, const double *a, const double *a_x, const double *a_y,
                         const double *x, double *y, double *work);
void s0s0_1ds_dot_sym8_d2_d_(const int *d, const int *n, const int *bc, const int *nx,
                             const int *ny, const int *narr, const double *a, const double *a_x,
                             const double *a_y, const double *x, double *y, double *dot_in,
                             double *work);
void s0s0_1ds_sym8_d1_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                         const int *narr, const double *a, const double *a_x, const double *a_y,
                         const double *x, double *y, double *work);
void s0s0_sym8_md_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                     const int *narr, const double *a, const double *a_x, const double *a_y,
                     const double *x, double *y, double *work);
void s0s0_sym8_mi_d_(const int *d, const int *n, const int *bc, const int *nx, const int *ny,
                     const int *narr, const double *a, const double *a_x, const double *a_y,
                     const double *x, double *y, double *work);

/* ------------------------------------------- stream API (device memory)
 * New code: device pointers only, asynchronous on `stream` (a hipStream_t,
 * NULL = the default stream), int return (LC_OK or an error code).  prec is
 * 8 (FP64) or 4 (FP32 storage and arithmetic); scalars are given in double
 * and rounded to the storage precision; dot outputs are device pointers to
 * FP64 (d values for lc_conv_1ds), written asynchronously.  The library's
 * internal device scratch (energy partials, sweep intermediates) is kept
 * per host thread and per (device, stream), so calls on different streams
 * may run concurrently; calls on one stream are ordered by that stream.
 * fil: 0 magic filter (sym8_md), 1 inverse magic filter (sym8_imd),
 *      2 second derivative (sym8_d2), 3 first derivative (sym8_d1);
 * wav: 0 sym8, 1 symicomb8 (magic filter combined with the wavelet).       */
enum { LC_FIL_MD = 0, LC_FIL_IMD = 1, LC_FIL_D2 = 2, LC_FIL_D1 = 3 };
enum { LC_WAV_SYM8 = 0, LC_WAV_SYMICOMB8 = 1 };
int lc_conv_1d(int prec, int fil, int d, int idim, const int *n, int bc, const int *nx,
               const int *ny, int narr, const void *x, void *y, double a, double a_x, double a_y,
               double *dot_dev, void *stream);
int lc_wavelet_1d(int prec, int wav, int inverse, int d, int idim, const int *n, int bc,
                  const int *nx, const int *ny, int narr, const void *x, void *y, double a,
                  double a_y, void *stream);
/* separable product over all d dims: y = a fil_{d-1}..fil_0 x + a_x x + a_y y */
int lc_conv_md(int prec, int fil, int d, const int *n, const int *bc, const int *nx,
               const int *ny, int narr, const void *x, void *y, double a, double a_x, double a_y,
               void *stream);
/* sum over dims: y = sum_i a[i] fil_i x + a_x x + a_y y; a[] on the host */
int lc_conv_1ds(int prec, int fil, int d, const int *n, const int *bc, const int *nx,
                const int *ny, int narr, const double *a, double a_x, double a_y, const void *x,
                void *y, double *dot_dev, void *stream);
int lc_wavelet_md(int prec, int wav, int inverse, int d, const int *n, const int *bc,
                  const int *nx, const int *ny, int narr, const void *x, void *y, double a,
                  double a_y, void *stream);
/* lc_conv_1ds without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as lc_conv_1ds; the energy partials go to a static 16 MiB device
 * buffer of the library (one per GPU, held until the process ends) instead
 * of the pool, so it calls no hipMalloc / hipFree.  Shapes whose partials
 * may need more (output grids beyond about 440^3 points, with energies) are
 * refused with LC_ERR_UNSUPPORTED before the device is touched.  All calls on one GPU share that buffer: calls
 * that overlap on one GPU (other streams, other threads) race, so serialise
 * them.  The buffer keeps the last call's contents; every call writes all of
 * what it reads back.                                                        */
int lc_conv_1ds_noalloc(int prec, int fil, int d, const int *n, const int *bc, const int *nx,
                        const int *ny, int narr, const double *a, double a_x, double a_y,
                        const void *x, void *y, double *dot_dev, void *stream);
/* ------------------------------------------- batched 3D operators
 * No counterpart in the reference, which applies each operator to one
 * orbital at a time (BigDFT/GenericConvolution2.rb:2131-2245 composes one
 * array): this groups many independent 3D applications -- orbitals of
 * different sizes and boundary conditions -- into one launch per operator.
 * Items with the same kind and operator that follow each other form one
 * stage (one launch); stages run in item order, so a later stage may read
 * what an earlier one wrote.  Items inside a stage must not overlap.
 * All pointers are device pointers fixed at plan creation; running a plan
 * does no host work beyond the launches and allocates nothing, so a plan
 * can run inside a caller's stream capture.  lc_batch_run_graph captures
 * the plan into a HIP graph on its first call and replays it afterwards.
 *   SEP   y = a fil_2 fil_1 fil_0 x + a_y y          (a = a[0])
 *   SUM   y = sum_d a[d] fil_d x + a_x x + a_y y,  dot[d] = <x, a[d] fil_d x>
 *   DWT / IDWT  one 3D level of sym8 wavelet analysis / synthesis,
 *               y = a t + a_y y (LC_WAV_SYM8 only)
 * Sizes follow lc_conv_md / lc_wavelet_md (n, nx, ny per dim; wavelet
 * arrays hold 2 nx / 2 ny values per dim).                                */
enum { LC_BATCH_SEP = 0, LC_BATCH_SUM = 1, LC_BATCH_DWT = 2, LC_BATCH_IDWT = 3 };
typedef struct lc_batch_item {
  int kind;              /* LC_BATCH_* */
  int op;                /* LC_FIL_* (SEP, SUM) or LC_WAV_* (DWT, IDWT) */
  int n[3], bc[3], nx[3], ny[3];
  const void *x;         /* device */
  void *y;               /* device */
  double a[3];           /* SUM: a[d]; others: a[0] = a */
  double a_x, a_y;       /* a_x: SUM items only */
  double *dot;           /* SUM: device pointer to 3 FP64 energies, or NULL */
} lc_batch_item;
typedef struct lc_batch lc_batch;
int lc_batch_create(lc_batch **plan, int prec, int count, const lc_batch_item *items);
int lc_batch_run(lc_batch *plan, void *stream);
int lc_batch_run_graph(lc_batch *plan, void *stream);
int lc_batch_info(const lc_batch *plan, int *stages, long *blocks);
void lc_batch_destroy(lc_batch *plan);
/* The whole plan in ONE persistent launch (after a small reset kernel): a
 * device-side work queue per XCD hands out (item, block) units, items wait
 * on device counters for the items whose memory they depend on (found from
 * the pointers and extents at plan creation), so stages overlap across
 * items.  Items: SEP md/imd, SUM d2, sym8 DWT/IDWT.  Outputs equal
 * lc_batch_run's bit for bit; energies are summed in another grouping.
 * The _graph form captures the launch once and replays it.
 * lc_batch_persistent_info reports the grid and (synchronising) the status
 * word: non-zero when a dependency wait exceeded its bound
 * (LIBCONV_PERSIST_SPIN_LIMIT sleeps, default 2^24), in which case the
 * results of that run are not valid.                                        */
int lc_batch_run_persistent(lc_batch *plan, void *stream);
int lc_batch_run_persistent_graph(lc_batch *plan, void *stream);
int lc_batch_persistent_info(lc_batch *plan, int *grid, int *status);
/* Tunes every stage of the plan over its kernel variants (min of `repeat`
 * timed runs of the stage each, after one untimed run) and keeps the
 * fastest; stage_ms (may be NULL) receives the best time per stage.  The
 * runs write the plan's outputs (and accumulate where a_y != 0).           */
int lc_batch_tune(lc_batch *plan, int repeat, double *stage_ms);
int lc_batch_stage_variant(const lc_batch *plan, int stage);

/* ------------------------------------------- autotuner
 * In the spirit of BOAST's optimizer (GenericOptimization and
 * ConvolutionOperator1d#optimize, BigDFT/GenericConvolution2.rb:744-837,
 * :1223-1267): the compiled kernel variants of a 3D operator form (tile
 * shape, outputs per lane, window step, prefetch depth) each run on random
 * data of the given shape (rotating over operand sets larger than the 256 MB
 * last-level cache) in three rounds of `repeat` back-to-back launches;
 * the minimum round average per variant is kept and the fastest variant is
 * recorded for that exact shape, which later
 * calls of the shape then use.  times[] (may be NULL) receives the min time
 * of each variant (ms; infinity where a variant rejects the shape).  The
 * table is saved and loaded as JSON; LIBCONV_TUNE_FILE names a file loaded
 * at first use.  form: separable product (MagicFilter-style lc_conv_md), sum
 * form (Kinetic-style lc_conv_1ds), sym8 DWT or IDWT (lc_wavelet_md).      */
enum { LC_FORM_SEP = 0, LC_FORM_SUM = 1, LC_FORM_DWT = 2, LC_FORM_IDWT = 3 };
int lc_tune_md(int prec, int form, int op, const int *n, const int *bc, const int *nx,
               const int *ny, int repeat, int *best, double *best_ms, double *times, int ntimes);
int lc_tune_save(const char *path);
int lc_tune_load(const char *path);
void lc_tune_clear(void);
int lc_variant_count(int form);
/* name of the kernel a call of this shape launches (as profilers show it) */
int lc_kernel_name(int prec, int form, int op, const int *n, const int *bc, const int *nx,
                   const int *ny, char *buf, int len);

/* ------------------------------------------- local Hamiltonian (one orbital)
 * The chain libconv.f90:238-295 sketches for BigDFT's free-BC wavelets, on
 * device data, in stream order:
 *   psi_fscf = IDWT_grow psig                     (n -> n + 7 pairs per dim)
 *   psir     = MF_grow psi_fscf                   (2 (n + 7) -> 2 (n + 7) + 15)
 *   rho     += psir^2;  psir = pot psir           (rho may be NULL)
 *   hpsi     = IMF_shrink psir                    (back to 2 (n + 7))
 *   hpsi    += sum_d a[d] D2_d psi_fscf           (bc 2: free, output on the input box)
 *   hpsig    = DWT_shrink hpsi                    (n pairs)
 * psig / hpsig: (2 n)^3 values in the band layout of lc_wavelet_md; pot and
 * rho: (2 n + 29)^3 values; a[d] = -0.5 / h_d^2 on the fine grid; ekin (3
 * FP64, device, may be NULL) = <psi_fscf, a_d D2_d psi_fscf>.  work: device
 * array of lc_local_hamiltonian_work(n) elements of the precision (psi_fscf,
 * psir, hpsi), owned by the caller as the reference's work arrays are.     */
long lc_local_hamiltonian_work(const int *n);
int lc_local_hamiltonian(int prec, const int *n, const double *a, const void *psig, void *hpsig,
                         const void *pot, void *rho, void *work, double *ekin, void *stream);

/* ------------------------------------------- Poisson nabla operators
 * First derivatives of order nord in {2, 4, 6, 8, 16} (BigDFT/Poisson2.rb,
 * filters of GenericConvolution2.rb:237-294 PoissonFilter) on a 3D grid n[3]
 * (every n[k] > nord).  bc 0 periodic; bc -2 non-periodic: the first and
 * last nord/2 outputs of a line use one-sided stencils.  geocode (div_u_i_ref
 * convention) 1: x, y, z free; 2: all periodic; other: x, z periodic, y free.
 * Vector fields are (n0, n1, n2, 3) arrays, component k at offset k n0 n1 n2.
 * The stencil weights are the exact Lagrange weights; the reference's table
 * differs in one entry (nord 8, first output of a non-periodic line, where
 * its decimal table carries a sign typo).
 * Stream API (device memory, asynchronous, int return):                      */
/* du = a D_idim u (+ a_y du)  (Poisson_broker) */
int lc_poisson_1d(int prec, int nord, int idim, const int *n, int bc, const void *u, void *du, double a,
                  double a_y, void *stream);
/* du = sum_k D_k u_k / h_k; cc (may be NULL) = sum_k u_k^2 d_k u_k
 *      + 2 (u_x u_y d_x u_y + u_x u_z d_x u_z + u_y u_z d_y u_z)            */
int lc_div_u_i(int prec, int geocode, const int *n, const void *u, void *du, int nord, const double *h,
               void *cc, void *stream);
/* du(:, :, :, k) = D_k u / h_k */
int lc_nabla_u(int prec, int geocode, const int *n, const void *u, void *du, int nord, const double *h,
               void *stream);
/* the gradient and du2 = |grad u|^2 */
int lc_nabla_u_and_square(int prec, int geocode, const int *n, const void *u, void *du, void *du2, int nord,
                          const double *h, void *stream);
/* ddu = |grad u|^2 alone: (d_x^2 + d_y^2) + d_z^2                         */
int lc_nabla_u_square(int prec, int geocode, const int *n, const void *u, void *ddu, int nord,
                      const double *h, void *stream);
/* du(:, :, :, k) = eps D_k u / h_k                                           */
int lc_nabla_u_epsilon(int prec, int geocode, const int *n, const void *u, void *du, int nord,
                       const double *h, const void *eps, void *stream);
/* per point res = ((e_1 d_x u + e_2 d_y u) + e_3 d_z u) / (4 pi), with
 * (e_1, e_2, e_3) = dlogeps(:, i0, i1, i2) ((3, n0, n1, n2), the component
 * fastest); res = eta (res - rhopol); rhopol = res + rhopol (in place);
 * *rhores2_dev (device FP64; a host address is refused) = the sum of res^2
 * over the grid, summed in FP64 in a fixed order (the same bits on every
 * call)                                                                    */
int lc_update_rhopol(int prec, int geocode, const int *n, const void *u, int nord, const double *h,
                     double eta, const void *dlogeps, void *rhopol, double *rhores2_dev, void *stream);
/* The generated procedure names (C flavour, FP64, host or device arrays,
 * synchronous): Poisson2.rb:1951-2043 (Poisson_broker), :2067-2195 (div_u_i),
 * :2553-2620 (nabla_u), :2197-2276 (nabla_u_and_square), :2278-2359
 * (nabla_u_square), :2362-2443 (nabla_u_epsilon), :2446-2550 (update_rhopol;
 * rhores2 host or device).                                                  */
void Poisson_broker(int nord, int idim, const int *nn, int bc, const double *u, double *du, double a);
void div_u_i(int geocode, int n01, int n02, int n03, const double *u, double *du, int nord,
             const double *hgrids, double *cc);
void nabla_u(int geocode, int n01, int n02, int n03, const double *u, double *du, int nord,
             const double *hgrids);
void nabla_u_and_square(int geocode, int n01, int n02, int n03, const double *u, double *du, double *du2,
                        int nord, const double *hgrids);
void nabla_u_square(int geocode, int n01, int n02, int n03, const double *u, double *ddu, int nord,
                    const double *hgrids);
void nabla_u_epsilon(int geocode, int n01, int n02, int n03, const double *u, double *du, int nord,
                     const double *hgrids, const double *eps);
void update_rhopol(int geocode, int n01, int n02, int n03, const double *u, int nord, const double *hgrids,
                   double eta, const double *dlogeps, double *rhopol, double *rhores2);

/* extent of the input (side 0) or output (side 1) along a dim for an
 * operator: op 0..3 = fil ids, 10 + wav = analysis, 20 + wav = synthesis */
int lc_extent(int op, int bc, int n, int side);
/* library version string */
const char *lc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LIBCONV_H */

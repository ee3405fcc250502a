// lc_kernels.hpp -- launch interfaces of the libconv HIP kernels (internal).
#pragma once

#include <string>
#include <vector>

#include "lc_common.hpp"

namespace lc {

enum OpId : int { OP_MD = 0, OP_IMD = 1, OP_D2 = 2, OP_D1 = 3 };
enum WavId : int { WAV_SYM8 = 0, WAV_SYMICOMB8 = 1 };

// One 1D sweep over narr arrays seen as (L, line, R).
template <typename R>
struct SweepArgs {
  const R* x;
  R* y;
  long L;         // stride between steps of a line (elements)
  long nr;        // lines per array along the slower dims
  long ncols;     // L * R * narr
  long line_in;   // distance between consecutive r in x (= L * IN * nx)
  long line_out;  // distance between consecutive r in y (= L * OUT * ny)
  long nti, nto;  // distance between consecutive arrays
  int narr;
  int n, nx, ny;  // logical length and extents along the line (steps)
  int ox, oy;     // logical index of the first stored input / output step
  int i_first;    // first output step (-up for grow, else 0)
  int nout;       // number of output steps
  int chunk;      // outputs per lane chunk (strided kernel; set by launcher)
  R a, a_x, a_y;
  int flags;      // F_A | F_AX | F_AY | F_DOT
  double* dot_part;  // per-block FP64 partials of <x, a f(x)> (or null)
};

template <typename R>
hipError_t sweep(int kind, int id, int bc, const SweepArgs<R>& p, hipStream_t st, double* dot_out);
template <typename R>
long sweep_partials_needed(const SweepArgs<R>& p, int W);
hipError_t sum_partial_sets(const double* part, long nparts, double* out, int nsets, long stride,
                            hipStream_t st);

// Filter geometry for the host side.
struct OpGeom {
  int lo, up, w;  // step offsets and window (pairs for wavelets)
};
OpGeom conv_geom(int id);
OpGeom wavelet_geom(int id, int recompose, int bc);

// ------------------------------------------------------------ 3D operators
// Box of one 3D array: logical extent n[3], BC per dim, storage extents
// nx[3] (input) and ny[3] (output).
struct Box3 {
  int n[3], bc[3], nx[3], ny[3];
};

// the two-plane bodies read input planes through 32-bit buffer records:
// planes of at most kPlaneMax bytes
constexpr long kPlaneMax = 0x7fffffffL;
inline bool plane_fits(const Box3& b, int prec) {
  return (long)b.nx[0] * b.nx[1] * prec <= kPlaneMax && (long)b.ny[0] * b.ny[1] * prec <= kPlaneMax;
}
// wavelet arrays hold 2 x 2 values per (x, y) step
inline bool wav_plane_fits(const Box3& b, int prec) { return 4L * b.nx[0] * b.nx[1] * prec <= kPlaneMax; }

// separable product y = a * F2 F1 F0 x (+ a_y y), fused single pass
template <typename R>
hipError_t sep3d(int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                 double* scratch, size_t scratch_bytes);
// kinetic-style sum y = sum_d a_d F_d x + a_x x + a_y y with per-dim
// energies dot[d] = <x, a_d F_d x> (of the last array) when dot != null
// LIBCONV_DIRECT=1: the unfused 3D kernels (a debugging aid)
bool direct_kernels();

// Poisson nabla operators (lc_poisson.hip); n[3], nord in {2, 4, 6, 8, 16}
bool poisson_order_ok(int nord);
template <typename R>
hipError_t poisson_1d(int nord, int idim, const int* n, int bc, const R* u, R* du, R a, R a_y, hipStream_t st);
template <typename R>
hipError_t div_u_i(int geocode, const int* n, const R* u, R* du, R* cc, int nord, const double* h, hipStream_t st);
template <typename R>
hipError_t nabla_u(int geocode, const int* n, const R* u, R* du, R* du2, int nord, const double* h, hipStream_t st);
template <typename R>
hipError_t nabla_u_square(int geocode, const int* n, const R* u, R* ddu, int nord, const double* h, hipStream_t st);
template <typename R>
hipError_t nabla_u_epsilon(int geocode, const int* n, const R* u, R* du, const R* eps, int nord, const double* h,
                           hipStream_t st);
// update_rhopol writes one FP64 partial sum of res^2 per block to part
// (update_rhopol_partials of them, on the current device) and their sum, in
// a fixed order, to *rhores2 (device)
template <typename R>
long update_rhopol_partials(int geocode, const int* n, int nord);
template <typename R>
hipError_t update_rhopol(int geocode, const int* n, const R* u, int nord, const double* h, R eta, const R* dlogeps,
                         R* rhopol, double* part, long nparts, double* rhores2, hipStream_t st);

template <typename R>
hipError_t sum3d(int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                 double* dot, hipStream_t st, double* scratch, size_t scratch_bytes);
template <typename R>
size_t sum3d_scratch_bytes(const Box3& b, int narr);

// single-pass kernels (lc_fused.hip)
template <typename R>
hipError_t sum3d_fused_run(int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                           R a_y, double* dot, hipStream_t st, double* scratch, size_t scratch_bytes);
template <typename R>
// pot (optional): the F_POT step, rho += t^2 and t = pot t, on every output
hipError_t sep3d_fused_run(int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                           const R* pot = nullptr, R* rho = nullptr);
size_t fused_sum3d_parts(const Box3& b, int narr, int lo, int up, int w);

// wave-specialised single-pass kernels (lc_fused_ws.hip); per_cu = target
// blocks per CU, minz = shortest z chunk
template <typename R, class F, int TX, int TY, int QX, int QY, int HD>
hipError_t launch_sum3d_ws(const Box3& b, int narr, const R* x, R* y, const double* a, R a_x, R a_y,
                           double* dot, hipStream_t st, double* scratch, size_t sb, int per_cu,
                           int minz);
template <typename R, class F, int TX, int TY, int QX, int QY, int AD>
hipError_t launch_sep3d_ws(const Box3& b, int narr, const R* x, R* y, R a, R a_y, hipStream_t st,
                           int per_cu, int minz);
size_t ws_sum3d_parts_bound(const Box3& b, int lo, int up);
template <typename R>
size_t sep3d_scratch_bytes(const Box3& b, int narr);

// ------------------------------------------------------------ batches
// A stage applies one operator to `count` independent 3D items in one
// launch (lc_batch.cpp builds stages from the public batch plan).
enum BatchKind : int { BK_SEP = 0, BK_SUM = 1, BK_DWT = 2, BK_IDWT = 3 };
template <typename R>
struct BatchItem {
  Box3 b;
  const R* x;
  R* y;
  double a[3];       // BK_SUM: a_d per dim; BK_SEP / wavelets: a[0] = a
  double a_x, a_y;
  double* dot;       // BK_SUM: device pointer to 3 energies, or null
};
struct BatchStage {
  int kind = 0, id = 0, bc = 0, count = 0, blocks = 0;
  int variant = 0;           // kernel variant of the stage (tuner index)
  void* tab = nullptr;       // device table of per-item kernel arguments
  int* first = nullptr;      // device: first block of every item (count + 1)
  double* part = nullptr;    // device: BK_SUM energy partials, 3 per block
  double** dots = nullptr;   // device: BK_SUM energy destinations per item
  bool any_dot = false;
};
// copies the argument table and block prefix to new device buffers
hipError_t batch_upload(const void* tab, size_t tab_bytes, const int* first, int count, BatchStage* s);
void batch_stage_free(BatchStage* s);
// build (allocating, synchronous) or launch (stream-ordered, no host work)
template <typename R>
hipError_t fused_stage(bool build, int kind, int id, const BatchItem<R>* items, int count,
                       BatchStage* s, hipStream_t st);
template <typename R>
hipError_t wav_stage(bool build, int kind, int id, int bc, const BatchItem<R>* items, int count,
                     BatchStage* s, hipStream_t st);
// single-pass 3D wavelet transform (kind KDWT / KIDWT), lc_fused_wav.hip
template <typename R>
hipError_t sepw3d(int kind, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                  hipStream_t st);

// ------------------------------------------------------------ variants / tuner
// Forms of the 3D operators the tuner selects kernels for.
enum Form : int { FORM_SEP = 0, FORM_SUM = 1, FORM_DWT = 2, FORM_IDWT = 3 };
int variant_count(int form);          // single-call variants
bool has_batch_form(int form, int v);  // the variant also has a batched kernel
int wav_variant_count();
// variant a call of this shape uses: tuner table, else environment, else default
int select_variant(int prec, int form, int id, const Box3& b);
// tuner table lookup (lc_tune.cpp): -1 when the shape was not tuned
int tuned_variant(int prec, int form, int id, const Box3& b);
std::string fused_kernel_name(int prec, int form, int id, int v, bool batch);
std::string wav_kernel_name(int prec, int form, int id, int v, bool batch);
template <typename R>
hipError_t sum3d_fused_v(int v, int id, const Box3& b, int narr, const R* x, R* y, const double* a, R a_x,
                         R a_y, double* dot, hipStream_t st, double* scratch, size_t sb);
template <typename R>
hipError_t sep3d_fused_v(int v, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                         hipStream_t st, const R* pot = nullptr, R* rho = nullptr);
template <typename R>
hipError_t sepw3d_v(int v, int kind, int id, const Box3& b, int narr, const R* x, R* y, R a, R a_y,
                    hipStream_t st);

// tuner (lc_tune.cpp)
template <typename R>
hipError_t tune_md(int form, int id, const Box3& b, int repeat, int* best, double* best_ms,
                   std::vector<double>* times);
int tune_save(const char* path);
int tune_load(const char* path);
void tune_clear();

// one persistent launch over a batch (lc_batch_persist.hip); kinds[] are
// BatchKind values, ids[] the operators
struct PersistPlan;
template <typename R>
hipError_t persist_build(const BatchItem<R>* items, const int* kinds, const int* ids, int count,
                         PersistPlan** out);
hipError_t persist_launch(const PersistPlan* pp, hipStream_t st, int spin_limit);
hipError_t persist_status(const PersistPlan* pp, int* status);
int persist_grid(const PersistPlan* pp);
void persist_free(PersistPlan* pp);

}  // namespace lc

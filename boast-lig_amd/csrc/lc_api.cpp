// lc_api.cpp -- C ABI of the libconv HIP library (include/libconv.h).
//
// Dispatch mirrors the generated libconv procedures:
//   1D  GenericConvolutionOperator1d#procedure (BigDFT/GenericConvolution2.rb
//       :1560-1766): ndat_left/right from nx[] of the other dims (x2 per dim
//       for wavelets), narr arrays at strides nti / nto, a == 1 / a_x == 0 /
//       a_y == 0 select the variant without that term;
//   md  GenericConvolutionOperator#procedure (:2131-2245): sweep i reads the
//       array whose dims < i carry the output extents ny[] and dims >= i the
//       input extents nx[].
// Pointers are classified per call: device memory runs in place on the
// default stream, host memory is staged through a per-thread device pool.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/libconv.h"
#include "lc_kernels.hpp"

using namespace lc;

namespace {

thread_local int g_err = LC_OK;
thread_local char g_msg[512] = "";

int fail(int code, const char* fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  const char* ab = getenv("LIBCONV_ABORT_ON_ERROR");
  if (ab && ab[0] == '1') {
    fprintf(stderr, "libconv: %s\n", g_msg);
    abort();
  }
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(LC_ERR_DEVICE, "%s: %s", where, hipGetErrorString(e));
}

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// ---------------------------------------------------------------- pools
// Device scratch and staging buffers, owned per host thread and keyed by
// (device, stream): work queued asynchronously on one stream never shares a
// buffer with work on another stream, so stream-API calls on different
// streams of one thread may overlap.  Calls on one stream are ordered by the
// stream itself.
struct Pool {
  void* p = nullptr;
  size_t bytes = 0;
};
constexpr int kMaxDev = 16;
constexpr int kSlots = 6;  // 0 scratch, 1 x staging, 2 y staging, 3 dot, 4-5 extra outputs
struct StreamPools {
  int dev;
  hipStream_t st;
  Pool slot[kSlots];
};
thread_local std::vector<StreamPools> g_pools;

int pool_get(int slot, size_t bytes, void** out, hipStream_t st) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(LC_ERR_UNSUPPORTED, "device %d out of range", dev);
  StreamPools* sp = nullptr;
  for (auto& e : g_pools)
    if (e.dev == dev && e.st == st) { sp = &e; break; }
  if (!sp) {
    g_pools.push_back(StreamPools{dev, st, {}});
    sp = &g_pools.back();
  }
  Pool& pl = sp->slot[slot];
  if (bytes == 0) bytes = 256;
  if (pl.bytes < bytes) {
    if (pl.p) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(pl.p));
      pl.p = nullptr;
      pl.bytes = 0;
    }
    size_t want = std::max(bytes, pl.bytes + pl.bytes / 2);
    hipError_t e = hipMalloc(&pl.p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      e = hipMalloc(&pl.p, bytes);
      if (e != hipSuccess) return fail(LC_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
      want = bytes;
    }
    pl.bytes = want;
  }
  *out = pl.p;
  return LC_OK;
}

// ------------------------------------------------ static scratch
// Device scratch of lc_conv_1ds_noalloc: a file-scope device array, one copy
// per GPU, that the library never allocates or frees.  In host code the
// array's name is only a stand-in for the symbol: its device address comes
// from hipGetSymbolAddress, taken once per device (with that device current)
// on the entry point's first call there, and only that address is handed to
// kernels.  Every call on one GPU shares the array.
constexpr size_t kStaticScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
void* g_static_scratch[kMaxDev] = {};
std::mutex g_static_mu;

int static_scratch(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(LC_ERR_UNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (!g_static_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(LC_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
    g_static_scratch[dev] = p;
  }
  *out = g_static_scratch[dev];
  return LC_OK;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// Host-or-device array argument of a drop-in call.
template <typename R>
struct Arg {
  const R* host_in = nullptr;
  R* host_out = nullptr;
  R* dev = nullptr;
  size_t count = 0;
};

template <typename R>
int stage_in(const R* p, size_t count, int slot, bool copy, hipStream_t st, Arg<R>* a) {
  a->count = count;
  if (is_device_ptr(p)) {
    a->dev = const_cast<R*>(p);
    return LC_OK;
  }
  void* d;
  int rc = pool_get(slot, count * sizeof(R), &d, st);
  if (rc) return rc;
  a->dev = static_cast<R*>(d);
  a->host_in = p;
  if (copy && count) HIPCHK(hipMemcpyAsync(d, p, count * sizeof(R), hipMemcpyHostToDevice, st));
  return LC_OK;
}

template <typename R>
int stage_out(Arg<R>& a, R* host, hipStream_t st) {
  if (a.host_in || host) {
    if (a.count && !is_device_ptr(host))
      HIPCHK(hipMemcpyAsync(host, a.dev, a.count * sizeof(R), hipMemcpyDeviceToHost, st));
  }
  return LC_OK;
}

// ---------------------------------------------------------------- geometry
bool valid_bc(int bc) { return bc == 0 || bc == 1 || bc == -1; }

int conv_window(int fil) {
  switch (fil) {
    case OP_MD: case OP_IMD: return 16;
    case OP_D2: case OP_D1: return 29;
  }
  return -1;
}

int wav_window(int wav) {
  switch (wav) {
    case WAV_SYM8: return 8;
    case WAV_SYMICOMB8: return 16;
  }
  return -1;
}

// required extents of the processed dim for window W (steps)
// free_ok: the caller runs on the single-pass 3D convolution kernels, which
// also take bc 2 (free, output on the input box)
int check_line(int bc, int n, int nx, int ny, int W, const char* who, bool free_ok = false) {
  if (!valid_bc(bc) && !(free_ok && bc == FREE))
    return fail(LC_ERR_ARG, "%s: bc %d not in {0, 1, -1%s}", who, bc, free_ok ? ", 2" : "");
  if (n < 1) return fail(LC_ERR_ARG, "%s: n = %d < 1", who, n);
  const int need_in = bc == -1 ? n + W - 1 : n;
  const int need_out = bc == 1 ? n + W - 1 : n;
  if (nx < need_in) return fail(LC_ERR_ARG, "%s: nx = %d < %d required by bc %d", who, nx, need_in, bc);
  if (ny < need_out) return fail(LC_ERR_ARG, "%s: ny = %d < %d required by bc %d", who, ny, need_out, bc);
  return LC_OK;
}

int check_dims(int d, const int* n, const int* nx, const int* ny, int narr, const char* who) {
  if (d < 1 || d > 16) return fail(LC_ERR_ARG, "%s: d = %d not in [1, 16]", who, d);
  if (!n || !nx || !ny) return fail(LC_ERR_ARG, "%s: null size array", who);
  if (narr < 0) return fail(LC_ERR_ARG, "%s: narr = %d < 0", who, narr);
  for (int i = 0; i < d; ++i)
    if (n[i] < 1 || nx[i] < 1 || ny[i] < 1)
      return fail(LC_ERR_ARG, "%s: dim %d has n/nx/ny = %d/%d/%d", who, i, n[i], nx[i], ny[i]);
  return LC_OK;
}

// ---------------------------------------------------------------- 1D core
// kind: KCONV (id = filter), KDWT / KIDWT (id = wavelet)
template <typename R>
int run_1d(int kind, int id, int d, int idim, const int* n, int bc, const int* nx, const int* ny,
           int narr, const R* x, R* y, double a, double a_x, double a_y, double* dot_host,
           double* dot_dev_user, hipStream_t st, bool device_api) {
  const char* who = kind == KCONV ? "s0s0_1d" : (kind == KDWT ? "s0s1_1d" : "s1s0_1d");
  int rc = check_dims(d, n, nx, ny, narr, who);
  if (rc) return rc;
  if (idim < 0 || idim >= d) return fail(LC_ERR_ARG, "%s: idim = %d not in [0, %d)", who, idim, d);
  const int W = kind == KCONV ? conv_window(id) : wav_window(id);
  if (W < 0) return fail(LC_ERR_ARG, "%s: unknown operator id %d", who, id);
  rc = check_line(bc, n[idim], nx[idim], ny[idim], W, who);
  if (rc) return rc;
  const int mul = kind == KCONV ? 1 : 2;
  long L = 1, Rr = 1;
  for (int i = 0; i < idim; ++i) L *= (long)mul * nx[i];
  for (int i = idim + 1; i < d; ++i) Rr *= (long)mul * nx[i];
  OpGeom g = kind == KCONV ? conv_geom(id) : wavelet_geom(id, kind == KIDWT, bc);
  SweepArgs<R> p{};
  p.L = L;
  p.nr = Rr;
  p.narr = narr;
  p.ncols = L * Rr * narr;
  p.n = n[idim];
  p.nx = nx[idim];
  p.ny = ny[idim];
  p.line_in = L * mul * nx[idim];
  p.line_out = L * mul * ny[idim];
  p.nti = p.line_in * Rr;
  p.nto = p.line_out * Rr;
  p.ox = bc == -1 ? g.lo : 0;
  p.oy = bc == 1 ? -g.up : 0;
  p.i_first = bc == 1 ? -g.up : 0;
  p.nout = bc == 1 ? n[idim] + W - 1 : n[idim];
  p.a = (R)a;
  p.a_x = (R)a_x;
  p.a_y = (R)a_y;
  p.flags = (a != 1.0 ? F_A : 0) | (kind == KCONV && a_x != 0.0 ? F_AX : 0) | (a_y != 0.0 ? F_AY : 0);
  const bool want_dot = kind == KCONV && (dot_host || dot_dev_user);
  if (want_dot) p.flags |= F_DOT;

  Arg<R> ax, ay;
  const size_t nin = (size_t)p.nti * narr, nout = (size_t)p.nto * narr;
  if (device_api) {
    ax.dev = const_cast<R*>(x);
    ay.dev = y;
  } else {
    rc = stage_in(x, nin, 1, true, st, &ax);
    if (rc) return rc;
    rc = stage_in<R>(y, nout, 2, true, st, &ay);  // keeps padding of y intact
    if (rc) return rc;
  }
  p.x = ax.dev;
  p.y = ay.dev;
  double* dot_dev = dot_dev_user;
  if (want_dot) {
    const long np = sweep_partials_needed(p, W);
    void* sc;
    rc = pool_get(0, sizeof(double) * (size_t)(np + 8), &sc, st);
    if (rc) return rc;
    p.dot_part = static_cast<double*>(sc);
    if (!dot_dev) dot_dev = static_cast<double*>(sc) + np;
  }
  hipError_t e = sweep<R>(kind, id, bc, p, st, want_dot ? dot_dev : nullptr);
  if (e != hipSuccess) return hip_fail(e, who);
  if (!device_api) {
    if (ay.host_in) HIPCHK(hipMemcpyAsync(y, ay.dev, nout * sizeof(R), hipMemcpyDeviceToHost, st));
    if (dot_host) HIPCHK(hipMemcpyAsync(dot_host, dot_dev, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return LC_OK;
}

// ---------------------------------------------------------------- md core
__global__ void axpy_kernel_d(long n, double a, const double* x, double* y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fmar(x[i], a, y[i]);
}
__global__ void axpy_kernel_f(long n, float a, const float* x, float* y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fmar(x[i], a, y[i]);
}
hipError_t axpy(long n, double a, const double* x, double* y, hipStream_t st) {
  if (n > 0) axpy_kernel_d<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, x, y);
  return hipGetLastError();
}
hipError_t axpy(long n, float a, const float* x, float* y, hipStream_t st) {
  if (n > 0) axpy_kernel_f<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, x, y);
  return hipGetLastError();
}

// kind: KCONV separable product / KDWT / KIDWT over all d dims
template <typename R>
int run_md(int kind, int id, int d, const int* n, const int* bc, const int* nx, const int* ny,
           int narr, const R* x, R* y, double a, double a_x, double a_y, hipStream_t st,
           bool device_api) {
  const char* who = kind == KCONV ? "s0s0_md" : (kind == KDWT ? "s0s1_md" : "s1s0_md");
  int rc = check_dims(d, n, nx, ny, narr, who);
  if (rc) return rc;
  if (!bc) return fail(LC_ERR_ARG, "%s: null bc", who);
  const int W = kind == KCONV ? conv_window(id) : wav_window(id);
  if (W < 0) return fail(LC_ERR_ARG, "%s: unknown operator id %d", who, id);
  const bool free_ok = kind == KCONV && d == 3 && !direct_kernels();
  for (int i = 0; i < d; ++i) {
    rc = check_line(bc[i], n[i], nx[i], ny[i], W, who, free_ok);
    if (rc) return rc;
  }
  const int mul = kind == KCONV ? 1 : 2;
  long tin = 1, tout = 1;
  for (int i = 0; i < d; ++i) {
    tin *= (long)mul * nx[i];
    tout *= (long)mul * ny[i];
  }
  if (a_x != 0.0 && tin != tout)
    return fail(LC_ERR_ARG, "%s: a_x needs input and output of the same shape", who);
  Arg<R> ax, ay;
  if (device_api) {
    ax.dev = const_cast<R*>(x);
    ay.dev = y;
  } else {
    rc = stage_in(x, (size_t)tin * narr, 1, true, st, &ax);
    if (rc) return rc;
    rc = stage_in<R>(y, (size_t)tout * narr, 2, true, st, &ay);
    if (rc) return rc;
  }
  if (d == 3 && kind == KCONV) {  // single-pass 3D product
    Box3 b;
    for (int i = 0; i < 3; ++i) { b.n[i] = n[i]; b.bc[i] = bc[i]; b.nx[i] = nx[i]; b.ny[i] = ny[i]; }
    const size_t sb = sep3d_scratch_bytes<R>(b, narr);
    void* sc = nullptr;
    if (getenv("LIBCONV_DIRECT")) {
      rc = pool_get(0, sb, &sc, st);
      if (rc) return rc;
    }
    hipError_t e = sep3d<R>(id, b, narr, ax.dev, ay.dev, (R)a, (R)a_y, st, static_cast<double*>(sc), sb);
    if (e != hipSuccess) return hip_fail(e, who);
    if (a_x != 0.0) {
      e = axpy((long)tout * narr, (R)a_x, ax.dev, ay.dev, st);
      if (e != hipSuccess) return hip_fail(e, who);
    }
    if (!device_api) {
      if (ay.host_in) HIPCHK(hipMemcpyAsync(y, ay.dev, (size_t)tout * narr * sizeof(R), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    return LC_OK;
  }
  // single-pass 3D sym8 wavelet transform (symicomb8 runs as three sweeps)
  Box3 wb;
  for (int i = 0; i < 3 && d == 3; ++i) { wb.n[i] = n[i]; wb.bc[i] = bc[i]; wb.nx[i] = nx[i]; wb.ny[i] = ny[i]; }
  if (d == 3 && kind != KCONV && id == WAV_SYM8 && wav_plane_fits(wb, sizeof(R)) && !getenv("LIBCONV_DIRECT")) {
    const Box3& b = wb;
    hipError_t e = sepw3d<R>(kind, id, b, narr, ax.dev, ay.dev, (R)a, (R)a_y, st);
    if (e != hipSuccess) return hip_fail(e, who);
    if (a_x != 0.0) {
      e = axpy((long)tout * narr, (R)a_x, ax.dev, ay.dev, st);
      if (e != hipSuccess) return hip_fail(e, who);
    }
    if (!device_api) {
      if (ay.host_in) HIPCHK(hipMemcpyAsync(y, ay.dev, (size_t)tout * narr * sizeof(R), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    return LC_OK;
  }
  // intermediates: after sweep s the dims < s+1 carry ny
  int cur[16];
  for (int i = 0; i < d; ++i) cur[i] = nx[i];
  size_t wmax = 0;
  for (int s = 0; s < d - 1; ++s) {
    cur[s] = ny[s];
    size_t t = 1;
    for (int i = 0; i < d; ++i) t *= (size_t)mul * cur[i];
    wmax = std::max(wmax, t);
  }
  R* w[2] = {nullptr, nullptr};
  if (d > 1) {
    void* sc;
    rc = pool_get(0, 2 * wmax * (size_t)std::max(narr, 1) * sizeof(R), &sc, st);
    if (rc) return rc;
    w[0] = static_cast<R*>(sc);
    w[1] = w[0] + wmax * (size_t)std::max(narr, 1);
  }
  for (int i = 0; i < d; ++i) cur[i] = nx[i];
  const R* src = ax.dev;
  for (int s = 0; s < d; ++s) {
    const bool last = s == d - 1;
    R* dst = last ? ay.dev : w[s & 1];
    OpGeom g = kind == KCONV ? conv_geom(id) : wavelet_geom(id, kind == KIDWT, bc[s]);
    SweepArgs<R> p{};
    long L = 1, Rr = 1;
    for (int i = 0; i < s; ++i) L *= (long)mul * cur[i];
    for (int i = s + 1; i < d; ++i) Rr *= (long)mul * cur[i];
    p.x = src;
    p.y = dst;
    p.L = L;
    p.nr = Rr;
    p.narr = narr;
    p.ncols = L * Rr * narr;
    p.n = n[s];
    p.nx = cur[s];
    p.ny = ny[s];
    p.line_in = L * mul * cur[s];
    p.line_out = L * mul * ny[s];
    p.nti = p.line_in * Rr;
    p.nto = p.line_out * Rr;
    p.ox = bc[s] == -1 ? g.lo : 0;
    p.oy = bc[s] == 1 ? -g.up : 0;
    p.i_first = bc[s] == 1 ? -g.up : 0;
    p.nout = bc[s] == 1 ? n[s] + W - 1 : n[s];
    p.a = last ? (R)a : R(1);
    p.a_y = (R)a_y;
    p.flags = (last && a != 1.0 ? F_A : 0) | (last && a_y != 0.0 ? F_AY : 0);
    hipError_t e = sweep<R>(kind, id, bc[s], p, st, nullptr);
    if (e != hipSuccess) return hip_fail(e, who);
    cur[s] = ny[s];
    src = dst;
  }
  if (a_x != 0.0) {
    hipError_t e = axpy((long)tout * narr, (R)a_x, ax.dev, ay.dev, st);
    if (e != hipSuccess) return hip_fail(e, who);
  }
  if (!device_api) {
    if (ay.host_in) HIPCHK(hipMemcpyAsync(y, ay.dev, (size_t)tout * narr * sizeof(R), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return LC_OK;
}

// kinetic-style sum over d <= 3 dims
// use_static: the energy partials go to the library's static device array
// instead of the pool (the allocation-free entry point); shapes whose
// partials do not fit are refused before the device is touched
template <typename R>
int run_1ds(int fil, int d, const int* n, const int* bc, const int* nx, const int* ny, int narr,
            const double* a, double a_x, double a_y, const R* x, R* y, double* dot_host,
            double* dot_dev_user, hipStream_t st, bool device_api, bool use_static = false) {
  const char* who = "s0s0_1ds";
  int rc = check_dims(d, n, nx, ny, narr, who);
  if (rc) return rc;
  if (d > 3) return fail(LC_ERR_UNSUPPORTED, "%s: d = %d > 3 not supported on the GPU", who, d);
  if (!bc || !a) return fail(LC_ERR_ARG, "%s: null bc or a", who);
  const int W = conv_window(fil);
  if (W < 0) return fail(LC_ERR_ARG, "%s: unknown filter %d", who, fil);
  Box3 b;
  double a3[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    if (i < d) {
      rc = check_line(bc[i], n[i], nx[i], ny[i], W, who, !direct_kernels());
      if (rc) return rc;
      b.n[i] = n[i]; b.bc[i] = bc[i]; b.nx[i] = nx[i]; b.ny[i] = ny[i];
      a3[i] = a[i];
    } else {
      b.n[i] = 1; b.bc[i] = 0; b.nx[i] = 1; b.ny[i] = 1;
    }
  }
  long tin = (long)b.nx[0] * b.nx[1] * b.nx[2], tout = (long)b.ny[0] * b.ny[1] * b.ny[2];
  const bool want_dot = dot_host || dot_dev_user;
  const size_t sbytes = want_dot ? sum3d_scratch_bytes<R>(b, narr) : 0;
  void* fixed = nullptr;
  if (use_static) {
    if (want_dot && sbytes + 8 * sizeof(double) > kStaticScratchBytes)
      return fail(LC_ERR_UNSUPPORTED,
                  "%s: the energy partials of this shape need %zu bytes of scratch, above the "
                  "%zu-byte static buffer of the allocation-free entry point",
                  who, sbytes + 8 * sizeof(double), kStaticScratchBytes);
    rc = static_scratch(&fixed);
    if (rc) return rc;
  }
  Arg<R> ax, ay;
  if (device_api) {
    ax.dev = const_cast<R*>(x);
    ay.dev = y;
  } else {
    rc = stage_in(x, (size_t)tin * narr, 1, true, st, &ax);
    if (rc) return rc;
    rc = stage_in<R>(y, (size_t)tout * narr, 2, true, st, &ay);
    if (rc) return rc;
  }
  double* dot3 = nullptr;
  double* scratch = nullptr;
  if (want_dot) {
    void* sc = fixed;
    if (!sc) {
      rc = pool_get(0, sbytes + 8 * sizeof(double), &sc, st);
      if (rc) return rc;
    }
    scratch = static_cast<double*>(sc);
    dot3 = reinterpret_cast<double*>(static_cast<char*>(sc) + sbytes);
  }
  hipError_t e = sum3d<R>(fil, b, narr, ax.dev, ay.dev, a3, (R)a_x, (R)a_y, want_dot ? dot3 : nullptr,
                          st, scratch, sbytes);
  if (e != hipSuccess) return hip_fail(e, who);
  if (want_dot && dot_dev_user)
    HIPCHK(hipMemcpyAsync(dot_dev_user, dot3, d * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (!device_api) {
    if (ay.host_in) HIPCHK(hipMemcpyAsync(y, ay.dev, (size_t)tout * narr * sizeof(R), hipMemcpyDeviceToHost, st));
    if (dot_host) HIPCHK(hipMemcpyAsync(dot_host, dot3, d * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return LC_OK;
}

void kin_extents(int d, const int* n, const int* bc, int* nx, int* ny) {
  for (int i = 0; i < d && i < 16; ++i) {
    nx[i] = bc[i] == -1 ? n[i] + 28 : n[i];
    ny[i] = bc[i] == 1 ? n[i] + 28 : n[i];
  }
}

void wav_extents(int d, const int* n, const int* bc, int recompose, int* nx, int* ny) {
  (void)recompose;
  for (int i = 0; i < d && i < 16; ++i) {
    nx[i] = bc[i] == -1 ? n[i] + 7 : n[i];
    ny[i] = bc[i] == 1 ? n[i] + 7 : n[i];
  }
}

hipStream_t hs(void* s) { return static_cast<hipStream_t>(s); }

long cost_1d(int d, int idim, const int* n, int narr, int per_step, int mul) {
  long ndat = 1;
  for (int i = 0; i < d; ++i)
    if (i != idim) ndat *= (long)mul * n[i];
  return (long)narr * n[idim] * ndat * per_step;
}

// ---------------------------------------------------------------- local Hamiltonian
// rho += psir^2, then psir = pot psir (the step between the magic filter and
// its inverse in libconv.f90:266-271), one point per lane; the unfused form
// of the F_POT store step (used with LIBCONV_DIRECT=1)
template <typename R>
__global__ __launch_bounds__(256) void pot_density(long n, R* __restrict__ psir, const R* __restrict__ pot,
                                                   R* __restrict__ rho) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const R t = psir[i];
    if (rho) rho[i] = fmar(t, t, rho[i]);
    psir[i] = pot[i] * t;
  }
}

struct HamDims {
  int n[3];   // coarse grid (pairs)
  int f[3];   // psi_fscf extent 2 (n + 7): the synthesis grown by 7 pairs
  int r[3];   // psir extent 2 (n + 7) + 15: the magic filter grown by 15
  long nf, nr, ng;  // points of psi_fscf, psir, psig
};

int ham_dims(const int* n, HamDims* h, const char* who) {
  if (!n) return fail(LC_ERR_ARG, "%s: null n", who);
  h->nf = h->nr = h->ng = 1;
  for (int i = 0; i < 3; ++i) {
    if (n[i] < 1) return fail(LC_ERR_ARG, "%s: n[%d] = %d < 1", who, i, n[i]);
    h->n[i] = n[i];
    h->f[i] = 2 * (n[i] + 7);
    h->r[i] = h->f[i] + 15;
    h->nf *= h->f[i];
    h->nr *= h->r[i];
    h->ng *= 2L * n[i];
  }
  return LC_OK;
}

// libconv.f90:238-295, device data on one stream:
//   psi_fscf = IDWT_grow psig            (s1s0_md sym8, bc 1)
//   psir     = MF_grow psi_fscf          (s0s0_md sym8_md, bc 1)
//   rho += psir^2; psir = pot psir
//   hpsi     = IMF_shrink psir           (s0s0_md sym8_mi, bc -1)
//   hpsi    += sum_d a_d D2_d psi_fscf   (s0s0_1ds_dot sym8_d2, bc 2 "free, nogrow", a_y = 1),
//              ekin[d] = <psi_fscf, a_d D2_d psi_fscf>
//   hpsig    = DWT_shrink hpsi           (s0s1_md sym8, bc -1)
template <typename R>
int run_local_hamiltonian(const int* n, const double* a, const R* psig, R* hpsig, const R* pot, R* rho,
                          R* work, double* ekin_dev, hipStream_t st) {
  const char* who = "lc_local_hamiltonian";
  HamDims h;
  int rc = ham_dims(n, &h, who);
  if (rc) return rc;
  if (!a || !psig || !hpsig || !pot || !work) return fail(LC_ERR_ARG, "%s: null argument", who);
  R* fscf = work;
  R* psir = work + h.nf;
  R* hpsi = psir + h.nr;
  const int g3[3] = {1, 1, 1}, s3[3] = {-1, -1, -1}, f3[3] = {FREE, FREE, FREE};
  const int nf7[3] = {n[0] + 7, n[1] + 7, n[2] + 7};
  rc = run_md<R>(KIDWT, WAV_SYM8, 3, h.n, g3, h.n, nf7, 1, psig, fscf, 1.0, 0.0, 0.0, st, true);
  if (rc) return rc;
  if (!direct_kernels()) {
    // MF grow with the potential step fused into its store (no extra pass over psir)
    Box3 b;
    for (int i = 0; i < 3; ++i) { b.n[i] = h.f[i]; b.bc[i] = 1; b.nx[i] = h.f[i]; b.ny[i] = h.r[i]; }
    hipError_t e = sep3d_fused_run<R>(OP_MD, b, 1, fscf, psir, R(1), R(0), st, pot, rho);
    if (e != hipSuccess) return hip_fail(e, who);
  } else {
    rc = run_md<R>(KCONV, OP_MD, 3, h.f, g3, h.f, h.r, 1, fscf, psir, 1.0, 0.0, 0.0, st, true);
    if (rc) return rc;
    const unsigned nb = (unsigned)std::min<long>((h.nr + 255) / 256, 8L * 1024);
    pot_density<R><<<nb, 256, 0, st>>>(h.nr, psir, pot, rho);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, who);
  }
  rc = run_md<R>(KCONV, OP_IMD, 3, h.f, s3, h.r, h.f, 1, psir, hpsi, 1.0, 0.0, 0.0, st, true);
  if (rc) return rc;
  rc = run_1ds<R>(OP_D2, 3, h.f, f3, h.f, h.f, 1, a, 0.0, 1.0, fscf, hpsi, nullptr, ekin_dev, st, true);
  if (rc) return rc;
  return run_md<R>(KDWT, WAV_SYM8, 3, h.n, s3, nf7, h.n, 1, hpsi, hpsig, 1.0, 0.0, 0.0, st, true);
}

// ---------------------------------------------------------------- Poisson
// validation shared by the Poisson entry points (div_u_i_ref: the grid must
// exceed the stencil along every dim; orders 2, 4, 6, 8, 16)
int poisson_check(int nord, const int* n, const char* who) {
  if (!poisson_order_ok(nord)) return fail(LC_ERR_ARG, "%s: nord %d not in {2, 4, 6, 8, 16}", who, nord);
  if (!n) return fail(LC_ERR_ARG, "%s: null n", who);
  for (int i = 0; i < 3; ++i)
    if (n[i] < nord + 1) return fail(LC_ERR_ARG, "%s: n[%d] = %d < nord + 1 = %d", who, i, n[i], nord + 1);
  return LC_OK;
}

enum PoissonOp { PO_1D, PO_DIV, PO_NABLA, PO_NABLA2 };

// one Poisson operator; arrays may be host (staged, synchronous) or device
// (device_api: in place, asynchronous).  in: u (3 n for PO_DIV); out0: du
// (3 n for the nabla forms); out1: cc (PO_DIV, optional) or du2 (PO_NABLA2)
template <typename R>
int run_poisson(int op, int nord, int idim, int bc, int geocode, const int* n, const R* u, R* out0, R* out1,
                double a, double a_y, const double* h, hipStream_t st, bool device_api) {
  const char* who = op == PO_1D ? "Poisson_broker" : op == PO_DIV ? "div_u_i" : op == PO_NABLA ? "nabla_u"
                                                                                            : "nabla_u_and_square";
  int rc = poisson_check(nord, n, who);
  if (rc) return rc;
  if (op == PO_1D && (idim < 0 || idim > 2)) return fail(LC_ERR_ARG, "%s: idim %d not in [0, 2]", who, idim);
  if (op == PO_1D && bc != 0 && bc != -2) return fail(LC_ERR_ARG, "%s: bc %d not in {0, -2}", who, bc);
  if (op != PO_1D && !h) return fail(LC_ERR_ARG, "%s: null hgrids", who);
  if (!u || !out0 || (op == PO_NABLA2 && !out1)) return fail(LC_ERR_ARG, "%s: null array", who);
  const size_t np = (size_t)n[0] * n[1] * n[2];
  Arg<R> au, a0, a1;
  if (device_api) {
    au.dev = const_cast<R*>(u);
    a0.dev = out0;
    a1.dev = out1;
  } else {
    rc = stage_in(u, (op == PO_DIV ? 3 : 1) * np, 1, true, st, &au);
    if (rc) return rc;
    rc = stage_in<R>(out0, (op == PO_NABLA || op == PO_NABLA2 ? 3 : 1) * np, 2, op == PO_1D && a_y != 0.0, st,
                     &a0);
    if (rc) return rc;
    if (out1) {
      rc = stage_in<R>(out1, np, 4, false, st, &a1);
      if (rc) return rc;
    }
  }
  hipError_t e;
  if (op == PO_1D)
    e = lc::poisson_1d<R>(nord, idim, n, bc, au.dev, a0.dev, (R)a, (R)a_y, st);
  else if (op == PO_DIV)
    e = lc::div_u_i<R>(geocode, n, au.dev, a0.dev, a1.dev, nord, h, st);
  else
    e = lc::nabla_u<R>(geocode, n, au.dev, a0.dev, op == PO_NABLA2 ? a1.dev : nullptr, nord, h, st);
  if (e != hipSuccess) return hip_fail(e, who);
  if (!device_api) {
    if (a0.host_in) HIPCHK(hipMemcpyAsync(out0, a0.dev, a0.count * sizeof(R), hipMemcpyDeviceToHost, st));
    if (out1 && a1.host_in) HIPCHK(hipMemcpyAsync(out1, a1.dev, a1.count * sizeof(R), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return LC_OK;
}

enum DielOp { PD_SQUARE, PD_EPS, PD_RHOPOL };

// nabla_u_square (out: ddu), nabla_u_epsilon (in1: eps; out: du, 3 n) and
// update_rhopol (in1: dlogeps, 3 n; out: rhopol, in and out; rhores2); arrays
// host (staged, synchronous) or device (device_api: asynchronous, rhores2 a
// device FP64 value)
template <typename R>
int run_dielectric(int op, int geocode, const int* n, const R* u, const R* in1, R* out, int nord, const double* h,
                   double eta, double* rhores2, hipStream_t st, bool device_api) {
  const char* who = op == PD_SQUARE ? "nabla_u_square" : op == PD_EPS ? "nabla_u_epsilon" : "update_rhopol";
  int rc = poisson_check(nord, n, who);
  if (rc) return rc;
  if (!h) return fail(LC_ERR_ARG, "%s: null hgrids", who);
  if (!u || !out || (op != PD_SQUARE && !in1) || (op == PD_RHOPOL && !rhores2))
    return fail(LC_ERR_ARG, "%s: null array", who);
  // the stream form writes rhores2 from the device: a host address there
  // would fault the kernel, so it is refused here
  if (device_api && op == PD_RHOPOL && !is_device_ptr(rhores2))
    return fail(LC_ERR_ARG, "%s: rhores2 is not device memory", who);
  const size_t np = (size_t)n[0] * n[1] * n[2];
  Arg<R> au, a1, ao;
  double* r2 = rhores2;
  if (device_api) {
    au.dev = const_cast<R*>(u);
    a1.dev = const_cast<R*>(in1);
    ao.dev = out;
  } else {
    rc = stage_in(u, np, 1, true, st, &au);
    if (!rc && op != PD_SQUARE) rc = stage_in(in1, (op == PD_EPS ? 1 : 3) * np, 4, true, st, &a1);
    if (!rc) rc = stage_in<R>(out, (op == PD_EPS ? 3 : 1) * np, 2, op == PD_RHOPOL, st, &ao);
    if (!rc && op == PD_RHOPOL && !is_device_ptr(rhores2)) {
      void* d;
      rc = pool_get(3, sizeof(double), &d, st);
      r2 = static_cast<double*>(d);
    }
    if (rc) return rc;
  }
  hipError_t e;
  if (op == PD_SQUARE) {
    e = lc::nabla_u_square<R>(geocode, n, au.dev, ao.dev, nord, h, st);
  } else if (op == PD_EPS) {
    e = lc::nabla_u_epsilon<R>(geocode, n, au.dev, ao.dev, a1.dev, nord, h, st);
  } else {
    const long nparts = lc::update_rhopol_partials<R>(geocode, n, nord);
    if (nparts < 0) return fail(LC_ERR_ARG, "%s: nord %d not in {2, 4, 6, 8, 16}", who, nord);
    void* part;
    rc = pool_get(0, (size_t)std::max(nparts, 1L) * sizeof(double), &part, st);
    if (rc) return rc;
    e = lc::update_rhopol<R>(geocode, n, au.dev, nord, h, (R)eta, a1.dev, ao.dev, static_cast<double*>(part),
                             nparts, r2, st);
  }
  if (e != hipSuccess) return hip_fail(e, who);
  if (!device_api) {
    if (ao.host_in) HIPCHK(hipMemcpyAsync(out, ao.dev, ao.count * sizeof(R), hipMemcpyDeviceToHost, st));
    if (r2 != rhores2) HIPCHK(hipMemcpyAsync(rhores2, r2, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return LC_OK;
}

}  // namespace

namespace lc {
OpGeom conv_geom(int id) {
  switch (id) {
    case OP_MD: return {FMD::LO, FMD::UP, FMD::N};
    case OP_IMD: return {FIMD::LO, FIMD::UP, FIMD::N};
    case OP_D2: return {FD2::LO, FD2::UP, FD2::N};
    case OP_D1: return {FD1::LO, FD1::UP, FD1::N};
  }
  return {0, 0, -1};
}
OpGeom wavelet_geom(int id, int recompose, int bc) {
  if (id == WAV_SYM8)
    return recompose ? OpGeom{Sym8Rec::LO, Sym8Rec::UP, Sym8Rec::W} : OpGeom{Sym8Dec::LO, Sym8Dec::UP, Sym8Dec::W};
  if (bc == 0)
    return recompose ? OpGeom{CombPRec::LO, CombPRec::UP, CombPRec::W} : OpGeom{CombPDec::LO, CombPDec::UP, CombPDec::W};
  return recompose ? OpGeom{CombGRec::LO, CombGRec::UP, CombGRec::W} : OpGeom{CombGDec::LO, CombGDec::UP, CombGDec::W};
}
}  // namespace lc

// ====================================================================== C ABI
extern "C" {

int lc_last_error(void) {
  int e = g_err;
  g_err = LC_OK;
  return e;
}
const char* lc_last_error_message(void) { return g_msg; }
const char* lc_version(void) { return "libconv-mi355x 0.1 (gfx950)"; }

int lc_extent(int op, int bc, int n, int side) {
  int W;
  if (op >= 0 && op < 4) W = conv_window(op);
  else if (op >= 10 && op < 12) W = wav_window(op - 10);
  else if (op >= 20 && op < 22) W = wav_window(op - 20);
  else return -1;
  if (!(valid_bc(bc) || (bc == FREE && op < 4)) || n < 1) return -1;
  if (side == 0) return bc == -1 ? n + W - 1 : n;
  return bc == 1 ? n + W - 1 : n;
}

// ---------------------------------------------------------------- 1D C names
#define LC_CONV1D(NAME, R, FIL)                                                                  \
  void NAME(int d, int idim, const int* n, int bc, const int* nx, const int* ny, int narr,       \
            const R* x, R* y, R a, R a_x, R a_y) {                                               \
    run_1d<R>(KCONV, FIL, d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, nullptr, nullptr,     \
              nullptr, false);                                                                   \
  }
LC_CONV1D(d_s0s0_1d_sym8_md, double, OP_MD)
LC_CONV1D(s_s0s0_1d_sym8_md, float, OP_MD)
LC_CONV1D(d_s0s0_1d_sym8_imd, double, OP_IMD)
LC_CONV1D(s_s0s0_1d_sym8_imd, float, OP_IMD)
LC_CONV1D(d_s0s0_1d_sym8_d2, double, OP_D2)
LC_CONV1D(s_s0s0_1d_sym8_d2, float, OP_D2)
LC_CONV1D(d_s0s0_1d_sym8_d1, double, OP_D1)
LC_CONV1D(s_s0s0_1d_sym8_d1, float, OP_D1)
#undef LC_CONV1D

void d_s0s0_1d_sym8_d2_dotin(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                             int narr, const double* x, double* y, double a, double a_x, double a_y,
                             double* dot_in) {
  double dot = 0.0;
  if (run_1d<double>(KCONV, OP_D2, d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, &dot, nullptr,
                     nullptr, false) == LC_OK && dot_in)
    *dot_in = dot;
}
void s_s0s0_1d_sym8_d2_dotin(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                             int narr, const float* x, float* y, float a, float a_x, float a_y,
                             float* dot_in) {
  double dot = 0.0;
  if (run_1d<float>(KCONV, OP_D2, d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, &dot, nullptr,
                    nullptr, false) == LC_OK && dot_in)
    *dot_in = (float)dot;
}
void s_s0s0_1d_sym8_d2_dotin64(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                               int narr, const float* x, float* y, float a, float a_x, float a_y,
                               double* dot_in) {
  double dot = 0.0;
  if (run_1d<float>(KCONV, OP_D2, d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, &dot, nullptr,
                    nullptr, false) == LC_OK && dot_in)
    *dot_in = dot;
}

#define LC_WAV1D(NAME, R, KIND, WAV)                                                             \
  void NAME(int d, int idim, const int* n, int bc, const int* nx, const int* ny, int narr,       \
            const R* x, R* y, R a, R a_y) {                                                      \
    run_1d<R>(KIND, WAV, d, idim, n, bc, nx, ny, narr, x, y, a, 0.0, a_y, nullptr, nullptr,      \
              nullptr, false);                                                                   \
  }
LC_WAV1D(d_s0s1_1d_sym8, double, KDWT, WAV_SYM8)
LC_WAV1D(s_s0s1_1d_sym8, float, KDWT, WAV_SYM8)
LC_WAV1D(d_s1s0_1d_sym8, double, KIDWT, WAV_SYM8)
LC_WAV1D(s_s1s0_1d_sym8, float, KIDWT, WAV_SYM8)
LC_WAV1D(d_s0s1_1d_symicomb8, double, KDWT, WAV_SYMICOMB8)
LC_WAV1D(s_s0s1_1d_symicomb8, float, KDWT, WAV_SYMICOMB8)
LC_WAV1D(d_s1s0_1d_symicomb8, double, KIDWT, WAV_SYMICOMB8)
LC_WAV1D(s_s1s0_1d_symicomb8, float, KIDWT, WAV_SYMICOMB8)
#undef LC_WAV1D

void d_s0s0_1d_sym8_md_cost(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                            int narr, const double* x, double* y, double a, double a_x, double a_y,
                            int* cost) {
  (void)bc; (void)nx; (void)ny; (void)x; (void)y; (void)a; (void)a_x; (void)a_y;
  if (cost) *cost = (int)cost_1d(d, idim, n, narr, 32, 1);
}
void d_s0s0_1d_sym8_d2_dotin_cost(int d, int idim, const int* n, int bc, const int* nx,
                                  const int* ny, int narr, const double* x, double* y, double a,
                                  double a_x, double a_y, double* dot_in, int* cost) {
  (void)bc; (void)nx; (void)ny; (void)x; (void)y; (void)a; (void)a_x; (void)a_y; (void)dot_in;
  if (cost) *cost = (int)cost_1d(d, idim, n, narr, 58, 1);
}
void d_s0s1_1d_sym8_cost(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                         int narr, const double* x, double* y, double a, double a_y, int* cost) {
  (void)bc; (void)nx; (void)ny; (void)x; (void)y; (void)a; (void)a_y;
  if (cost) *cost = (int)cost_1d(d, idim, n, narr, 64, 2);
}
void d_s1s0_1d_sym8_cost(int d, int idim, const int* n, int bc, const int* nx, const int* ny,
                         int narr, const double* x, double* y, double a, double a_y, int* cost) {
  (void)bc; (void)nx; (void)ny; (void)x; (void)y; (void)a; (void)a_y;
  if (cost) *cost = (int)cost_1d(d, idim, n, narr, 64, 2);
}

// ---------------------------------------------------------------- md C names
void d_sym8_md_ld_narr(int d, const int* n, const int* bc, const int* nx, const int* ny, int narr,
                       const double* x, double* y, double* w1, double* w2) {
  (void)w1; (void)w2;
  run_md<double>(KCONV, OP_MD, d, n, bc, nx, ny, narr, x, y, 1.0, 0.0, 0.0, nullptr, false);
}
void s_sym8_md_ld_narr(int d, const int* n, const int* bc, const int* nx, const int* ny, int narr,
                       const float* x, float* y, float* w1, float* w2) {
  (void)w1; (void)w2;
  run_md<float>(KCONV, OP_MD, d, n, bc, nx, ny, narr, x, y, 1.0, 0.0, 0.0, nullptr, false);
}
void d_sym8_imd_ld_narr(int d, const int* n, const int* bc, const int* nx, const int* ny, int narr,
                        const double* x, double* y, double* w1, double* w2) {
  (void)w1; (void)w2;
  run_md<double>(KCONV, OP_IMD, d, n, bc, nx, ny, narr, x, y, 1.0, 0.0, 0.0, nullptr, false);
}
void s_sym8_imd_ld_narr(int d, const int* n, const int* bc, const int* nx, const int* ny, int narr,
                        const float* x, float* y, float* w1, float* w2) {
  (void)w1; (void)w2;
  run_md<float>(KCONV, OP_IMD, d, n, bc, nx, ny, narr, x, y, 1.0, 0.0, 0.0, nullptr, false);
}

void d_sym8_d2(int d, const int* n, const int* bc, const double* x, double* y, double* w1, double* w2,
               const double* a, double a_x, double a_y) {
  (void)w1; (void)w2;
  if (d < 1 || d > 16 || !n || !bc) { fail(LC_ERR_ARG, "sym8_d2: bad d or null n/bc"); return; }
  int nx[16], ny[16];
  kin_extents(d, n, bc, nx, ny);
  run_1ds<double>(OP_D2, d, n, bc, nx, ny, 1, a, a_x, a_y, x, y, nullptr, nullptr, nullptr, false);
}
void s_sym8_d2(int d, const int* n, const int* bc, const float* x, float* y, float* w1, float* w2,
               const float* a, float a_x, float a_y) {
  (void)w1; (void)w2;
  if (d < 1 || d > 16 || !n || !bc || !a) { fail(LC_ERR_ARG, "sym8_d2: bad d or null n/bc/a"); return; }
  int nx[16], ny[16];
  kin_extents(d, n, bc, nx, ny);
  double ad[16];
  for (int i = 0; i < d; ++i) ad[i] = a[i];
  run_1ds<float>(OP_D2, d, n, bc, nx, ny, 1, ad, a_x, a_y, x, y, nullptr, nullptr, nullptr, false);
}
void d_sym8_d2_dotin(int d, const int* n, const int* bc, const double* x, double* y, double* w1,
                     double* w2, const double* a, double a_x, double a_y, double* dot_in) {
  (void)w1; (void)w2;
  if (d < 1 || d > 16 || !n || !bc) { fail(LC_ERR_ARG, "sym8_d2: bad d or null n/bc"); return; }
  int nx[16], ny[16];
  kin_extents(d, n, bc, nx, ny);
  double dot[16] = {0};
  if (run_1ds<double>(OP_D2, d, n, bc, nx, ny, 1, a, a_x, a_y, x, y, dot, nullptr, nullptr, false) == LC_OK && dot_in)
    for (int i = 0; i < d; ++i) dot_in[i] = dot[i];
}
void s_sym8_d2_dotin64(int d, const int* n, const int* bc, const float* x, float* y, float* w1,
                       float* w2, const float* a, float a_x, float a_y, double* dot_in) {
  (void)w1; (void)w2;
  if (d < 1 || d > 16 || !n || !bc || !a) { fail(LC_ERR_ARG, "sym8_d2: bad d or null n/bc/a"); return; }
  int nx[16], ny[16];
  kin_extents(d, n, bc, nx, ny);
  double ad[16], dot[16] = {0};
  for (int i = 0; i < d; ++i) ad[i] = a[i];
  if (run_1ds<float>(OP_D2, d, n, bc, nx, ny, 1, ad, a_x, a_y, x, y, dot, nullptr, nullptr, false) == LC_OK && dot_in)
    for (int i = 0; i < d; ++i) dot_in[i] = dot[i];
}

#define LC_WAVMD(NAME, R, KIND)                                                                  \
  void NAME(int d, const int* n, const int* bc, const R* x, R* y, R* w1, R* w2) {                 \
    (void)w1; (void)w2;                                                                          \
    if (d < 1 || d > 16 || !n || !bc) { fail(LC_ERR_ARG, #NAME ": bad d or null n/bc"); return; } \
    int nx[16], ny[16];                                                                          \
    wav_extents(d, n, bc, KIND == KIDWT, nx, ny);                                                \
    run_md<R>(KIND, WAV_SYM8, d, n, bc, nx, ny, 1, x, y, 1.0, 0.0, 0.0, nullptr, false);          \
  }
LC_WAVMD(d_dwt_sym8, double, KDWT)
LC_WAVMD(s_dwt_sym8, float, KDWT)
LC_WAVMD(d_idwt_sym8, double, KIDWT)
LC_WAVMD(s_idwt_sym8, float, KIDWT)
#undef LC_WAVMD

// ---------------------------------------------------------------- Fortran names
void s0s1_1d_sym8_lp_d_(const int* d, const int* idim, const int* n, const int* bc_id, const int* nx,
                        const int* ny, const int* narr, const double* a, const double* a_y,
                        const double* x, double* y) {
  run_1d<double>(KDWT, WAV_SYM8, *d, *idim, n, *bc_id, nx, ny, *narr, x, y, *a, 0.0, *a_y, nullptr,
                 nullptr, nullptr, false);
}
void s1s0_1d_sym8_lp_d_(const int* d, const int* idim, const int* n, const int* bc_id, const int* nx,
                        const int* ny, const int* narr, const double* a, const double* a_y,
                        const double* x, double* y) {
  run_1d<double>(KIDWT, WAV_SYM8, *d, *idim, n, *bc_id, nx, ny, *narr, x, y, *a, 0.0, *a_y, nullptr,
                 nullptr, nullptr, false);
}
void s0s1_sym8_lp_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                     const int* narr, const double* a, const double* a_y, const double* x, double* y,
                     double* work) {
  (void)work;
  run_md<double>(KDWT, WAV_SYM8, *d, n, bc, nx, ny, *narr, x, y, *a, 0.0, *a_y, nullptr, false);
}
void s1s0_sym8_lp_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                     const int* narr, const double* a, const double* a_y, const double* x, double* y,
                     double* work) {
  (void)work;
  run_md<double>(KIDWT, WAV_SYM8, *d, n, bc, nx, ny, *narr, x, y, *a, 0.0, *a_y, nullptr, false);
}
void s0s0_1d_sym8_md_d_(const int* d, const int* idim, const int* n, const int* bc_id, const int* nx,
                        const int* ny, const int* narr, const double* a, const double* a_x,
                        const double* a_y, const double* x, double* y) {
  run_1d<double>(KCONV, OP_MD, *d, *idim, n, *bc_id, nx, ny, *narr, x, y, *a, *a_x, *a_y, nullptr,
                 nullptr, nullptr, false);
}
void s0s0_1d_sym8_mi_d_(const int* d, const int* idim, const int* n, const int* bc_id, const int* nx,
                        const int* ny, const int* narr, const double* a, const double* a_x,
                        const double* a_y, const double* x, double* y) {
  run_1d<double>(KCONV, OP_IMD, *d, *idim, n, *bc_id, nx, ny, *narr, x, y, *a, *a_x, *a_y, nullptr,
                 nullptr, nullptr, false);
}
void s0s0_1ds_sym8_d2_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                         const int* narr, const double* a, const double* a_x, const double* a_y,
                         const double* x, double* y, double* work) {
  (void)work;
  run_1ds<double>(OP_D2, *d, n, bc, nx, ny, *narr, a, *a_x, *a_y, x, y, nullptr, nullptr, nullptr, false);
}
void s0s0_1ds_dot_sym8_d2_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                             const int* narr, const double* a, const double* a_x, const double* a_y,
                             const double* x, double* y, double* dot_in, double* work) {
  (void)work;
  run_1ds<double>(OP_D2, *d, n, bc, nx, ny, *narr, a, *a_x, *a_y, x, y, dot_in, nullptr, nullptr, false);
}
void s0s0_1ds_sym8_d1_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                         const int* narr, const double* a, const double* a_x, const double* a_y,
                         const double* x, double* y, double* work) {
  (void)work;
  run_1ds<double>(OP_D1, *d, n, bc, nx, ny, *narr, a, *a_x, *a_y, x, y, nullptr, nullptr, nullptr, false);
}
void s0s0_sym8_md_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                     const int* narr, const double* a, const double* a_x, const double* a_y,
                     const double* x, double* y, double* work) {
  (void)work;
  run_md<double>(KCONV, OP_MD, *d, n, bc, nx, ny, *narr, x, y, *a, *a_x, *a_y, nullptr, false);
}
void s0s0_sym8_mi_d_(const int* d, const int* n, const int* bc, const int* nx, const int* ny,
                     const int* narr, const double* a, const double* a_x, const double* a_y,
                     const double* x, double* y, double* work) {
  (void)work;
  run_md<double>(KCONV, OP_IMD, *d, n, bc, nx, ny, *narr, x, y, *a, *a_x, *a_y, nullptr, false);
}

// ---------------------------------------------------------------- stream API
int lc_conv_1d(int prec, int fil, int d, int idim, const int* n, int bc, const int* nx,
               const int* ny, int narr, const void* x, void* y, double a, double a_x, double a_y,
               double* dot_dev, void* stream) {
  if (prec == 8)
    return run_1d<double>(KCONV, fil, d, idim, n, bc, nx, ny, narr, (const double*)x, (double*)y, a,
                          a_x, a_y, nullptr, dot_dev, hs(stream), true);
  if (prec == 4)
    return run_1d<float>(KCONV, fil, d, idim, n, bc, nx, ny, narr, (const float*)x, (float*)y, a,
                         a_x, a_y, nullptr, dot_dev, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_conv_1d: prec %d not in {4, 8}", prec);
}
int lc_wavelet_1d(int prec, int wav, int inverse, int d, int idim, const int* n, int bc,
                  const int* nx, const int* ny, int narr, const void* x, void* y, double a,
                  double a_y, void* stream) {
  const int kind = inverse ? KIDWT : KDWT;
  if (prec == 8)
    return run_1d<double>(kind, wav, d, idim, n, bc, nx, ny, narr, (const double*)x, (double*)y, a,
                          0.0, a_y, nullptr, nullptr, hs(stream), true);
  if (prec == 4)
    return run_1d<float>(kind, wav, d, idim, n, bc, nx, ny, narr, (const float*)x, (float*)y, a,
                         0.0, a_y, nullptr, nullptr, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_wavelet_1d: prec %d not in {4, 8}", prec);
}
int lc_conv_md(int prec, int fil, int d, const int* n, const int* bc, const int* nx, const int* ny,
               int narr, const void* x, void* y, double a, double a_x, double a_y, void* stream) {
  if (prec == 8)
    return run_md<double>(KCONV, fil, d, n, bc, nx, ny, narr, (const double*)x, (double*)y, a, a_x,
                          a_y, hs(stream), true);
  if (prec == 4)
    return run_md<float>(KCONV, fil, d, n, bc, nx, ny, narr, (const float*)x, (float*)y, a, a_x,
                         a_y, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_conv_md: prec %d not in {4, 8}", prec);
}
int lc_conv_1ds(int prec, int fil, int d, const int* n, const int* bc, const int* nx, const int* ny,
                int narr, const double* a, double a_x, double a_y, const void* x, void* y,
                double* dot_dev, void* stream) {
  if (prec == 8)
    return run_1ds<double>(fil, d, n, bc, nx, ny, narr, a, a_x, a_y, (const double*)x, (double*)y,
                           nullptr, dot_dev, hs(stream), true);
  if (prec == 4)
    return run_1ds<float>(fil, d, n, bc, nx, ny, narr, a, a_x, a_y, (const float*)x, (float*)y,
                          nullptr, dot_dev, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_conv_1ds: prec %d not in {4, 8}", prec);
}
int lc_conv_1ds_noalloc(int prec, int fil, int d, const int* n, const int* bc, const int* nx,
                        const int* ny, int narr, const double* a, double a_x, double a_y, const void* x,
                        void* y, double* dot_dev, void* stream) {
  if (prec == 8)
    return run_1ds<double>(fil, d, n, bc, nx, ny, narr, a, a_x, a_y, (const double*)x, (double*)y,
                           nullptr, dot_dev, hs(stream), true, true);
  if (prec == 4)
    return run_1ds<float>(fil, d, n, bc, nx, ny, narr, a, a_x, a_y, (const float*)x, (float*)y,
                          nullptr, dot_dev, hs(stream), true, true);
  return fail(LC_ERR_ARG, "lc_conv_1ds_noalloc: prec %d not in {4, 8}", prec);
}
int lc_wavelet_md(int prec, int wav, int inverse, int d, const int* n, const int* bc,
                  const int* nx, const int* ny, int narr, const void* x, void* y, double a,
                  double a_y, void* stream) {
  const int kind = inverse ? KIDWT : KDWT;
  if (prec == 8)
    return run_md<double>(kind, wav, d, n, bc, nx, ny, narr, (const double*)x, (double*)y, a, 0.0,
                          a_y, hs(stream), true);
  if (prec == 4)
    return run_md<float>(kind, wav, d, n, bc, nx, ny, narr, (const float*)x, (float*)y, a, 0.0, a_y,
                         hs(stream), true);
  return fail(LC_ERR_ARG, "lc_wavelet_md: prec %d not in {4, 8}", prec);
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Independent 3D operator applications grouped into stages: consecutive
// items of the same kind and operator form one stage, executed as one
// launch; stages run in item order on the stream.
struct lc_batch {
  int prec = 8;
  int count = 0;
  std::vector<lc_batch_item> items;  // kept for re-building stages when tuning
  std::vector<BatchStage> stages;
  hipStream_t cap = nullptr;     // private stream the graph is captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PersistPlan* persist = nullptr;  // one-launch form, built on first use
  hipGraph_t pgraph = nullptr;
  hipGraphExec_t pexec = nullptr;
};

namespace {

template <typename R>
BatchItem<R> to_item(const lc_batch_item& u) {
  BatchItem<R> it{};
  for (int i = 0; i < 3; ++i) {
    it.b.n[i] = u.n[i]; it.b.bc[i] = u.bc[i]; it.b.nx[i] = u.nx[i]; it.b.ny[i] = u.ny[i];
    it.a[i] = u.a[i];
  }
  it.x = static_cast<const R*>(u.x);
  it.y = static_cast<R*>(u.y);
  it.a_x = u.a_x;
  it.a_y = u.a_y;
  it.dot = u.dot;
  return it;
}

int check_item(const lc_batch_item& u, int k) {
  const char* who = "lc_batch_create";
  const bool wav = u.kind == LC_BATCH_DWT || u.kind == LC_BATCH_IDWT;
  if (u.kind < LC_BATCH_SEP || u.kind > LC_BATCH_IDWT)
    return fail(LC_ERR_ARG, "%s: item %d has kind %d", who, k, u.kind);
  const int W = wav ? wav_window(u.op) : conv_window(u.op);
  if (W < 0) return fail(LC_ERR_ARG, "%s: item %d has unknown operator %d", who, k, u.op);
  for (int i = 0; i < 3; ++i) {
    int rc = check_line(u.bc[i], u.n[i], u.nx[i], u.ny[i], W, who, !wav);
    if (rc) return rc;
  }
  if (!u.x || !u.y || !is_device_ptr(u.x) || !is_device_ptr(u.y))
    return fail(LC_ERR_ARG, "%s: item %d: x and y must be device pointers", who, k);
  if (u.dot && (u.kind != LC_BATCH_SUM || !is_device_ptr(u.dot)))
    return fail(LC_ERR_ARG, "%s: item %d: dot must be a device pointer of a SUM item", who, k);
  if (u.kind != LC_BATCH_SUM && u.a_x != 0.0)
    return fail(LC_ERR_UNSUPPORTED, "%s: item %d: a_x is supported for SUM items only", who, k);
  if (wav && u.op != WAV_SYM8)
    return fail(LC_ERR_UNSUPPORTED, "%s: item %d: batched wavelets support sym8 only", who, k);
  return LC_OK;
}

// stage key: kind, operator and, for the symicomb8 wavelet (whose filter
// centre differs between periodic and free lines), periodic vs free
long stage_key(const lc_batch_item& u) {
  const bool wav = u.kind == LC_BATCH_DWT || u.kind == LC_BATCH_IDWT;
  const bool per = wav && u.op == WAV_SYMICOMB8 && u.bc[0] == 0;
  return ((long)u.kind << 16) | ((long)u.op << 1) | (per ? 1 : 0);
}

int item_form(const lc_batch_item& u) {
  switch (u.kind) {
    case LC_BATCH_SEP: return FORM_SEP;
    case LC_BATCH_SUM: return FORM_SUM;
    case LC_BATCH_DWT: return FORM_DWT;
  }
  return FORM_IDWT;
}

Box3 item_box(const lc_batch_item& u) {
  Box3 b;
  for (int i = 0; i < 3; ++i) { b.n[i] = u.n[i]; b.bc[i] = u.bc[i]; b.nx[i] = u.nx[i]; b.ny[i] = u.ny[i]; }
  return b;
}

// variant of a new stage: the tuned one for its largest item's shape when it
// has a batched form, else the default
int stage_default_variant(const lc_batch_item* it, int count, int prec) {
  int big = 0;
  long most = -1;
  for (int i = 0; i < count; ++i) {
    const long p = (long)it[i].ny[0] * it[i].ny[1] * it[i].ny[2];
    if (p > most) { most = p; big = i; }
  }
  const int form = item_form(it[big]);
  const int v = select_variant(prec, form, it[big].op, item_box(it[big]));
  if (has_batch_form(form, v)) return v;
  bool fit = true;
  for (int i = 0; i < count; ++i) fit = fit && plane_fits(item_box(it[i]), prec);
  if (form == FORM_SUM) return !fit ? 8 : prec == 4 ? 32 : 29;  // fastest batched kinetic forms
  return form == FORM_SEP ? 12 : 0;
}

template <typename R>
int build_stages(lc_batch* b, int count, const lc_batch_item* items) {
  int k = 0;
  while (k < count) {
    int e = k + 1;
    while (e < count && stage_key(items[e]) == stage_key(items[k])) ++e;
    std::vector<BatchItem<R>> its;
    for (int i = k; i < e; ++i) its.push_back(to_item<R>(items[i]));
    BatchStage st;
    st.kind = items[k].kind;
    st.id = items[k].op;
    st.bc = items[k].bc[0];
    st.count = e - k;
    st.variant = stage_default_variant(items + k, e - k, sizeof(R));
    hipError_t er;
    if (st.kind == LC_BATCH_SEP || st.kind == LC_BATCH_SUM)
      er = fused_stage<R>(true, st.kind == LC_BATCH_SUM ? BK_SUM : BK_SEP, st.id, its.data(), st.count, &st,
                          nullptr);
    else
      er = wav_stage<R>(true, st.kind == LC_BATCH_DWT ? KDWT : KIDWT, st.id, st.bc, its.data(), st.count,
                        &st, nullptr);
    b->stages.push_back(st);
    if (er != hipSuccess) return hip_fail(er, "lc_batch_create");
    k = e;
  }
  return LC_OK;
}

template <typename R>
int launch_stages(lc_batch* b, hipStream_t st) {
  for (BatchStage& s : b->stages) {
    hipError_t er;
    if (s.kind == LC_BATCH_SEP || s.kind == LC_BATCH_SUM)
      er = fused_stage<R>(false, s.kind == LC_BATCH_SUM ? BK_SUM : BK_SEP, s.id, nullptr, 0, &s, st);
    else
      er = wav_stage<R>(false, s.kind == LC_BATCH_DWT ? KDWT : KIDWT, s.id, s.bc, nullptr, 0, &s, st);
    if (er != hipSuccess) return hip_fail(er, "lc_batch_run");
  }
  return LC_OK;
}

// (re)build stage s of plan b with kernel variant v
template <typename R>
hipError_t rebuild_stage(lc_batch* b, size_t si, int first_item, int v) {
  BatchStage& st = b->stages[si];
  BatchStage fresh;
  fresh.kind = st.kind; fresh.id = st.id; fresh.bc = st.bc; fresh.count = st.count;
  fresh.variant = v;
  std::vector<BatchItem<R>> its;
  for (int i = first_item; i < first_item + st.count; ++i) its.push_back(to_item<R>(b->items[i]));
  hipError_t er;
  if (st.kind == LC_BATCH_SEP || st.kind == LC_BATCH_SUM)
    er = fused_stage<R>(true, st.kind == LC_BATCH_SUM ? BK_SUM : BK_SEP, st.id, its.data(), st.count,
                        &fresh, nullptr);
  else
    er = wav_stage<R>(true, st.kind == LC_BATCH_DWT ? KDWT : KIDWT, st.id, st.bc, its.data(), st.count,
                      &fresh, nullptr);
  if (er != hipSuccess) {
    batch_stage_free(&fresh);
    return er;
  }
  batch_stage_free(&st);
  st = fresh;
  return hipSuccess;
}

template <typename R>
hipError_t launch_one(BatchStage& s, hipStream_t st) {
  if (s.kind == LC_BATCH_SEP || s.kind == LC_BATCH_SUM)
    return fused_stage<R>(false, s.kind == LC_BATCH_SUM ? BK_SUM : BK_SEP, s.id, nullptr, 0, &s, st);
  return wav_stage<R>(false, s.kind == LC_BATCH_DWT ? KDWT : KIDWT, s.id, s.bc, nullptr, 0, &s, st);
}

// every batched variant of every stage, min of `repeat` timed runs; keeps
// the fastest per stage
template <typename R>
int tune_stages(lc_batch* b, int repeat, double* ms_out) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int first = 0;
  int rc = LC_OK;
  for (size_t si = 0; si < b->stages.size() && rc == LC_OK; ++si) {
    const int form = item_form(b->items[first]);
    int best = b->stages[si].variant;
    double bt = 1e300;
    for (int v = 0; v < variant_count(form) && rc == LC_OK; ++v) {
      if (!has_batch_form(form, v)) continue;
      hipError_t er = rebuild_stage<R>(b, si, first, v);
      if (er == hipErrorInvalidValue) {  // the variant does not take these items (the stage is unchanged)
        (void)hipGetLastError();
        continue;
      }
      if (er != hipSuccess) { rc = hip_fail(er, "lc_batch_tune"); break; }
      double tmin = 1e300;
      for (int r = 0; r <= std::max(1, repeat); ++r) {  // first run untimed
        er = hipEventRecord(e0, nullptr);
        if (er == hipSuccess) er = launch_one<R>(b->stages[si], nullptr);
        if (er == hipSuccess) er = hipEventRecord(e1, nullptr);
        if (er == hipSuccess) er = hipEventSynchronize(e1);
        float ms = 0.f;
        if (er == hipSuccess) er = hipEventElapsedTime(&ms, e0, e1);
        if (er != hipSuccess) { rc = hip_fail(er, "lc_batch_tune"); break; }
        if (r > 0) tmin = std::min(tmin, (double)ms);
      }
      if (tmin < bt) { bt = tmin; best = v; }
    }
    if (rc == LC_OK && b->stages[si].variant != best) {
      hipError_t er = rebuild_stage<R>(b, si, first, best);
      if (er != hipSuccess) rc = hip_fail(er, "lc_batch_tune");
    }
    if (ms_out) ms_out[si] = bt;
    first += b->stages[si].count;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

int lc_batch_create(lc_batch** plan, int prec, int count, const lc_batch_item* items) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_create: null plan");
  *plan = nullptr;
  if (prec != 4 && prec != 8) return fail(LC_ERR_ARG, "lc_batch_create: prec %d not in {4, 8}", prec);
  if (count < 0 || (count > 0 && !items)) return fail(LC_ERR_ARG, "lc_batch_create: bad item list");
  for (int k = 0; k < count; ++k) {
    int rc = check_item(items[k], k);
    if (rc) return rc;
  }
  lc_batch* b = new lc_batch;
  b->prec = prec;
  b->count = count;
  b->items.assign(items, items + count);
  int rc = prec == 8 ? build_stages<double>(b, count, items) : build_stages<float>(b, count, items);
  if (rc) {
    lc_batch_destroy(b);
    return rc;
  }
  *plan = b;
  return LC_OK;
}

int lc_batch_run(lc_batch* plan, void* stream) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_run: null plan");
  return plan->prec == 8 ? launch_stages<double>(plan, hs(stream)) : launch_stages<float>(plan, hs(stream));
}

int lc_batch_run_graph(lc_batch* plan, void* stream) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_run_graph: null plan");
  if (!plan->exec) {
    if (!plan->cap) HIPCHK(hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
    int rc = plan->prec == 8 ? launch_stages<double>(plan, plan->cap) : launch_stages<float>(plan, plan->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(plan->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "lc_batch_run_graph: capture");
    plan->graph = g;
    HIPCHK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
  }
  HIPCHK(hipGraphLaunch(plan->exec, hs(stream)));
  return LC_OK;
}

int lc_batch_info(const lc_batch* plan, int* stages, long* blocks) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_info: null plan");
  long nb = 0;
  for (const BatchStage& s : plan->stages) nb += s.blocks;
  if (stages) *stages = (int)plan->stages.size();
  if (blocks) *blocks = nb;
  return LC_OK;
}

static int persist_ready(lc_batch* plan) {
  if (plan->persist) return LC_OK;
  const int n = plan->count;
  std::vector<int> kinds(n), ids(n);
  for (int i = 0; i < n; ++i) {
    const lc_batch_item& u = plan->items[i];
    kinds[i] = u.kind == LC_BATCH_SEP ? BK_SEP : u.kind == LC_BATCH_SUM ? BK_SUM
             : u.kind == LC_BATCH_DWT ? BK_DWT : BK_IDWT;
    ids[i] = u.op;
    const bool ok = (u.kind == LC_BATCH_SEP && (u.op == LC_FIL_MD || u.op == LC_FIL_IMD)) ||
                    (u.kind == LC_BATCH_SUM && u.op == LC_FIL_D2) ||
                    ((u.kind == LC_BATCH_DWT || u.kind == LC_BATCH_IDWT) && u.op == LC_WAV_SYM8);
    if (!ok)
      return fail(LC_ERR_UNSUPPORTED, "lc_batch_run_persistent: item %d (kind %d, operator %d) has no "
                  "persistent form (SEP md/imd, SUM d2, sym8 DWT/IDWT)", i, u.kind, u.op);
  }
  hipError_t e;
  if (plan->prec == 8) {
    std::vector<BatchItem<double>> its;
    for (int i = 0; i < n; ++i) its.push_back(to_item<double>(plan->items[i]));
    e = persist_build<double>(its.data(), kinds.data(), ids.data(), n, &plan->persist);
  } else {
    std::vector<BatchItem<float>> its;
    for (int i = 0; i < n; ++i) its.push_back(to_item<float>(plan->items[i]));
    e = persist_build<float>(its.data(), kinds.data(), ids.data(), n, &plan->persist);
  }
  if (e != hipSuccess) return hip_fail(e, "lc_batch_run_persistent: build");
  return LC_OK;
}

static int persist_spin_limit() {
  const char* e = getenv("LIBCONV_PERSIST_SPIN_LIMIT");
  const long v = e ? atol(e) : 0;
  return v > 0 && v < 0x7fffffffL ? (int)v : (1 << 24);
}

int lc_batch_run_persistent(lc_batch* plan, void* stream) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_run_persistent: null plan");
  int rc = persist_ready(plan);
  if (rc) return rc;
  hipError_t e = persist_launch(plan->persist, hs(stream), persist_spin_limit());
  return e == hipSuccess ? LC_OK : hip_fail(e, "lc_batch_run_persistent");
}

int lc_batch_run_persistent_graph(lc_batch* plan, void* stream) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_run_persistent_graph: null plan");
  int rc = persist_ready(plan);
  if (rc) return rc;
  if (!plan->pexec) {
    if (!plan->cap) HIPCHK(hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
    hipError_t le = persist_launch(plan->persist, plan->cap, persist_spin_limit());
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(plan->cap, &g);
    if (le != hipSuccess || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(le != hipSuccess ? le : e, "lc_batch_run_persistent_graph: capture");
    }
    plan->pgraph = g;
    HIPCHK(hipGraphInstantiate(&plan->pexec, plan->pgraph, nullptr, nullptr, 0));
  }
  HIPCHK(hipGraphLaunch(plan->pexec, hs(stream)));
  return LC_OK;
}

int lc_batch_persistent_info(lc_batch* plan, int* grid, int* status) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_persistent_info: null plan");
  int rc = persist_ready(plan);
  if (rc) return rc;
  if (grid) *grid = persist_grid(plan->persist);
  if (status) {
    HIPCHK(hipDeviceSynchronize());
    hipError_t e = persist_status(plan->persist, status);
    if (e != hipSuccess) return hip_fail(e, "lc_batch_persistent_info");
  }
  return LC_OK;
}

int lc_batch_tune(lc_batch* plan, int repeat, double* stage_ms) {
  if (!plan) return fail(LC_ERR_ARG, "lc_batch_tune: null plan");
  HIPCHK(hipDeviceSynchronize());
  if (plan->exec) { (void)hipGraphExecDestroy(plan->exec); plan->exec = nullptr; }
  if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
  const int rc = plan->prec == 8 ? tune_stages<double>(plan, repeat, stage_ms)
                                 : tune_stages<float>(plan, repeat, stage_ms);
  if (rc) return rc;
  HIPCHK(hipDeviceSynchronize());
  return LC_OK;
}

int lc_batch_stage_variant(const lc_batch* plan, int stage) {
  if (!plan || stage < 0 || stage >= (int)plan->stages.size())
    return fail(LC_ERR_ARG, "lc_batch_stage_variant: bad plan or stage");
  return plan->stages[stage].variant;
}

// ---------------------------------------------------------------- tuner
static int form_check(int form, int op, const int* n, const int* bc, const int* nx, const int* ny,
                      const char* who, Box3* b) {
  if (form < LC_FORM_SEP || form > LC_FORM_IDWT) return fail(LC_ERR_ARG, "%s: form %d", who, form);
  const bool wav = form == LC_FORM_DWT || form == LC_FORM_IDWT;
  const int W = wav ? (op == WAV_SYM8 ? 8 : -1) : conv_window(op);
  if (W < 0) return fail(LC_ERR_ARG, "%s: operator %d not tunable for form %d", who, op, form);
  if (!n || !bc || !nx || !ny) return fail(LC_ERR_ARG, "%s: null size array", who);
  for (int i = 0; i < 3; ++i) {
    int rc = check_line(bc[i], n[i], nx[i], ny[i], W, who, !wav);
    if (rc) return rc;
    b->n[i] = n[i]; b->bc[i] = bc[i]; b->nx[i] = nx[i]; b->ny[i] = ny[i];
  }
  return LC_OK;
}

int lc_tune_md(int prec, int form, int op, const int* n, const int* bc, const int* nx, const int* ny,
               int repeat, int* best, double* best_ms, double* times, int ntimes) {
  Box3 b;
  int rc = form_check(form, op, n, bc, nx, ny, "lc_tune_md", &b);
  if (rc) return rc;
  if (prec != 4 && prec != 8) return fail(LC_ERR_ARG, "lc_tune_md: prec %d not in {4, 8}", prec);
  std::vector<double> t;
  int bv = -1;
  double bt = 0.0;
  hipError_t e = prec == 8 ? tune_md<double>(form, op, b, repeat, &bv, &bt, &t)
                           : tune_md<float>(form, op, b, repeat, &bv, &bt, &t);
  if (e != hipSuccess) return hip_fail(e, "lc_tune_md");
  if (best) *best = bv;
  if (best_ms) *best_ms = bt;
  for (int i = 0; times && i < ntimes && i < (int)t.size(); ++i) times[i] = t[i];
  return LC_OK;
}

int lc_tune_save(const char* path) {
  if (!path) return fail(LC_ERR_ARG, "lc_tune_save: null path");
  const int n = tune_save(path);
  return n < 0 ? fail(LC_ERR_ARG, "lc_tune_save: cannot write %s", path) : LC_OK;
}

int lc_tune_load(const char* path) {
  if (!path) return fail(LC_ERR_ARG, "lc_tune_load: null path");
  const int n = tune_load(path);
  return n < 0 ? fail(LC_ERR_ARG, "lc_tune_load: cannot read %s", path) : LC_OK;
}

void lc_tune_clear(void) { tune_clear(); }

int lc_variant_count(int form) {
  if (form < LC_FORM_SEP || form > LC_FORM_IDWT) return fail(LC_ERR_ARG, "lc_variant_count: form %d", form);
  return variant_count(form);
}

int lc_kernel_name(int prec, int form, int op, const int* n, const int* bc, const int* nx, const int* ny,
                   char* buf, int len) {
  Box3 b;
  int rc = form_check(form, op, n, bc, nx, ny, "lc_kernel_name", &b);
  if (rc) return rc;
  if (!buf || len < 1) return fail(LC_ERR_ARG, "lc_kernel_name: no buffer");
  const int v = select_variant(prec, form, op, b);
  const std::string s = form == LC_FORM_DWT || form == LC_FORM_IDWT ? wav_kernel_name(prec, form, op, v, false)
                                                                     : fused_kernel_name(prec, form, op, v, false);
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return LC_OK;
}

// ---- Poisson nabla operators (BigDFT/Poisson2.rb)
int lc_poisson_1d(int prec, int nord, int idim, const int* n, int bc, const void* u, void* du, double a,
                  double a_y, void* stream) {
  if (prec == 8)
    return run_poisson<double>(PO_1D, nord, idim, bc, 0, n, (const double*)u, (double*)du, nullptr, a, a_y,
                               nullptr, hs(stream), true);
  if (prec == 4)
    return run_poisson<float>(PO_1D, nord, idim, bc, 0, n, (const float*)u, (float*)du, nullptr, a, a_y,
                              nullptr, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_poisson_1d: prec %d not in {4, 8}", prec);
}
int lc_div_u_i(int prec, int geocode, const int* n, const void* u, void* du, int nord, const double* h, void* cc,
               void* stream) {
  if (prec == 8)
    return run_poisson<double>(PO_DIV, nord, -1, 0, geocode, n, (const double*)u, (double*)du, (double*)cc, 1.0,
                               0.0, h, hs(stream), true);
  if (prec == 4)
    return run_poisson<float>(PO_DIV, nord, -1, 0, geocode, n, (const float*)u, (float*)du, (float*)cc, 1.0,
                              0.0, h, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_div_u_i: prec %d not in {4, 8}", prec);
}
int lc_nabla_u(int prec, int geocode, const int* n, const void* u, void* du, int nord, const double* h,
               void* stream) {
  if (prec == 8)
    return run_poisson<double>(PO_NABLA, nord, -1, 0, geocode, n, (const double*)u, (double*)du, nullptr, 1.0,
                               0.0, h, hs(stream), true);
  if (prec == 4)
    return run_poisson<float>(PO_NABLA, nord, -1, 0, geocode, n, (const float*)u, (float*)du, nullptr, 1.0, 0.0,
                              h, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_nabla_u: prec %d not in {4, 8}", prec);
}
int lc_nabla_u_and_square(int prec, int geocode, const int* n, const void* u, void* du, void* du2, int nord,
                          const double* h, void* stream) {
  if (prec == 8)
    return run_poisson<double>(PO_NABLA2, nord, -1, 0, geocode, n, (const double*)u, (double*)du, (double*)du2,
                               1.0, 0.0, h, hs(stream), true);
  if (prec == 4)
    return run_poisson<float>(PO_NABLA2, nord, -1, 0, geocode, n, (const float*)u, (float*)du, (float*)du2, 1.0,
                              0.0, h, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_nabla_u_and_square: prec %d not in {4, 8}", prec);
}

int lc_nabla_u_square(int prec, int geocode, const int* n, const void* u, void* ddu, int nord, const double* h,
                      void* stream) {
  if (prec == 8)
    return run_dielectric<double>(PD_SQUARE, geocode, n, (const double*)u, nullptr, (double*)ddu, nord, h, 0.0,
                                  nullptr, hs(stream), true);
  if (prec == 4)
    return run_dielectric<float>(PD_SQUARE, geocode, n, (const float*)u, nullptr, (float*)ddu, nord, h, 0.0,
                                 nullptr, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_nabla_u_square: prec %d not in {4, 8}", prec);
}
int lc_nabla_u_epsilon(int prec, int geocode, const int* n, const void* u, void* du, int nord, const double* h,
                       const void* eps, void* stream) {
  if (prec == 8)
    return run_dielectric<double>(PD_EPS, geocode, n, (const double*)u, (const double*)eps, (double*)du, nord, h,
                                  0.0, nullptr, hs(stream), true);
  if (prec == 4)
    return run_dielectric<float>(PD_EPS, geocode, n, (const float*)u, (const float*)eps, (float*)du, nord, h, 0.0,
                                 nullptr, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_nabla_u_epsilon: prec %d not in {4, 8}", prec);
}
int lc_update_rhopol(int prec, int geocode, const int* n, const void* u, int nord, const double* h, double eta,
                     const void* dlogeps, void* rhopol, double* rhores2_dev, void* stream) {
  if (prec == 8)
    return run_dielectric<double>(PD_RHOPOL, geocode, n, (const double*)u, (const double*)dlogeps,
                                  (double*)rhopol, nord, h, eta, rhores2_dev, hs(stream), true);
  if (prec == 4)
    return run_dielectric<float>(PD_RHOPOL, geocode, n, (const float*)u, (const float*)dlogeps, (float*)rhopol,
                                 nord, h, eta, rhores2_dev, hs(stream), true);
  return fail(LC_ERR_ARG, "lc_update_rhopol: prec %d not in {4, 8}", prec);
}

// the procedure names the BOAST Poisson generator emits (C flavour, FP64,
// host or device arrays): Poisson_broker (Poisson2.rb:1951-2043), div_u_i
// (:2067-2195), nabla_u (:2553-2620), nabla_u_and_square (:2197-2276)
void Poisson_broker(int nord, int idim, const int* nn, int bc, const double* u, double* du, double a) {
  run_poisson<double>(PO_1D, nord, idim, bc, 0, nn, u, du, nullptr, a, 0.0, nullptr, nullptr, false);
}
void div_u_i(int geocode, int n01, int n02, int n03, const double* u, double* du, int nord, const double* hgrids,
             double* cc) {
  const int n[3] = {n01, n02, n03};
  run_poisson<double>(PO_DIV, nord, -1, 0, geocode, n, u, du, cc, 1.0, 0.0, hgrids, nullptr, false);
}
void nabla_u(int geocode, int n01, int n02, int n03, const double* u, double* du, int nord, const double* hgrids) {
  const int n[3] = {n01, n02, n03};
  run_poisson<double>(PO_NABLA, nord, -1, 0, geocode, n, u, du, nullptr, 1.0, 0.0, hgrids, nullptr, false);
}
void nabla_u_and_square(int geocode, int n01, int n02, int n03, const double* u, double* du, double* du2, int nord,
                        const double* hgrids) {
  const int n[3] = {n01, n02, n03};
  run_poisson<double>(PO_NABLA2, nord, -1, 0, geocode, n, u, du, du2, 1.0, 0.0, hgrids, nullptr, false);
}
// Poisson2.rb:2278-2359, :2362-2443, :2446-2550
void nabla_u_square(int geocode, int n01, int n02, int n03, const double* u, double* ddu, int nord,
                    const double* hgrids) {
  const int n[3] = {n01, n02, n03};
  run_dielectric<double>(PD_SQUARE, geocode, n, u, nullptr, ddu, nord, hgrids, 0.0, nullptr, nullptr, false);
}
void nabla_u_epsilon(int geocode, int n01, int n02, int n03, const double* u, double* du, int nord,
                     const double* hgrids, const double* eps) {
  const int n[3] = {n01, n02, n03};
  run_dielectric<double>(PD_EPS, geocode, n, u, eps, du, nord, hgrids, 0.0, nullptr, nullptr, false);
}
void update_rhopol(int geocode, int n01, int n02, int n03, const double* u, int nord, const double* hgrids,
                   double eta, const double* dlogeps, double* rhopol, double* rhores2) {
  const int n[3] = {n01, n02, n03};
  run_dielectric<double>(PD_RHOPOL, geocode, n, u, dlogeps, rhopol, nord, hgrids, eta, rhores2, nullptr, false);
}

long lc_local_hamiltonian_work(const int* n) {
  HamDims h;
  if (ham_dims(n, &h, "lc_local_hamiltonian_work")) return -1;
  return 2 * h.nf + h.nr;
}

int lc_local_hamiltonian(int prec, const int* n, const double* a, const void* psig, void* hpsig,
                         const void* pot, void* rho, void* work, double* ekin_dev, void* stream) {
  if (prec == 8)
    return run_local_hamiltonian<double>(n, a, (const double*)psig, (double*)hpsig, (const double*)pot,
                                         (double*)rho, (double*)work, ekin_dev, hs(stream));
  if (prec == 4)
    return run_local_hamiltonian<float>(n, a, (const float*)psig, (float*)hpsig, (const float*)pot,
                                        (float*)rho, (float*)work, ekin_dev, hs(stream));
  return fail(LC_ERR_ARG, "lc_local_hamiltonian: prec %d not in {4, 8}", prec);
}

void lc_batch_destroy(lc_batch* plan) {
  if (!plan) return;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->pexec) (void)hipGraphExecDestroy(plan->pexec);
  if (plan->pgraph) (void)hipGraphDestroy(plan->pgraph);
  if (plan->cap) (void)hipStreamDestroy(plan->cap);
  if (plan->persist) persist_free(plan->persist);
  for (BatchStage& s : plan->stages) batch_stage_free(&s);
  delete plan;
}

}  // extern "C"

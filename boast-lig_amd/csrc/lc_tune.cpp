// lc_tune.cpp -- autotuner of the 3D libconv kernels.
//
// In the spirit of BOAST's optimizer (GenericOptimization and
// ConvolutionOperator1d#optimize, BigDFT/GenericConvolution2.rb:744-837,
// :1223-1267): the search space is the set of compiled kernel variants of an
// operator form (tile shape, outputs per lane, window step, prefetch depth);
// every variant runs on random data of the caller's shape, in three rounds of
// `repeat` back-to-back launches (steady clocks); the minimum round average
// is kept (as the reference keeps the min of its repeats), and
// the launches rotate over enough input / output sets to exceed the 256 MB
// last-level cache (MALL) of the chip, so that a variant is timed with its
// operands in HBM, as in an application, and
// the fastest variant is recorded for that exact shape.  Later calls of the
// same shape use it.  The table persists as JSON, one entry per line, and
// LIBCONV_TUNE_FILE names a file loaded at first use.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lc_kernels.hpp"

namespace lc {

namespace {

struct Entry {
  int prec, form, id;
  Box3 b;
  int variant;
  double ms;
};

constexpr long kMallBytes = 256L << 20;  // last-level cache of the chip
constexpr int kTuneSetsMax = 64;

std::mutex g_mu;
std::map<std::string, Entry> g_table;
bool g_env_loaded = false;

std::string key_of(int prec, int form, int id, const Box3& b) {
  char k[256];
  snprintf(k, sizeof k, "%d %d %d | %d %d %d | %d %d %d | %d %d %d | %d %d %d", prec, form, id, b.n[0],
           b.n[1], b.n[2], b.bc[0], b.bc[1], b.bc[2], b.nx[0], b.nx[1], b.nx[2], b.ny[0], b.ny[1], b.ny[2]);
  return k;
}

// integers after "name": in one JSON line (name must be unique in the line)
bool get_ints(const char* line, const char* name, int* out, int n) {
  char pat[32];
  snprintf(pat, sizeof pat, "\"%s\":", name);
  const char* p = strstr(line, pat);
  if (!p) return false;
  p += strlen(pat);
  for (int i = 0; i < n; ++i) {
    while (*p && (*p == ' ' || *p == '[' || *p == ',')) ++p;
    char* end;
    const long v = strtol(p, &end, 10);
    if (end == p) return false;
    out[i] = (int)v;
    p = end;
  }
  return true;
}

int load_locked(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  int n = 0;
  while (fgets(line, sizeof line, f)) {
    Entry e{};
    if (!get_ints(line, "prec", &e.prec, 1) || !get_ints(line, "form", &e.form, 1) ||
        !get_ints(line, "op", &e.id, 1) || !get_ints(line, "n", e.b.n, 3) ||
        !get_ints(line, "bc", e.b.bc, 3) || !get_ints(line, "nx", e.b.nx, 3) ||
        !get_ints(line, "ny", e.b.ny, 3) || !get_ints(line, "variant", &e.variant, 1))
      continue;
    if (e.variant < 0 || e.variant >= variant_count(e.form)) continue;
    const char* m = strstr(line, "\"ms\":");
    e.ms = m ? atof(m + 5) : 0.0;
    g_table[key_of(e.prec, e.form, e.id, e.b)] = e;
    ++n;
  }
  fclose(f);
  return n;
}

void ensure_env_loaded() {
  if (g_env_loaded) return;
  g_env_loaded = true;
  const char* p = getenv("LIBCONV_TUNE_FILE");
  if (p && *p) (void)load_locked(p);
}

__global__ void fill_kernel(double* d, float* f, long n, unsigned seed) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + seed;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // U[-1, 1)
  if (d) d[i] = u; else f[i] = (float)u;
}

template <typename R>
hipError_t fill(R* p, long n, unsigned seed, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (sizeof(R) == 8)
    fill_kernel<<<nb, 256, 0, st>>>(reinterpret_cast<double*>(p), nullptr, n, seed);
  else
    fill_kernel<<<nb, 256, 0, st>>>(nullptr, reinterpret_cast<float*>(p), n, seed);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_variant(int form, int id, int v, const Box3& b, const R* x, R* y, double* dot,
                          double* sc, size_t sb, hipStream_t st) {
  switch (form) {
    case FORM_SUM: {
      const double a[3] = {-5.5, -5.5, -5.5};
      return sum3d_fused_v<R>(v, id, b, 1, x, y, a, R(0), R(0), dot, st, sc, sb);
    }
    case FORM_SEP: return sep3d_fused_v<R>(v, id, b, 1, x, y, R(1), R(0), st);
    case FORM_DWT: return sepw3d_v<R>(v, KDWT, id, b, 1, x, y, R(1), R(0), st);
    case FORM_IDWT: return sepw3d_v<R>(v, KIDWT, id, b, 1, x, y, R(1), R(0), st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

int tuned_variant(int prec, int form, int id, const Box3& b) {
  std::lock_guard<std::mutex> g(g_mu);
  ensure_env_loaded();
  auto it = g_table.find(key_of(prec, form, id, b));
  return it == g_table.end() ? -1 : it->second.variant;
}

// Times every variant of (form, id) on random data of shape b; records and
// returns the fastest.  times[v] receives the min time of variant v in ms
// (infinity when the variant rejected the shape).
template <typename R>
hipError_t tune_md(int form, int id, const Box3& b, int repeat, int* best, double* best_ms,
                   std::vector<double>* times) {
  const bool wav = form == FORM_DWT || form == FORM_IDWT;
  const long mul = wav ? 8 : 1;
  const long nti = mul * b.nx[0] * b.nx[1] * b.nx[2], nto = mul * b.ny[0] * b.ny[1] * b.ny[2];
  // sets of (x, y) rotated so that consecutive launches miss the MALL
  const long set_bytes = (long)sizeof(R) * (nti + nto);
  const int nsets = (int)std::min<long>(kTuneSetsMax, std::max<long>(1, (kMallBytes * 3 + set_bytes - 1) / std::max(1L, set_bytes)));
  R *x = nullptr, *y = nullptr;
  double *dot = nullptr, *sc = nullptr;
  size_t sb = form == FORM_SUM ? sum3d_scratch_bytes<R>(b, 1) : 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&x), sizeof(R) * nti * nsets);
  if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&y), sizeof(R) * nto * nsets);
  if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&dot), sizeof(double) * 4);
  if (err == hipSuccess && sb) err = hipMalloc(reinterpret_cast<void**>(&sc), sb);
  if (err == hipSuccess) err = fill<R>(x, nti * nsets, 12345u, st);
  int k = 0;  // launch counter: set k % nsets
  auto run = [&](int v) {
    const int s = k++ % nsets;
    return launch_variant<R>(form, id, v, b, x + nti * s, y + nto * s, dot, sc, sb, st);
  };
  const int nv = variant_count(form);
  times->assign(nv, std::numeric_limits<double>::infinity());
  int bv = -1;
  double bt = std::numeric_limits<double>::infinity();
  for (int v = 0; err == hipSuccess && v < nv; ++v) {
    if (run(v) != hipSuccess) {
      (void)hipGetLastError();
      continue;  // the variant does not take this shape
    }
    // steady state: `repeat` back-to-back launches between two events give
    // one average; the minimum of three such averages is kept
    double tmin = std::numeric_limits<double>::infinity();
    const int reps = std::max(1, repeat);
    for (int w = 0; w < 2 && err == hipSuccess; ++w) err = run(v);
    for (int r = 0; r < 3 && err == hipSuccess; ++r) {
      err = hipEventRecord(e0, st);
      for (int r2 = 0; r2 < reps && err == hipSuccess; ++r2) err = run(v);
      if (err == hipSuccess) err = hipEventRecord(e1, st);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess) tmin = std::min(tmin, (double)ms / reps);
    }
    (*times)[v] = tmin;
    if (tmin < bt) {
      bt = tmin;
      bv = v;
    }
  }
  if (st) (void)hipStreamSynchronize(st);
  if (x) (void)hipFree(x);
  if (y) (void)hipFree(y);
  if (dot) (void)hipFree(dot);
  if (sc) (void)hipFree(sc);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (err != hipSuccess) return err;
  if (bv < 0) return hipErrorInvalidValue;
  *best = bv;
  *best_ms = bt;
  std::lock_guard<std::mutex> g(g_mu);
  ensure_env_loaded();
  Entry e{(int)sizeof(R), form, id, b, bv, bt};
  g_table[key_of(e.prec, form, id, b)] = e;
  return hipSuccess;
}

template hipError_t tune_md<double>(int, int, const Box3&, int, int*, double*, std::vector<double>*);
template hipError_t tune_md<float>(int, int, const Box3&, int, int*, double*, std::vector<double>*);

int tune_save(const char* path) {
  std::lock_guard<std::mutex> g(g_mu);
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "{\"libconv_tune\": 1, \"entries\": [\n");
  size_t k = 0;
  for (const auto& kv : g_table) {
    const Entry& e = kv.second;
    const std::string name = e.form == FORM_DWT || e.form == FORM_IDWT
                                 ? wav_kernel_name(e.prec, e.form, e.id, e.variant, false)
                                 : fused_kernel_name(e.prec, e.form, e.id, e.variant, false);
    fprintf(f,
            " {\"prec\": %d, \"form\": %d, \"op\": %d, \"n\": [%d, %d, %d], \"bc\": [%d, %d, %d], "
            "\"nx\": [%d, %d, %d], \"ny\": [%d, %d, %d], \"variant\": %d, \"ms\": %.6f, \"kernel\": \"%s\"}%s\n",
            e.prec, e.form, e.id, e.b.n[0], e.b.n[1], e.b.n[2], e.b.bc[0], e.b.bc[1], e.b.bc[2], e.b.nx[0],
            e.b.nx[1], e.b.nx[2], e.b.ny[0], e.b.ny[1], e.b.ny[2], e.variant, e.ms, name.c_str(),
            ++k < g_table.size() ? "," : "");
  }
  fprintf(f, "]}\n");
  return fclose(f) == 0 ? (int)k : -1;
}

int tune_load(const char* path) {
  std::lock_guard<std::mutex> g(g_mu);
  g_env_loaded = true;
  return load_locked(path);
}

void tune_clear() {
  std::lock_guard<std::mutex> g(g_mu);
  g_env_loaded = true;
  g_table.clear();
}

}  // namespace lc

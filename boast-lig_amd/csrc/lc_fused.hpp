// lc_fused.hpp -- shared pieces of the single-pass 3D kernels (internal).
#pragma once

#include <algorithm>
#include <vector>

#include "lc_kernels.hpp"

namespace lc {

// LDS carving: every array starts on a 16-byte boundary
__host__ __device__ constexpr size_t lds_align(size_t b) { return (b + 15) & ~size_t(15); }

struct FDim {   // one dim of a fused 3D operator
  int n, bc;
  int nx, ny;   // storage extents of input / output
  int ox, oy;   // logical index of the first stored input / output element
  int olo, nout;  // first logical output index and output count
};

__device__ __forceinline__ int fsrc(const FDim& g, int j) {
  if (g.bc == PER) return pmod(j, g.n);
  if (g.bc == GROW || g.bc == FREE) return (j >= 0 && j < g.n) ? j : -1;
  const int s = j - g.ox;
  return (s >= 0 && s < g.nx) ? s : -1;  // beyond the stored input: never used
}

// fsrc for j, j + 1, j + 2, ... without a division: `raw` is j wrapped on a
// periodic dim, else j less the first stored index; a z-marching kernel keeps
// one cursor per plane stream in scalar registers
struct ZCursor {
  int raw, lim, n;
  bool per;
  __device__ __forceinline__ ZCursor(const FDim& g, int j)
      : raw(g.bc == PER ? pmod(j, g.n) : j - (g.bc == SHRINK ? g.ox : 0)),
        lim(g.bc == SHRINK ? g.nx : g.n), n(g.n), per(g.bc == PER) {}
  __device__ __forceinline__ int src() const { return (unsigned)raw < (unsigned)lim ? raw : -1; }
  __device__ __forceinline__ void next() { raw = (per && raw + 1 == n) ? 0 : raw + 1; }
};

// One input plane as a raw buffer resource (stride 0, 32-bit records).  A
// buffer load whose byte offset is >= the record count reads zero, so a tile
// value outside the input (zero padding of the grow / shrink conditions) is
// loaded without a branch or a select: its offset is kOffNone, or the whole
// plane is absent (record count 0).  The two-plane bodies use this for their
// prefetch, where a branch or a select on the loaded value would make the
// wave wait for the load at once.  Plane sizes are limited to kPlaneMax bytes.
constexpr int kOffNone = int(0x80000000u);
constexpr int kBufDword3 = 0x00020000;  // raw 32-bit data format of gfx9 buffers

// `count` elements from `base` (count 0: every load reads zero)
template <typename R>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t span_rsrc(const R* base, long count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<R*>(base), (short)0, int(count * (long)sizeof(R)),
                                           kBufDword3);
}

// input plane s2 of sxy elements (s2 < 0: outside the input, reads zero)
template <typename R>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const R* xa, long sxy, int s2) {
  return span_rsrc(xa + (s2 >= 0 ? (long)s2 * sxy : 0), s2 >= 0 ? sxy : 0);
}

// element at byte offset `boff` of the plane; kOffNone reads zero
template <typename R>
__device__ __forceinline__ R plane_ld(__amdgpu_buffer_rsrc_t rs, int boff) {
  if constexpr (sizeof(R) == 8)
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, boff, 0, 0));
  else
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, boff, 0, 0));
}

// store t at byte offset `boff` of the plane (kOffNone, or a record count
// of 0: the store is dropped); nt: streaming (non-temporal) store
template <typename R, bool NT = false>
__device__ __forceinline__ void plane_st(__amdgpu_buffer_rsrc_t rs, int boff, R t) {
  typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
  if constexpr (sizeof(R) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, t), rs, boff, 0, NT ? 2 : 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, t), rs, boff, 0, NT ? 2 : 0);
}

template <typename R>
__host__ __device__ __forceinline__ int plane_boff(int off) {  // element offset (< 0: none)
  return off >= 0 ? off * (int)sizeof(R) : kOffNone;
}

template <typename R>
struct FusedArgs {
  const R* x;
  R* y;
  FDim d0, d1, d2;
  long nti, nto;
  int narr;
  int tiles_x, tiles_y, zchunk;  // zchunk = outputs per block along z
  R a0, a1, a2, a_x, a_y;
  int flags;
  double* part;  // 3 sets of per-block energies (sum3d) or null
  long nparts;   // blocks per array = sets stride
  const R* pot;  // F_POT: potential and density on the output box (one array's
  R* rho;        // offsets, shared by all arrays); rho may be null
};

// the F_POT step on output value t at offset o of an output array
template <typename R>
__device__ __forceinline__ R pot_step(const FusedArgs<R>& p, long o, R t) {
  if (p.rho) p.rho[o] = fmar(t, t, p.rho[o]);
  return p.pot[o] * t;
}

// Block -> (tile, z chunk, array).  Workgroups are dispatched round-robin
// over the 8 XCDs of the chip, each with its own L2: the linear block id is
// permuted so that every XCD receives one contiguous run of tiles, and the
// filter halos that neighbouring tiles share are fetched once per XCD.
__device__ __forceinline__ void block_coords(int tiles_x, int& bx, int& by, int& zc, int& arr) {
  const int tiles = gridDim.x, nz = gridDim.y;
  const int nb = tiles * nz * gridDim.z;
  const int lin = blockIdx.x + tiles * (blockIdx.y + nz * blockIdx.z);
  const int xcd = lin & 7, k = lin >> 3;
  const int per = nb >> 3, rem = nb & 7;
  const int nl = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
  const int tile = nl % tiles;
  const int rest = nl / tiles;
  zc = rest % nz;
  arr = rest / nz;
  bx = tile % tiles_x;
  by = tile / tiles_x;
}

// Batched forms: one launch over the blocks of `count` independent operator
// applications.  first[o] is the first block of application o (first[count]
// = total); the block finds its application by binary search, then its
// (tile, z chunk) inside it.  Applications are ordered tiles-fastest, so the
// XCD-contiguous order of block_coords keeps each application's tiles on
// few XCDs.
__device__ __forceinline__ int batch_item(const int* __restrict__ first, int count, int lin) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= lin) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int xcd_linear() {
  const int nb = gridDim.x, lin = blockIdx.x;
  const int xcd = lin & 7, k = lin >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// ------------------------------------------------------------ host helpers
inline FDim fdim(int n, int bc, int nx, int ny, int lo, int up) {
  FDim g;
  g.n = n; g.bc = bc; g.nx = nx; g.ny = ny;
  g.ox = bc == SHRINK ? lo : 0;
  g.oy = bc == GROW ? -up : 0;
  g.olo = bc == GROW ? -up : 0;
  g.nout = bc == GROW ? n + (up - lo) : n;
  return g;
}

// z chunk: enough blocks for ~2 per CU, chunks no shorter than 4 * filter
// length; a grid too small to give every CU a block that way takes chunks
// down to 2 * filter length (a chunk repeats W - 1 planes of work, which
// costs less than idle CUs)
inline int pick_zchunk(long tiles, int nz, int narr, int w, int ncu) {
  const long per = tiles * std::max(narr, 1);
  long chunks = std::min((2L * ncu + per - 1) / per, (long)std::max(1, nz / (4 * w)));
  if (chunks * per < ncu)
    chunks = std::max(chunks, std::min((ncu + per - 1) / per, (long)std::max(1, nz / (2 * w))));
  chunks = std::max(chunks, 1L);
  return (int)((nz + chunks - 1) / chunks);
}

inline int num_cu() {
  static int g_ncu = 0;
  if (!g_ncu) {
    int dev = 0;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess)
      g_ncu = pr.multiProcessorCount;
    if (g_ncu <= 0) g_ncu = 256;
  }
  return g_ncu;
}

}  // namespace lc

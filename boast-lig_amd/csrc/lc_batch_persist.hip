// lc_batch_persist.hip -- a whole batch plan in one persistent launch.
//
// Every block of every item of the plan becomes a work unit (item, block of
// the item).  Units sit in 8 device work queues, one per XCD: an item goes
// whole to the least loaded queue (largest items first), and each queue
// lists its items by dependency depth, so that a unit only ever waits for
// units of a lower depth.  A persistent grid of 8 k resident blocks (block b
// serves queue b % 8, the XCD the dispatcher places it on) pops units with
// an atomic counter until its queue is empty, then serves the other queues
// in turn until all are empty.
//
// Dependencies come from memory: item j depends on an earlier item i when
// i writes what j reads or writes, or j writes what i reads.  A unit of j
// waits until done[i] equals i's block count (thread 0 spins, bounded: past
// the limit it records an error in the status word and goes on, so every
// wave ends).  Publication (only for items another item waits for; the fence
// writes back the XCD's L2, which costs): every wave waits for its own
// stores (s_waitcnt), the block synchronises, then one agent-scope fence (L2
// write back) precedes the increment of done[item]; the waiting side fences
// (L2 invalidate) after its wait.  The Kinetic items' energies are summed
// after the launch by one small kernel, one block per item, each over the
// item's per-block partials in a fixed order.
//
// Progress: all 8 k blocks are resident (k from the occupancy query), every
// queue is ordered by depth and popped in that order (by its own blocks, and
// by others only after that), and depth-0 units never wait; by induction on
// depth every unit completes: a popped unit waits only for units of lower
// depth, and a queue's own blocks stay on it until it is empty.
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "lc_body_conv.hpp"
#include "lc_body_sepq.hpp"
#include "lc_body_sum2.hpp"
#include "lc_body_wav.hpp"

namespace lc {

enum PBody : int { PB_SEP_MD = 0, PB_SEP_IMD = 1, PB_SUM_D2 = 2, PB_DWT = 3, PB_IDWT = 4 };

struct PItem {
  int body;
  int nblocks;
  int dep_off, dep_cnt;
  int publish;   // a later item waits for it
  double* dot;   // energies of a Kinetic item (3 doubles) or null
};

struct PState {  // device; zeroed before every run
  int next[8];
  int status;
  int pad[7];
};

// The bodies of the persistent kernel, 256 lanes each: the 128-bit-LDS magic
// filter (FP64 16 x 16, FP32 32 x 16 tiles), the two-plane Kinetic form
// (16 x 16, two planes per step, four outputs per task; FP64 with the column
// planes two steps in flight) and the wavelet body on 16 x 8 steps.
constexpr int kPThreads = 256;
template <typename R> struct PShape;
template <> struct PShape<double> {
  static constexpr int MX = 16, MY = 16, MQX = 4, MQY = 2, KPW = 2, WBZ = 1;
};
template <> struct PShape<float> {
  static constexpr int MX = 32, MY = 16, MQX = 8, MQY = 4, KPW = 1, WBZ = 1;
};
template <typename R, class F>
using PSepGeo = SepqGeo<R, F, PShape<R>::MX, PShape<R>::MY, PShape<R>::MQX, PShape<R>::MQY, 2>;
template <typename R>
using PSumGeo = Sum2Geo<R, FD2, 16, 16, 4, 4, 2>;

template <typename R, class F, bool ME>
__device__ __forceinline__ void p_sep(const FusedArgs<R>& p, int bx, int by, int zc, unsigned char* lds) {
  using S = PShape<R>;
  // plain stores (streaming ones measured no faster for the publishing
  // fence's write-back)
  sepq_body<R, F, S::MX, S::MY, S::MQX, S::MQY, 2, ME, false>(p, bx, by, zc, 0, lds);
}
template <typename R, bool ME>
__device__ __forceinline__ void p_sum(const FusedArgs<R>& p, int bx, int by, int zc, long bid, unsigned char* lds) {
  sum3d2_body<R, FD2, 16, 16, 4, 4, 2, 0, PShape<R>::KPW, ME>(p, bx, by, zc, 0, bid, lds);
}
template <typename R, class OP>
__device__ __forceinline__ void p_wav(const FusedArgs<R>& p, int bx, int by, int zc, unsigned char* lds) {
  sepw3d_body<R, OP, 16, 8, 4, PShape<R>::WBZ, PShape<R>::WBZ>(p, bx, by, zc, 0, lds);
}

template <typename R>
constexpr size_t persist_lds_bytes() {
  static_assert(PSepGeo<R, FMD>::NT == kPThreads && PSepGeo<R, FIMD>::NT == kPThreads &&
                    PSumGeo<R>::NT == kPThreads && 2 * 16 * 8 == kPThreads,
                "every body of the persistent kernel takes kPThreads lanes");
  return std::max(std::max(PSepGeo<R, FMD>::lds_bytes(), PSepGeo<R, FIMD>::lds_bytes()),
                  std::max(PSumGeo<R>::lds_bytes(),
                           sepw3d_lds_bytes<R, Dwt<Sym8Dec>, 16, 8, 4, PShape<R>::WBZ, PShape<R>::WBZ>()));
}

__global__ void persist_reset(PState* st, int* done, int nitems) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 16) reinterpret_cast<int*>(st)[i] = 0;
  if (i < nitems) done[i] = 0;
}

// two waves per SIMD (two blocks per CU): the wavelet body takes more than
// 128 VGPRs (capped there it spills)
template <typename R>
__global__ __launch_bounds__(kPThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_persistent(const FusedArgs<R>* __restrict__ tab,
                                                        const PItem* __restrict__ items,
                                                        const int* __restrict__ deps,
                                                        const int2* __restrict__ work,
                                                        const int* __restrict__ qoff, PState* st,
                                                        int* done, int spin_limit) {
  __shared__ __align__(16) unsigned char lds[persist_lds_bytes<R>()];
  __shared__ int s_pos;
  const int t = threadIdx.x;
  // the block's own queue first; once that is empty, the other queues in
  // turn (in their order), so that no XCD idles while another has units
  // left; the block leaves when it has found all 8 empty
  int queue = blockIdx.x & 7, empty = 0;
  for (;;) {
    __syncthreads();  // the previous unit is done with LDS and s_pos
    if (t == 0) s_pos = atomicAdd(&st->next[queue], 1);
    __syncthreads();
    const int pos = __builtin_amdgcn_readfirstlane(qoff[queue] + s_pos);
    if (pos >= qoff[queue + 1]) {  // uniform
      if (++empty == 8) break;     // every wave of the block leaves
      queue = (queue + 1) & 7;
      continue;
    }
    const int2 wk = work[pos];
    const int it = __builtin_amdgcn_readfirstlane(wk.x), local = __builtin_amdgcn_readfirstlane(wk.y);
    const PItem pi = items[it];
    if (t == 0 && pi.dep_cnt > 0) {
      for (int k = 0; k < pi.dep_cnt; ++k) {
        const int d = deps[pi.dep_off + k];
        const int need = items[d].nblocks;
        int spins = 0;
        while (__hip_atomic_load(&done[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          __builtin_amdgcn_s_sleep(8);
          if (++spins > spin_limit) {
            atomicOr(&st->status, 1);
            break;
          }
        }
      }
      // acquire only (an L2 invalidate): a full fence would also write
      // back this XCD's L2
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    const FusedArgs<R> p = tab[it];
    const int tiles = p.tiles_x * p.tiles_y;
    const int tile = local % tiles, zc = local / tiles;
    const int bx = tile % p.tiles_x, by = tile / p.tiles_x;
    const bool me = (p.flags & (F_AY | F_POT)) != 0;  // uniform
    switch (pi.body) {
      case PB_SEP_MD:
        if (me) p_sep<R, FMD, true>(p, bx, by, zc, lds); else p_sep<R, FMD, false>(p, bx, by, zc, lds);
        break;
      case PB_SEP_IMD:
        if (me) p_sep<R, FIMD, true>(p, bx, by, zc, lds); else p_sep<R, FIMD, false>(p, bx, by, zc, lds);
        break;
      case PB_SUM_D2:
        if (me) p_sum<R, true>(p, bx, by, zc, (long)local, lds); else p_sum<R, false>(p, bx, by, zc, (long)local, lds);
        break;
      case PB_DWT: p_wav<R, Dwt<Sym8Dec>>(p, bx, by, zc, lds); break;
      case PB_IDWT: p_wav<R, Idwt<Sym8Rec>>(p, bx, by, zc, lds); break;
    }
    // publish: this wave's stores complete, then the block, then one release
    // fence (an L2 write-back; a full fence would also invalidate the XCD's
    // L2 under every other block running there).  Units nobody waits for
    // skip it (their stores are complete when the kernel ends).
    if (!pi.publish) continue;  // uniform
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicAdd(&done[it], 1);
    }
  }
}

// The energies of the Kinetic items, after the persistent launch: one block
// per item sums its per-block partials in a fixed order.
struct PEnergy {
  const double* part;  // 3 sets of nb partials
  long nb;
  double* dot;
};

__global__ __launch_bounds__(256) void persist_energies(const PEnergy* __restrict__ en) {
  __shared__ double s_red[4];
  const PEnergy e = en[blockIdx.x];
  const int t = threadIdx.x;
  for (int k = 0; k < 3; ++k) {
    double v = 0.0;
    for (long i = t; i < e.nb; i += 256) v += e.part[k * e.nb + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) s_red[t >> 6] = v;
    __syncthreads();
    if (t == 0) e.dot[k] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
    __syncthreads();
  }
}

// ====================================================================== host
struct PersistPlan {
  int prec = 8;
  int nitems = 0, units = 0, grid = 0;
  void* tab = nullptr;
  PItem* items = nullptr;
  int* deps = nullptr;
  int2* work = nullptr;
  int* qoff = nullptr;
  PState* st = nullptr;
  int* done = nullptr;
  double* part = nullptr;
  PEnergy* en = nullptr;  // the Kinetic items with energies
  int nen = 0;
};

void persist_free(PersistPlan* pp) {
  if (!pp) return;
  for (void* q : {pp->tab, (void*)pp->items, (void*)pp->deps, (void*)pp->work, (void*)pp->qoff,
                  (void*)pp->st, (void*)pp->done, (void*)pp->part, (void*)pp->en})
    if (q) (void)hipFree(q);
  delete pp;
}

namespace {

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& v) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * std::max<size_t>(v.size(), 1));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

struct Range {
  uintptr_t lo, hi;
};
bool overlap(const Range& a, const Range& b) { return a.lo < b.hi && b.lo < a.hi; }

template <typename R, class F, int TX, int TY>
FusedArgs<R> conv_args(const BatchItem<R>& it, bool sum) {
  FusedArgs<R> p{};
  const Box3& b = it.b;
  p.x = it.x; p.y = it.y;
  p.d0 = fdim(b.n[0], b.bc[0], b.nx[0], b.ny[0], F::LO, F::UP);
  p.d1 = fdim(b.n[1], b.bc[1], b.nx[1], b.ny[1], F::LO, F::UP);
  p.d2 = fdim(b.n[2], b.bc[2], b.nx[2], b.ny[2], F::LO, F::UP);
  p.nti = (long)b.nx[0] * b.nx[1] * b.nx[2];
  p.nto = (long)b.ny[0] * b.ny[1] * b.ny[2];
  p.narr = 1;
  p.tiles_x = (p.d0.nout + TX - 1) / TX;
  p.tiles_y = (p.d1.nout + TY - 1) / TY;
  p.a0 = (R)it.a[0]; p.a1 = (R)it.a[1]; p.a2 = (R)it.a[2];
  p.a_x = (R)it.a_x; p.a_y = (R)it.a_y;
  if (sum)
    p.flags = (p.a_x != R(0) ? F_AX : 0) | (p.a_y != R(0) ? F_AY : 0) | (it.dot ? F_DOT : 0);
  else
    p.flags = (p.a0 != R(1) ? F_A : 0) | (p.a_y != R(0) ? F_AY : 0);
  p.zchunk = p.d2.nout;  // one chunk: the queue balances the load
  return p;
}

template <typename R, class Op>
FusedArgs<R> wav_args_p(const BatchItem<R>& it) {
  FusedArgs<R> p = conv_args<R, Op, 16, 8>(it, false);  // same geometry rules, per step
  p.nti = 8L * it.b.nx[0] * it.b.nx[1] * it.b.nx[2];
  p.nto = 8L * it.b.ny[0] * it.b.ny[1] * it.b.ny[2];
  return p;
}

}  // namespace

// kinds[i]: BK_SEP / BK_SUM / BK_DWT / BK_IDWT; ids[i] the operator
template <typename R>
hipError_t persist_build(const BatchItem<R>* it, const int* kinds, const int* ids, int count,
                         PersistPlan** out) {
  *out = nullptr;
  std::vector<FusedArgs<R>> tab(count);
  std::vector<PItem> items(count);
  std::vector<Range> rin(count), rout(count);
  const size_t mul_c = sizeof(R);
  long total_parts = 0;
  for (int i = 0; i < count; ++i) {
    const int k = kinds[i], id = ids[i];
    PItem& pi = items[i];
    pi.dot = nullptr;
    if (k == BK_SEP && (id == OP_MD || id == OP_IMD) && plane_fits(it[i].b, sizeof(R))) {
      constexpr int MX = PShape<R>::MX, MY = PShape<R>::MY;
      pi.body = id == OP_MD ? PB_SEP_MD : PB_SEP_IMD;
      tab[i] = id == OP_MD ? conv_args<R, FMD, MX, MY>(it[i], false) : conv_args<R, FIMD, MX, MY>(it[i], false);
    } else if (k == BK_SUM && id == OP_D2 && plane_fits(it[i].b, sizeof(R))) {
      pi.body = PB_SUM_D2;
      tab[i] = conv_args<R, FD2, 16, 16>(it[i], true);
      pi.dot = it[i].dot;
    } else if ((k == BK_DWT || k == BK_IDWT) && id == WAV_SYM8 && wav_plane_fits(it[i].b, sizeof(R))) {
      pi.body = k == BK_DWT ? PB_DWT : PB_IDWT;
      tab[i] = k == BK_DWT ? wav_args_p<R, Dwt<Sym8Dec>>(it[i]) : wav_args_p<R, Idwt<Sym8Rec>>(it[i]);
    } else {
      return hipErrorNotSupported;
    }
    const FusedArgs<R>& p = tab[i];
    pi.nblocks = p.tiles_x * p.tiles_y * ((p.d2.nout + p.zchunk - 1) / p.zchunk);
    if (pi.body == PB_SUM_D2 && pi.dot) total_parts += 3L * pi.nblocks;
    rin[i] = {(uintptr_t)it[i].x, (uintptr_t)it[i].x + mul_c * (uintptr_t)p.nti};
    rout[i] = {(uintptr_t)it[i].y, (uintptr_t)it[i].y + mul_c * (uintptr_t)p.nto};
  }
  // dependencies (earlier items only) and depth
  std::vector<int> deps, depth(count, 0);
  for (int j = 0; j < count; ++j) {
    items[j].dep_off = (int)deps.size();
    for (int i = 0; i < j; ++i)
      if (overlap(rout[i], rin[j]) || overlap(rout[i], rout[j]) || overlap(rin[i], rout[j])) {
        deps.push_back(i);
        depth[j] = std::max(depth[j], depth[i] + 1);
        items[i].publish = 1;
      }
    items[j].dep_cnt = (int)deps.size() - items[j].dep_off;
  }
  // items to queues: by depth, largest first, to the least loaded queue
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (depth[a] != depth[b]) return depth[a] < depth[b];
    return items[a].nblocks > items[b].nblocks;
  });
  std::vector<std::vector<int2>> q(8);
  std::vector<long> load(8, 0);
  long units = 0;
  for (int i : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (int b = 0; b < items[i].nblocks; ++b) q[x].push_back(make_int2(i, b));
    load[x] += items[i].nblocks;
    units += items[i].nblocks;
  }
  if (units > 0x7fffffffL) return hipErrorInvalidValue;
  std::vector<int2> work;
  std::vector<int> qoff(9, 0);
  for (int x = 0; x < 8; ++x) {
    qoff[x] = (int)work.size();
    work.insert(work.end(), q[x].begin(), q[x].end());
  }
  qoff[8] = (int)work.size();
  PersistPlan* pp = new PersistPlan;
  pp->prec = sizeof(R);
  pp->nitems = count;
  pp->units = (int)units;
  hipError_t e = hipSuccess;
  if (total_parts > 0) e = hipMalloc(reinterpret_cast<void**>(&pp->part), sizeof(double) * total_parts);
  long off = 0;
  std::vector<PEnergy> en;
  for (int i = 0; i < count && e == hipSuccess; ++i)
    if (items[i].body == PB_SUM_D2 && items[i].dot) {
      tab[i].part = pp->part + off;
      tab[i].nparts = items[i].nblocks;
      en.push_back({pp->part + off, (long)items[i].nblocks, items[i].dot});
      off += 3L * items[i].nblocks;
    }
  pp->nen = (int)en.size();
  if (e == hipSuccess && pp->nen) e = upload(&pp->en, en);
  if (e == hipSuccess) e = upload(reinterpret_cast<FusedArgs<R>**>(&pp->tab), tab);
  if (e == hipSuccess) e = upload(&pp->items, items);
  if (e == hipSuccess) e = upload(&pp->deps, deps);
  if (e == hipSuccess) e = upload(&pp->work, work);
  if (e == hipSuccess) e = upload(&pp->qoff, qoff);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&pp->st), sizeof(PState));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&pp->done), sizeof(int) * std::max(count, 1));
  int per_cu = 0;
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_persistent<R>, kPThreads, 0);
  if (e == hipSuccess && per_cu < 1) e = hipErrorInvalidConfiguration;
  if (e != hipSuccess) {
    persist_free(pp);
    return e;
  }
  // 8 k blocks, all resident: k per XCD = blocks per CU x CUs per XCD
  pp->grid = 8 * std::max(1, per_cu * (num_cu() / 8));
  *out = pp;
  return hipSuccess;
}

hipError_t persist_launch(const PersistPlan* pp, hipStream_t st, int spin_limit) {
  const int n = std::max(16, pp->nitems);
  persist_reset<<<(n + 255) / 256, 256, 0, st>>>(pp->st, pp->done, pp->nitems);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || pp->units == 0) return e;
  if (pp->prec == 8)
    batch_persistent<double><<<pp->grid, kPThreads, 0, st>>>(static_cast<const FusedArgs<double>*>(pp->tab),
                                                             pp->items, pp->deps, pp->work, pp->qoff, pp->st,
                                                             pp->done, spin_limit);
  else
    batch_persistent<float><<<pp->grid, kPThreads, 0, st>>>(static_cast<const FusedArgs<float>*>(pp->tab),
                                                            pp->items, pp->deps, pp->work, pp->qoff, pp->st,
                                                            pp->done, spin_limit);
  e = hipGetLastError();
  if (e != hipSuccess || pp->nen == 0) return e;
  persist_energies<<<pp->nen, 256, 0, st>>>(pp->en);
  return hipGetLastError();
}

hipError_t persist_status(const PersistPlan* pp, int* status) {
  PState h;
  hipError_t e = hipMemcpy(&h, pp->st, sizeof h, hipMemcpyDeviceToHost);
  if (e == hipSuccess) *status = h.status;
  return e;
}

int persist_grid(const PersistPlan* pp) { return pp->grid; }

template hipError_t persist_build<double>(const BatchItem<double>*, const int*, const int*, int, PersistPlan**);
template hipError_t persist_build<float>(const BatchItem<float>*, const int*, const int*, int, PersistPlan**);

}  // namespace lc

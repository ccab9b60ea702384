// batch.hip -- many independent grids ("shots") advanced together.
//
// The reference runs one grid per process (main.cpp); a survey of many
// shots of uneven size is the production use of the leapfrog sweep.  Two
// ways to run a time step over all shots:
//  * persistent: ONE launch per step.  A grid sized to the resident capacity
//    of the GPU pulls work items -- (shot, k-tile, j-tile, i-chunk) -- from a
//    queue (an atomic ticket counter), largest items first, and runs each with
//    the same tile engine as the single-grid sweep (sweep_tile).
//  * sequential: one plain sweep launch per shot per step.
//  * two to five steps per pass: one persistent launch per pass, items of
//    the fused sweep (sweep2_tile) of every shot from a queue.
// Either way the time loop can be captured into one hipGraph per call.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "sth_internal.h"
#include "sweep2_dispatch.h"
#include "sweep2_kernel.h"
#include "sweep_dispatch.h"

using namespace sth;
using namespace sth::detail;

namespace sth {

struct BatchItem {
  int shot, tk, tj, ch;
};

struct BatchArgs {
  const SweepArgs *shots;  // per-shot args: u = buffer a, v = buffer b
  const BatchItem *items;  // sorted work list
  int total_items;
  unsigned *ticket;        // queue counter
  unsigned ticket_base;    // tickets taken by earlier steps since the last reset
  int parity;              // 0: u = a, v = b; 1: u = b, v = a
  float c0, c1, c2n;
};

// Every block takes tickets until it draws one past the end, so a step
// consumes exactly total_items + gridDim.x tickets: step n of a launch
// sequence starts at ticket_base = n * (total_items + grid).
template <int NPTS, int MODE, int FMT, int MJ, int NW, int V>
__global__ __launch_bounds__(64 * NW) void batch_kernel(const BatchArgs b) {
  __shared__ __attribute__((aligned(16))) typename Codec<FMT>::A lds[2][TileLds<NPTS, MJ, NW, V>::SIZE];
  __shared__ float red[NW];
  __shared__ int sh_item;
  for (;;) {
    if (threadIdx.x == 0) sh_item = (int)(atomicAdd(b.ticket, 1u) - b.ticket_base);
    __syncthreads();
    const int item = sh_item;
    __syncthreads();
    if (item >= b.total_items) break;
    const BatchItem it = b.items[item];
    SweepArgs a = b.shots[it.shot];
    if (b.parity) {
      const void *t = a.u;
      a.u = a.v;
      a.v = const_cast<void *>(t);
    }
    a.c0 = b.c0;
    a.c1 = b.c1;
    a.c2n = b.c2n;
    const int ib = it.ch * a.chunk_len;
    const int ie = min(a.xs, ib + a.chunk_len);
    sweep_tile<NPTS, MODE, FMT, MJ, NW, V>(a, it.tk * 64 * V, it.tj * NW * MJ, ib, ie, lds, red);
    __syncthreads();  // LDS is reused by the next item
  }
}

template <int NPTS>
const void *batch_fn(int mj, int nw, int vec) {
#define STH_BATCH_FN(MJ, NW, VV)                 \
  if (mj == MJ && nw == NW && vec == VV)         \
    return (const void *)&batch_kernel<NPTS, MODE_LEAP, FMT_F32, MJ, NW, VV>;
  STH_TILE_LIST(STH_BATCH_FN)
#undef STH_BATCH_FN
  return nullptr;
}

template <int NPTS>
hipError_t launch_batch_n(int mj, int nw, int vec, unsigned grid, hipStream_t st,
                          const BatchArgs &b) {
#define STH_BATCH_CASE(MJ, NW, VV)                                                     \
  if (mj == MJ && nw == NW && vec == VV) {                                             \
    hipLaunchKernelGGL((batch_kernel<NPTS, MODE_LEAP, FMT_F32, MJ, NW, VV>), dim3(grid), \
                       dim3(64 * NW), 0, st, b);                                       \
    return hipGetLastError();                                                          \
  }
  STH_TILE_LIST(STH_BATCH_CASE)
#undef STH_BATCH_CASE
  return hipErrorInvalidValue;
}

}  // namespace sth

struct sth_batch {
  std::vector<sth_ctx *> shots;
  uint32_t npts = 0, dtype = 0;
  int device = 0, cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evs = nullptr;
  sth_tuning tune{};
  uint32_t planes = 0;       // i-planes per work item (0: whole x-range)
  uint32_t grid_blocks = 0;  // persistent grid (0: resident capacity)
  // device tables for the persistent kernel (rebuilt when the tuning changes)
  SweepArgs *shot_args = nullptr;
  BatchItem *items = nullptr;
  unsigned *ticket = nullptr;
  int total_items = 0;
  unsigned ticket_base = 0;
  bool tables_ok = false;
  // device tables of the fused-pass kernel (rebuilt on every call:
  // the shots' buffers trade places with their twins after each pass)
  Batch2Shot *shot2 = nullptr;
  Batch2Item *items2 = nullptr;
  size_t shot2_cap = 0, items2_cap = 0;
  int total_items2 = 0;
  std::vector<hipGraphExec_t> execs;  // the graphs of the last graph-mode call
  uint32_t best_method = STH_BATCH_PERSISTENT;  // the autotuner's fastest method
  bool timing = false;
  double timed_ms = 0.0;
  uint64_t launches = 0;
};

namespace {

// Kernel launches per captured graph (see sth_batch_leapfrog).
constexpr uint32_t kGraphLaunches = 1024;

hipError_t launch_batch(uint32_t npts, int mj, int nw, int vec, unsigned grid, hipStream_t st,
                        const BatchArgs &b) {
  if (npts == 7) return launch_batch_n<7>(mj, nw, vec, grid, st, b);
  if (npts == 13) return launch_batch_n<13>(mj, nw, vec, grid, st, b);
  if (npts == 25) return launch_batch_n<25>(mj, nw, vec, grid, st, b);
  return hipErrorInvalidValue;
}

const void *batch_kernel_fn(uint32_t npts, int mj, int nw, int vec) {
  if (npts == 7) return batch_fn<7>(mj, nw, vec);
  if (npts == 13) return batch_fn<13>(mj, nw, vec);
  if (npts == 25) return batch_fn<25>(mj, nw, vec);
  return nullptr;
}

struct TileSel {
  int mj, nw, vec;
};
TileSel tile_of(const sth_tuning &t) {
  return {t.rows_per_wave ? (int)t.rows_per_wave : kDefaultRows,
          t.waves ? (int)t.waves : kDefaultWaves, t.vec ? (int)t.vec : kDefaultVec};
}

#define STH_TRY(expr)                            \
  do {                                           \
    hipError_t e_ = (expr);                      \
    if (e_ != hipSuccess) return hip_status(e_); \
  } while (0)

// Per-shot plan for the batch tuning: the tile of the batch, i-chunks of
// `planes` planes.
Plan shot_plan(const sth_batch *b, const sth_ctx *c) {
  sth_tuning t = b->tune;
  // the persistent kernel's tile (tile_of): name it, so the plan does not
  // take the single-grid size-dependent default
  const TileSel ts = tile_of(b->tune);
  t.rows_per_wave = (uint32_t)ts.mj;
  t.waves = (uint32_t)ts.nw;
  t.vec = (uint32_t)ts.vec;
  t.launch = 1;  // one workgroup per tile (the batch has its own work queue)
  const int xs = (int)c->g.xs;
  const int planes = b->planes ? (int)b->planes : xs;
  t.chunks = (uint32_t)std::min(64, std::max(1, (xs + planes - 1) / planes));
  return make_plan(c, t);
}

unsigned persistent_grid(const sth_batch *b, const TileSel &ts) {
  if (b->grid_blocks) return b->grid_blocks;
  int per_cu = 0;
  const int nw = ts.nw;
  const void *fn = batch_kernel_fn(b->npts, ts.mj, ts.nw, ts.vec);
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * nw, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (unsigned)(per_cu * b->cus);
}

int build_tables(sth_batch *b) {
  if (b->tables_ok) return STH_OK;
  std::vector<SweepArgs> args(b->shots.size());
  std::vector<BatchItem> items;
  std::vector<int> len;
  for (size_t s = 0; s < b->shots.size(); s++) {
    sth_ctx *c = b->shots[s];
    const Plan p = shot_plan(b, c);
    SweepArgs &a = args[s];
    memset(&a, 0, sizeof(a));
    const int ub = c->ubuf;
    c->ubuf = 0;  // table holds u = buffer a, v = buffer b; the kernel swaps by parity
    base_args(c, p, a);
    c->ubuf = ub;
    fill_weights(c, true, a);
    a.remap = 0;
    for (int ch = 0; ch < p.chunks; ch++)
      for (int tj = 0; tj < p.ntj; tj++)
        for (int tk = 0; tk < p.ntk; tk++) {
          items.push_back({(int)s, tk, tj, ch});
          len.push_back(std::min((int)c->g.xs, (ch + 1) * p.chunk_len) - ch * p.chunk_len);
        }
  }
  // largest items first (stable: items of one shot stay together)
  std::vector<int> order(items.size());
  for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return len[x] > len[y]; });
  std::vector<BatchItem> sorted(items.size());
  for (size_t q = 0; q < order.size(); q++) sorted[q] = items[order[q]];
  if (b->shot_args) (void)hipFree(b->shot_args);
  if (b->items) (void)hipFree(b->items);
  b->shot_args = nullptr;
  b->items = nullptr;
  STH_TRY(hipMalloc((void **)&b->shot_args, args.size() * sizeof(SweepArgs)));
  STH_TRY(hipMalloc((void **)&b->items, std::max<size_t>(1, sorted.size()) * sizeof(BatchItem)));
  STH_TRY(hipMemcpyAsync(b->shot_args, args.data(), args.size() * sizeof(SweepArgs),
                         hipMemcpyHostToDevice, b->stream));
  STH_TRY(hipMemcpyAsync(b->items, sorted.data(), sorted.size() * sizeof(BatchItem),
                         hipMemcpyHostToDevice, b->stream));
  STH_TRY(hipStreamSynchronize(b->stream));
  b->total_items = (int)sorted.size();
  b->tables_ok = true;
  return STH_OK;
}

// The fused kernel's shape for S steps per pass: the stencil's default
// (rows per wave, waves).
void shape2(uint32_t npts, int S, int &mj, int &nw) {
  mj = S >= 4 ? 2 : npts == 25 || S == 3 ? 3 : 4;
  nw = 16;
}

unsigned persistent_grid2(const sth_batch *b, int mj, int nw, int S) {
  int per_cu = 0;
  const void *fn = batch2_fn((int)b->npts, mj, nw, S);
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * nw, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (unsigned)(per_cu * b->cus);
}

// Work list of a pass of S steps over all shots, from the shots' current
// buffers (the twins must exist).  i-chunks per shot: the tuning's chunk
// count when set, else as many as give about six rounds of the persistent
// grid (at least 8R planes per chunk: each chunk recomputes (S-1)(2R+1)
// planes).
int build_tables2(sth_batch *b, int mj, int nw, int S) {
  const size_t ns = b->shots.size();
  sth_tuning t = b->tune;
  t.rows_per_wave = (uint32_t)mj;
  t.waves = (uint32_t)nw;
  t.steps_per_pass = (uint32_t)S;
  int64_t tiles = 0;
  for (sth_ctx *c : b->shots) {
    t.chunks = 1;
    const Plan2 p = make_plan2(c, t);
    tiles += (int64_t)p.ntk * p.ntj;
  }
  const int64_t slots = persistent_grid2(b, mj, nw, S);
  const int target = (int)std::max<int64_t>(1, (6 * slots + tiles - 1) / std::max<int64_t>(1, tiles));
  std::vector<Batch2Shot> shots(ns);
  std::vector<Batch2Item> items;
  std::vector<int> len;
  for (size_t q = 0; q < ns; q++) {
    sth_ctx *c = b->shots[q];
    const int xs = (int)c->g.xs;
    const int most = std::max(1, xs / (8 * c->R));
    // the batch tuning's chunk count, else about six rounds (planes per
    // item is the one-step persistent method's knob: tuned for that kernel,
    // it cost the fused passes 8-12 %)
    t.chunks = (uint32_t)std::min(64, b->tune.chunks ? (int)b->tune.chunks
                                                     : std::min(target, most));
    const Plan2 p = make_plan2(c, t);
    const size_t o = c->L.origin * c->eb;
    Batch2Shot &sh = shots[q];
    sh.buf[0] = (const char *)c->base[0] + o;
    sh.buf[1] = (const char *)c->base[1] + o;
    sh.buf[2] = (const char *)c->twin[0] + o;
    sh.buf[3] = (const char *)c->twin[1] + o;
    sh.s = (const char *)c->base[2] + o;
    sh.pitch = (int64_t)c->L.pitch;
    sh.plane = (int64_t)c->L.plane;
    sh.xs = xs;
    sh.ys = (int)c->g.ys;
    sh.zs = (int)c->g.zs;
    sh.ntk = p.ntk;
    sh.ntj = p.ntj;
    sh.chunk_len = p.chunk_len;
    for (int ch = 0; ch < p.chunks; ch++)
      for (int tj = 0; tj < p.ntj; tj++)
        for (int tk = 0; tk < p.ntk; tk++) {
          items.push_back({(int)q, tk, tj, ch});
          len.push_back(std::min(xs, (ch + 1) * p.chunk_len) - ch * p.chunk_len);
        }
  }
  std::vector<int> order(items.size());
  for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return len[x] > len[y]; });
  std::vector<Batch2Item> sorted(items.size());
  for (size_t q = 0; q < order.size(); q++) sorted[q] = items[order[q]];
  if (b->shot2_cap < ns) {
    if (b->shot2) (void)hipFree(b->shot2);
    b->shot2 = nullptr;
    b->shot2_cap = 0;
    STH_TRY(hipMalloc((void **)&b->shot2, ns * sizeof(Batch2Shot)));
    b->shot2_cap = ns;
  }
  if (b->items2_cap < sorted.size()) {
    if (b->items2) (void)hipFree(b->items2);
    b->items2 = nullptr;
    b->items2_cap = 0;
    STH_TRY(hipMalloc((void **)&b->items2, sorted.size() * sizeof(Batch2Item)));
    b->items2_cap = sorted.size();
  }
  STH_TRY(hipMemcpyAsync(b->shot2, shots.data(), ns * sizeof(Batch2Shot), hipMemcpyHostToDevice,
                         b->stream));
  STH_TRY(hipMemcpyAsync(b->items2, sorted.data(), sorted.size() * sizeof(Batch2Item),
                         hipMemcpyHostToDevice, b->stream));
  STH_TRY(hipStreamSynchronize(b->stream));
  b->total_items2 = (int)sorted.size();
  return STH_OK;
}

// Queues `passes` passes of S steps (step S*q+r uses c2[S*q+r]); u starts
// in buffer ub0.
int enqueue_passes(sth_batch *b, int S, uint32_t passes, float c0, float c1, const float *c2,
                   int ub0, uint64_t *launches) {
  int mj, nw;
  shape2(b->npts, S, mj, nw);
  const unsigned G = persistent_grid2(b, mj, nw, S);
  const unsigned per_launch = (unsigned)b->total_items2 + G;
  if ((uint64_t)b->ticket_base + (uint64_t)per_launch * passes >= 0xF0000000ull) {
    STH_TRY(launch_zero_u32(b->ticket, 1, b->stream));
    b->ticket_base = 0;
  }
  SweepArgs wts;
  memset(&wts, 0, sizeof(wts));
  fill_weights(b->shots[0], true, wts);
  Batch2Args a;
  memset(&a, 0, sizeof(a));
  a.shots = b->shot2;
  a.items = b->items2;
  a.total_items = b->total_items2;
  a.ticket = b->ticket;
  a.c0 = c0;
  a.c1 = c1;
  memcpy(a.w, wts.w, sizeof(a.w));
  for (uint32_t q = 0; q < passes; q++) {
    a.ticket_base = b->ticket_base;
    a.ub = (ub0 + S * (int)q) & 1;
    a.phase = (int)(q & 1);
    const float *cq = c2 ? c2 + (size_t)S * q : nullptr;
    a.c2a = cq ? cq[0] : 0.0f;
    a.c2b = cq ? cq[1] : 0.0f;
    a.c2c = cq && S >= 3 ? cq[2] : 0.0f;
    a.c2d = cq && S >= 4 ? cq[3] : 0.0f;
    a.c2e = cq && S >= 5 ? cq[4] : 0.0f;
    STH_TRY(launch_batch2((int)b->npts, mj, nw, S, G, b->stream, a));
    b->ticket_base += per_launch;
    (*launches)++;
  }
  return STH_OK;
}

// Orders the batch stream after all work queued on the shots' streams.
int join_shots(sth_batch *b) {
  for (sth_ctx *c : b->shots) {
    STH_TRY(hipEventRecord(b->evs, c->stream));
    STH_TRY(hipStreamWaitEvent(b->stream, b->evs, 0));
  }
  return STH_OK;
}

// Orders every shot's stream after the work queued on the batch stream.
int fork_shots(sth_batch *b) {
  STH_TRY(hipEventRecord(b->evs, b->stream));
  for (sth_ctx *c : b->shots) STH_TRY(hipStreamWaitEvent(c->stream, b->evs, 0));
  return STH_OK;
}

// Queues `steps` steps on the batch stream (eagerly or while capturing).
int enqueue_steps(sth_batch *b, uint32_t steps, float c0, float c1, const float *c2, int method,
                  int parity0, uint64_t *launches) {
  const TileSel ts = tile_of(b->tune);
  if (method == STH_BATCH_PERSISTENT) {
    const unsigned G = persistent_grid(b, ts);
    const unsigned per_step = (unsigned)b->total_items + G;
    if ((uint64_t)b->ticket_base + (uint64_t)per_step * steps >= 0xF0000000ull) {
      STH_TRY(launch_zero_u32(b->ticket, 1, b->stream));
      b->ticket_base = 0;
    }
    BatchArgs a;
    a.shots = b->shot_args;
    a.items = b->items;
    a.total_items = b->total_items;
    a.ticket = b->ticket;
    a.c0 = c0;
    a.c1 = c1;
    for (uint32_t n = 0; n < steps; n++) {
      a.ticket_base = b->ticket_base;
      a.parity = (parity0 + (int)n) & 1;
      a.c2n = c2 ? c2[n] : 0.0f;
      STH_TRY(launch_batch(b->npts, ts.mj, ts.nw, ts.vec, G, b->stream, a));
      b->ticket_base += per_step;
      (*launches)++;
    }
    return STH_OK;
  }
  std::vector<Plan> plans(b->shots.size());
  std::vector<SweepArgs> args(b->shots.size());
  for (size_t s = 0; s < b->shots.size(); s++) {
    // one plain launch per shot: the shot's own tuning (the single-grid
    // default unless set on the shot), one workgroup per tile (a persistent
    // launch carries its queue base on the host, which a replayed graph
    // would not advance)
    sth_tuning t = b->shots[s]->tune;
    t.launch = 1;
    plans[s] = make_plan(b->shots[s], t);
    memset(&args[s], 0, sizeof(SweepArgs));
    fill_weights(b->shots[s], true, args[s]);
  }
  for (uint32_t n = 0; n < steps; n++) {
    for (size_t s = 0; s < b->shots.size(); s++) {
      sth_ctx *c = b->shots[s];
      const int ub = c->ubuf;
      c->ubuf = (parity0 + (int)n) & 1;
      base_args(c, plans[s], args[s]);
      c->ubuf = ub;
      args[s].c0 = c0;
      args[s].c1 = c1;
      args[s].c2n = c2 ? c2[n] : 0.0f;
      STH_TRY(launch_sweep((int)b->npts, MODE_LEAP, FMT_F32, plans[s].mj, plans[s].nw,
                           plans[s].vec, plans[s].edge, plans[s].grid, b->stream, args[s]));
      (*launches)++;
    }
  }
  return STH_OK;
}

// The call's launches: `passes` passes of S steps, then `single` steps of
// the one-step method (persistent after fused passes).
int enqueue_all(sth_batch *b, int S, uint32_t passes, uint32_t single, float c0, float c1,
                const float *c2, int base, int parity0, uint64_t *launches) {
  int st = STH_OK;
  if (passes && (st = enqueue_passes(b, S, passes, c0, c1, c2, parity0, launches))) return st;
  if (!single) return STH_OK;
  const int method = S ? STH_BATCH_PERSISTENT : base;
  const uint32_t done = (uint32_t)S * passes;
  return enqueue_steps(b, single, c0, c1, c2 ? c2 + done : nullptr, method,
                       (parity0 + (int)done) & 1, launches);
}

}  // namespace

extern "C" {

int sth_batch_destroy(sth_batch *b) {
  if (!b) return STH_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipGraphExec_t x : b->execs) (void)hipGraphExecDestroy(x);
  for (sth_ctx *c : b->shots)
    if (c) sth_destroy(c);
  if (b->shot_args) (void)hipFree(b->shot_args);
  if (b->items) (void)hipFree(b->items);
  if (b->shot2) (void)hipFree(b->shot2);
  if (b->items2) (void)hipFree(b->items2);
  if (b->ticket) (void)hipFree(b->ticket);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->evs) (void)hipEventDestroy(b->evs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return STH_OK;
}

int sth_batch_create(const uint32_t *dims, uint32_t nshots, uint32_t npts, uint32_t dtype,
                     sth_batch **out) {
  if (!dims || !out || nshots == 0 || nshots > 65536) return STH_ERR_ARG;
  *out = nullptr;
  if (dtype != STH_FP32) return STH_ERR_UNSUPPORTED;
  if (radius_of(npts) < 0) return STH_ERR_ARG;
  sth_batch *b = new (std::nothrow) sth_batch();
  if (!b) return STH_ERR_NOMEM;
  b->npts = npts;
  b->dtype = dtype;
  b->tune.xcd_remap = 1;
  int st = STH_OK;
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evs, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void **)&b->ticket, 256);
  if (e == hipSuccess) e = launch_zero_u32(b->ticket, 1, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) st = hip_status(e);
  for (uint32_t s = 0; s < nshots && st == STH_OK; s++) {
    sth_geom g = {dims[3 * s], dims[3 * s + 1], dims[3 * s + 2], npts, dtype};
    sth_ctx *c = nullptr;
    st = sth_create(&g, &c);
    if (st == STH_OK) {
      c->no_fuse = true;  // the batch's tables hold the shot's buffers
      b->shots.push_back(c);
    }
  }
  if (st != STH_OK) {
    sth_batch_destroy(b);
    return st;
  }
  *out = b;
  return STH_OK;
}

sth_ctx *sth_batch_shot(sth_batch *b, uint32_t shot) {
  if (!b || shot >= b->shots.size()) return nullptr;
  return b->shots[shot];
}

uint32_t sth_batch_size(const sth_batch *b) { return b ? (uint32_t)b->shots.size() : 0; }

int sth_batch_set_tuning(sth_batch *b, const sth_tuning *t, uint32_t planes, uint32_t grid_blocks) {
  if (!b || !t) return STH_ERR_ARG;
  const TileSel ts = tile_of(*t);
  if (!tile_supported(ts.mj, ts.nw, ts.vec)) return STH_ERR_ARG;
  b->tune = *t;
  b->planes = planes;
  b->grid_blocks = grid_blocks;
  b->tables_ok = false;
  // the shots' own tuning (the sequential method's) takes the tile; the
  // chunk count is the fused passes' knob, counted over the whole batch
  sth_tuning st = *t;
  st.chunks = 0;
  for (sth_ctx *c : b->shots) c->tune = st;
  return STH_OK;
}

int sth_batch_get_tuning(const sth_batch *b, sth_tuning *t, uint32_t *planes,
                         uint32_t *grid_blocks) {
  if (!b) return STH_ERR_ARG;
  if (t) {
    const TileSel ts = tile_of(b->tune);
    *t = b->tune;
    t->rows_per_wave = (uint32_t)ts.mj;
    t->waves = (uint32_t)ts.nw;
    t->vec = (uint32_t)ts.vec;
  }
  if (planes) *planes = b->planes;
  if (grid_blocks) {
    *grid_blocks = persistent_grid(b, tile_of(b->tune));
  }
  return STH_OK;
}

int sth_batch_leapfrog(sth_batch *b, uint32_t steps, float c0, float c1, const float *c2,
                       uint32_t method) {
  if (!b) return STH_ERR_ARG;
  const int base = (int)(method & 0xff);
  const bool graph = (method & STH_BATCH_GRAPH) != 0;
  if (base != STH_BATCH_PERSISTENT && base != STH_BATCH_SEQUENTIAL &&
      base != STH_BATCH_TWO_STEP && base != STH_BATCH_THREE_STEP && base != STH_BATCH_FUSED)
    return STH_ERR_ARG;
  if (base == STH_BATCH_THREE_STEP && b->npts == 25) return STH_ERR_UNSUPPORTED;
  if (steps == 0) return STH_OK;
  // One u/v parity for the whole batch (its launches pick every shot's u
  // buffer from it): a shot stepped on its own since the last batch call
  // (sth_leapfrog on sth_batch_shot) holds u in the other buffer.
  const int parity0 = b->shots[0]->ubuf;
  for (const sth_ctx *c : b->shots)
    if (c->ubuf != parity0) return STH_ERR_STATE;
  // STH_BATCH_FUSED: the single-grid default pass of the stencil
  const int S = base == STH_BATCH_TWO_STEP     ? 2
                : base == STH_BATCH_THREE_STEP ? 3
                : base == STH_BATCH_FUSED      ? (b->npts == 7 ? 5 : b->npts == 13 ? 3 : 2)
                                               : 0;
  const uint32_t passes = S ? steps / (uint32_t)S : 0;
  int st = STH_OK;
  if (passes) {
    for (sth_ctx *c : b->shots)
      if ((st = ensure_twins(c))) return st;
    int mj, nw;
    shape2(b->npts, S, mj, nw);
    if ((st = build_tables2(b, mj, nw, S))) return st;
    if (passes & 1) {
      // after an odd number of passes each shot's u and v live in the
      // buffers that are now its twins: swap the names now, so the one-step
      // table below (for an odd last step) already points at them
      for (sth_ctx *c : b->shots) {
        std::swap(c->base[0], c->twin[0]);
        std::swap(c->base[1], c->twin[1]);
      }
      b->tables_ok = false;
    }
  }
  const uint32_t single = steps - (uint32_t)S * passes;  // one-step steps of this call
  if (single && (st = build_tables(b))) {
    if (passes & 1)
      for (sth_ctx *c : b->shots) {  // nothing was queued: restore the names
        std::swap(c->base[0], c->twin[0]);
        std::swap(c->base[1], c->twin[1]);
      }
    return st;
  }
  if ((st = join_shots(b))) return st;
  if (!b->execs.empty()) {  // the previous call's graphs have been launched: free them
    STH_TRY(hipStreamSynchronize(b->stream));
    for (hipGraphExec_t x : b->execs) STH_TRY(hipGraphExecDestroy(x));
    b->execs.clear();
  }
  uint64_t launches = 0;
  if (graph) {
    // The call is captured in segments of at most kGraphLaunches launches,
    // one graph each (whole steps, and an even number of fused passes so
    // every segment starts on the call's first buffer phase), then the
    // graphs are launched in order.  Each graph starts from a zeroed ticket
    // counter, so every replay draws the same tickets.  (One graph of 6400
    // launches, 64 shots x 100 sequential steps, crashed rocprofv3's kernel
    // trace on the host; segments of 1024 run under it.)
    const unsigned saved_base = b->ticket_base;
    auto capture = [&](auto &&enqueue) -> int {
      hipGraph_t g = nullptr;
      STH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = launch_zero_u32(b->ticket, 1, b->stream);
      b->ticket_base = 0;
      const int sq = e == hipSuccess ? enqueue() : STH_OK;
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (e != hipSuccess || sq != STH_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return sq != STH_OK ? sq : hip_status(e != hipSuccess ? e : ec);
      }
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return hip_status(e);
      b->execs.push_back(x);
      return STH_OK;
    };
    const uint32_t seg_passes = kGraphLaunches & ~1u;
    for (uint32_t q = 0; q < passes && st == STH_OK;) {
      const uint32_t np = std::min(passes - q, seg_passes), done = (uint32_t)S * q;
      st = capture([&] {
        return enqueue_passes(b, S, np, c0, c1, c2 ? c2 + done : nullptr,
                              (parity0 + (int)done) & 1, &launches);
      });
      q += np;
    }
    if (single) {
      const int method = S ? STH_BATCH_PERSISTENT : base;
      const uint32_t done = (uint32_t)S * passes;
      const uint32_t per_step = method == STH_BATCH_SEQUENTIAL ? (uint32_t)b->shots.size() : 1;
      const uint32_t seg = std::max<uint32_t>(1, kGraphLaunches / per_step);
      for (uint32_t k = 0; k < single && st == STH_OK;) {
        const uint32_t nk = std::min(single - k, seg), at = done + k;
        st = capture([&] {
          return enqueue_steps(b, nk, c0, c1, c2 ? c2 + at : nullptr, method,
                               (parity0 + (int)at) & 1, &launches);
        });
        k += nk;
      }
    }
    if (st != STH_OK) {
      b->ticket_base = saved_base;
      return st;
    }
    if (b->timing) STH_TRY(hipEventRecord(b->ev0, b->stream));
    for (hipGraphExec_t x : b->execs) STH_TRY(hipGraphLaunch(x, b->stream));
    if (b->timing) STH_TRY(hipEventRecord(b->ev1, b->stream));
  } else {
    if (b->timing) STH_TRY(hipEventRecord(b->ev0, b->stream));
    st = enqueue_all(b, S, passes, single, c0, c1, c2, base, parity0, &launches);
    if (st) return st;
    if (b->timing) STH_TRY(hipEventRecord(b->ev1, b->stream));
  }
  if (b->timing) {
    STH_TRY(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    STH_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->timed_ms += ms;
    b->launches += launches;
  }
  for (sth_ctx *c : b->shots) c->ubuf = (parity0 + (int)steps) & 1;
  return fork_shots(b);
}

int sth_batch_sync(sth_batch *b) {
  if (!b) return STH_ERR_ARG;
  STH_TRY(hipStreamSynchronize(b->stream));
  for (sth_ctx *c : b->shots) STH_TRY(hipStreamSynchronize(c->stream));
  return STH_OK;
}

int sth_batch_timer_start(sth_batch *b) {
  if (!b) return STH_ERR_ARG;
  b->timing = true;
  b->timed_ms = 0.0;
  b->launches = 0;
  return STH_OK;
}

int sth_batch_timer_read(sth_batch *b, double *ms, uint64_t *launches) {
  if (!b || !b->timing) return STH_ERR_STATE;
  b->timing = false;
  if (ms) *ms = b->timed_ms;
  if (launches) *launches = b->launches;
  return STH_OK;
}

int sth_batch_autotune(sth_batch *b, uint32_t reps, sth_tuning *best, uint32_t *best_planes,
                       float *best_ms) {
  if (!b || reps == 0) return STH_ERR_ARG;
  static const int tiles[][3] = {{1, 4, 1}, {2, 4, 1}, {4, 4, 1}, {2, 8, 1}, {4, 8, 1},
                                 {1, 4, 2}, {2, 4, 2}, {4, 4, 2}, {2, 8, 2}, {4, 8, 2}};
  static const uint32_t plane_opts[] = {0, 128, 64, 32};
  const sth_tuning saved = b->tune;
  const uint32_t saved_planes = b->planes, saved_grid = b->grid_blocks;
  const bool saved_timing = b->timing;
  sth_tuning bt = saved;
  uint32_t bp = saved_planes;
  float bms = 1e30f;
  for (const auto &tl : tiles)
    for (uint32_t pl : plane_opts) {
      sth_tuning t = saved;
      t.rows_per_wave = (uint32_t)tl[0];
      t.waves = (uint32_t)tl[1];
      t.vec = (uint32_t)tl[2];
      int st = sth_batch_set_tuning(b, &t, pl, saved_grid);
      if (st) return st;
      b->timing = false;
      st = sth_batch_leapfrog(b, 1, 1.0f, 0.0f, nullptr, STH_BATCH_PERSISTENT);
      if (st) return st;
      b->timing = true;
      b->timed_ms = 0.0;
      b->launches = 0;
      st = sth_batch_leapfrog(b, reps, 1.0f, 0.0f, nullptr, STH_BATCH_PERSISTENT);
      if (st) return st;
      const float ms = (float)(b->timed_ms / reps);
      if (ms < bms) {
        bms = ms;
        bt = t;
        bp = pl;
      }
    }
  // The fused pass of the stencil's default length (STH_BATCH_FUSED): its
  // i-chunks per shot (0: about six rounds of its persistent grid), timed
  // over whole passes.
  static const uint32_t chunk_opts[] = {0, 1, 2, 3, 4, 6, 8, 12};
  const uint32_t S = b->npts == 7 ? 5 : b->npts == 13 ? 3 : 2;
  const uint32_t nsteps = (std::max(reps, S) + S - 1) / S * S;
  float fms = 1e30f;
  uint32_t fch = 0;
  for (uint32_t ch : chunk_opts) {
    sth_tuning t = bt;
    t.chunks = ch;
    int st = sth_batch_set_tuning(b, &t, bp, saved_grid);
    if (st) return st;
    b->timing = false;
    st = sth_batch_leapfrog(b, S, 1.0f, 0.0f, nullptr, STH_BATCH_FUSED);
    if (st) return st;
    b->timing = true;
    b->timed_ms = 0.0;
    b->launches = 0;
    st = sth_batch_leapfrog(b, nsteps, 1.0f, 0.0f, nullptr, STH_BATCH_FUSED);
    if (st) return st;
    const float ms = (float)(b->timed_ms / nsteps);
    if (ms < fms) {
      fms = ms;
      fch = ch;
    }
  }
  bt.chunks = fch;
  b->best_method = fms < bms ? STH_BATCH_FUSED : STH_BATCH_PERSISTENT;
  b->timing = saved_timing;
  b->timed_ms = 0.0;
  b->launches = 0;
  int st = sth_batch_set_tuning(b, &bt, bp, saved_grid);
  if (st) return st;
  if (best) *best = bt;
  if (best_planes) *best_planes = bp;
  if (best_ms) *best_ms = std::min(bms, fms);
  return STH_OK;
}

int sth_batch_best_method(const sth_batch *b, uint32_t *method) {
  if (!b || !method) return STH_ERR_ARG;
  *method = b->best_method;
  return STH_OK;
}

}  // extern "C"

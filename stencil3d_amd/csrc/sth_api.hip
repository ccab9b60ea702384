// sth_api.hip -- the extern "C" boundary of libstencil_hip (include/stencil_hip.h).
//
// Host side of the sweep: context (stream + hipMalloc'd fields), host<->device
// layout translation, initialisers, the per-step launch loop, check/norms,
// tuning and timing.  No call exits the process; every failure is a status.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "fields.h"
#include "stencil_hip.h"
#include "sth_common.h"
#include "sweep_dispatch.h"
#include "sweep2_dispatch.h"
#include "sweep2_kernel.h"
#include "sth_internal.h"

using namespace sth;


namespace sth {
namespace detail {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return STH_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return STH_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return STH_ERR_NODEV;
  return STH_ERR_HIP;
}

#define STH_TRY(expr)                            \
  do {                                           \
    hipError_t e_ = (expr);                      \
    if (e_ != hipSuccess) return hip_status(e_); \
  } while (0)

inline bool is16(const sth_ctx *c) { return c->fmt == FMT_F16 || c->fmt == FMT_BF16; }

inline int buf_of(const sth_ctx *c, sth_field f) {
  return f == STH_FIELD_U ? c->ubuf : f == STH_FIELD_V ? 1 - c->ubuf : 2;
}

inline FieldDesc desc_of(const sth_ctx *c) {
  FieldDesc d;
  d.pitch = (int64_t)c->L.pitch;
  d.plane = (int64_t)c->L.plane;
  d.align = kRowAlignBytes / c->eb;
  d.elems = c->L.elems;
  d.xs = (int)c->g.xs;
  d.ys = (int)c->g.ys;
  d.zs = (int)c->g.zs;
  d.R = c->R;
  return d;
}

// Element offset of the first element of plane i (cell (i, -R, -R)).
inline uint64_t plane_start(const sth_ctx *c, int64_t i) {
  return c->L.origin - c->R - c->R * c->L.pitch + (uint64_t)(i * (int64_t)c->L.plane);
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t field_key(uint64_t seed, int field) {
  return mix64(seed + (uint64_t)(field + 1) * 0x9E3779B97F4A7C15ull);
}

inline uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
inline float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Host encoders/decoders of the 16-bit storage formats (round to nearest even).
inline uint16_t enc16(int fmt, float x, int e) {
  const float y = ldexpf(x, -e);
  if (fmt == FMT_F16) {
    _Float16 h = (_Float16)y;
    uint16_t b;
    memcpy(&b, &h, 2);
    return b;
  }
  uint32_t b = f2u(y);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
inline float dec16(int fmt, uint16_t q, int e) {
  if (fmt == FMT_F16) {
    _Float16 h;
    memcpy(&h, &q, 2);
    return ldexpf((float)h, e);
  }
  return ldexpf(u2f((uint32_t)q << 16), e);
}

// Pinned staging for host transfers: two halves of at least one device
// plane each, so the host fills (or reads) one while the other is in flight.
int ensure_staging(sth_ctx *c) {
  if (c->staging) return STH_OK;
  const size_t want = std::max<size_t>((size_t)128 << 20, 2 * c->L.plane * c->eb);
  STH_TRY(hipHostMalloc(&c->staging, want, hipHostMallocDefault));
  c->staging_bytes = want;
  return STH_OK;
}

int check_host_layout(const sth_ctx *c, const sth_host_layout *hl) {
  if (!hl) return STH_ERR_ARG;
  if ((int)hl->xh < c->R || (int)hl->yh < c->R || (int)hl->zh < c->R) return STH_ERR_ARG;
  return STH_OK;
}

struct HostIdx {
  int64_t jm, im, ofs;
};
inline HostIdx host_idx(const sth_ctx *c, const sth_host_layout *hl) {
  HostIdx h;
  h.jm = (int64_t)c->g.zs + 2 * hl->zh;
  h.im = h.jm * ((int64_t)c->g.ys + 2 * hl->yh);
  h.ofs = (int64_t)hl->zh + (int64_t)hl->yh * h.jm + (int64_t)hl->xh * h.im;
  return h;
}

// Host threads for the row conversions of a transfer (memory-bound loops
// over host arrays; no HIP call is made on them).
int host_threads() {
  static const int n = [] {
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
  }();
  return n;
}

// fn(i, j, t) for every cell row (i, j) of planes [ia, ib), rows [-R, ys+R),
// split into contiguous runs over `host_threads()` threads (t: the thread).
template <class Fn>
void parallel_rows(int64_t ia, int64_t ib, int64_t R, int64_t ys, Fn &&fn) {
  const int64_t nj = ys + 2 * R, rows = (ib - ia) * nj;
  const int nt = (int)std::min<int64_t>(host_threads(), std::max<int64_t>(1, rows / 64));
  auto run = [&](int t) {
    const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
    for (int64_t r = r0; r < r1; r++) fn(ia + r / nj, r % nj - R, t);
  };
  if (nt == 1) {
    run(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  int t = 1;
  try {
    for (; t < nt; t++) pool.emplace_back(run, t);
  } catch (...) {  // no thread to spare: the calling thread takes the rest
  }
  for (int u = t; u < nt; u++) run(u);
  run(0);
  for (auto &th : pool) th.join();
}

// Moves planes [-R, xs+R) between a host dense array and a device buffer via
// the pinned staging buffer, in batches of whole planes that alternate
// between its two halves: uploads fill one half while the other's copy runs,
// downloads convert one half while the next batch arrives in the other.
// `row(i, j, srow, t)` is called per (i, j) cell row with the staging row
// (k = -R) on host thread t (rows of one batch run concurrently).
template <class RowFn>
int transfer_planes(sth_ctx *c, int buf, bool to_device, RowFn row) {
  int st = ensure_staging(c);
  if (st) return st;
  const int64_t R = c->R, xs = c->g.xs, ys = c->g.ys;
  const size_t plane_bytes = c->L.plane * c->eb;
  const int64_t per = std::max<int64_t>(1, (int64_t)(c->staging_bytes / 2 / plane_bytes));
  char *half[2] = {(char *)c->staging, (char *)c->staging + (size_t)per * plane_bytes};
  char *dev = (char *)c->base[buf];
  if (to_device) c->twins_ok = false;
  const int64_t nb = (xs + 2 * R + per - 1) / per;  // batches
  auto first = [&](int64_t b) { return -R + b * per; };
  auto last = [&](int64_t b) { return std::min<int64_t>(xs + R, first(b) + per); };
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (int q = 0; q < 2 && !st; q++)
    if (hipEventCreateWithFlags(&ev[q], hipEventDisableTiming) != hipSuccess) st = STH_ERR_HIP;
  auto copy = [&](int64_t b) {  // batch b between the device and its half
    const size_t bytes = (size_t)(last(b) - first(b)) * plane_bytes;
    char *dptr = dev + plane_start(c, first(b)) * c->eb;
    const hipError_t e =
        to_device ? hipMemcpyAsync(dptr, half[b & 1], bytes, hipMemcpyHostToDevice, c->stream)
                  : hipMemcpyAsync(half[b & 1], dptr, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) return hipEventRecord(ev[b & 1], c->stream);
    return e;
  };
  auto fill = [&](int64_t b) {  // the host side of batch b
    const int64_t ia = first(b);
    char *base = half[b & 1];
    parallel_rows(ia, last(b), R, ys, [&](int64_t i, int64_t j, int t) {
      row(i, j, base + ((size_t)(i - ia) * c->L.plane + (size_t)(j + R) * c->L.pitch) * c->eb,
          t);
    });
  };
  hipError_t e = hipSuccess;
  if (!st && to_device) {
    for (int64_t b = 0; b < nb && e == hipSuccess; b++) {
      if (b >= 2) e = hipEventSynchronize(ev[b & 1]);  // this half's previous copy is done
      if (e != hipSuccess) break;
      fill(b);
      e = copy(b);
    }
  } else if (!st) {
    e = copy(0);
    for (int64_t b = 0; b < nb && e == hipSuccess; b++) {
      if (b + 1 < nb) e = copy(b + 1);  // its half held batch b-1, converted already
      if (e == hipSuccess) e = hipEventSynchronize(ev[b & 1]);
      if (e == hipSuccess) fill(b);
    }
  }
  // drain the stream on every path: no copy may still use the staging buffer
  // when the call returns
  const hipError_t es = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = es;
  for (int q = 0; q < 2; q++)
    if (ev[q]) (void)hipEventDestroy(ev[q]);
  if (st) return st;
  return e == hipSuccess ? STH_OK : hip_status(e);
}

// Current 16-bit scale state of the context (from the device after sweeps).
int fetch_slot(sth_ctx *c) {
  if (!is16(c) || c->host_dirty) return STH_OK;
  STH_TRY(hipMemcpyAsync(&c->host_slot, c->qs_dev + c->slot, sizeof(Q16Slot),
                         hipMemcpyDeviceToHost, c->stream));
  STH_TRY(hipStreamSynchronize(c->stream));
  c->host_dirty = true;  // host copy is now authoritative until the next sweep
  return STH_OK;
}

int push_slot(sth_ctx *c) {
  if (!is16(c) || !c->host_dirty) return STH_OK;
  Q16Slot img[2];
  img[0] = c->host_slot;
  memset(&img[1], 0, sizeof(Q16Slot));
  // slot s holds the state, slot s+1 must start with zero maxima
  STH_TRY(hipMemcpyAsync(c->qs_dev + c->slot, &img[0], sizeof(Q16Slot), hipMemcpyHostToDevice,
                         c->stream));
  STH_TRY(hipMemcpyAsync(c->qs_dev + (c->slot + 1) % 3, &img[1], sizeof(Q16Slot),
                         hipMemcpyHostToDevice, c->stream));
  STH_TRY(hipStreamSynchronize(c->stream));
  c->host_dirty = false;
  return STH_OK;
}

int device_fill(sth_ctx *c, int buf, const FillSpec &f, bool halo_zero) {
  const FieldDesc d = desc_of(c);
  c->twins_ok = false;
  int e = 0;
  uint32_t mbits = 0;
  if (is16(c)) {
    int st = fetch_slot(c);
    if (st) return st;
    STH_TRY(launch_zero_u32(c->maxbits_dev, 1, c->stream));
    STH_TRY(launch_fill(d, f, c->base[buf], c->fmt, 0, 0, c->maxbits_dev, c->stream));
    STH_TRY(hipMemcpyAsync(&mbits, c->maxbits_dev, 4, hipMemcpyDeviceToHost, c->stream));
    STH_TRY(hipStreamSynchronize(c->stream));
    e = scale_exp((double)u2f(mbits));
  }
  STH_TRY(launch_fill(d, f, c->base[buf], c->fmt, e, 1, c->maxbits_dev, c->stream));
  if (is16(c)) {
    // the buffer's max is max |q| of the stored numbers (exponent 0 decode)
    STH_TRY(launch_zero_u32(c->maxbits_dev, 4, c->stream));
    STH_TRY(launch_norms(d, c->base[buf], c->fmt, 0, c->partial_dev, c->maxbits_dev, c->sum_dev,
                         c->stream));
    STH_TRY(hipMemcpyAsync(&mbits, c->maxbits_dev, 4, hipMemcpyDeviceToHost, c->stream));
    STH_TRY(hipStreamSynchronize(c->stream));
    c->host_slot.exp[buf] = e;
    c->host_slot.maxbits[buf] = mbits;
  }
  c->halo_zero[buf] = halo_zero;
  return STH_OK;
}

void fill_weights(const sth_ctx *c, bool leap, SweepArgs &a) {
  Coeffs cf;
  stencil_coeffs(c->g.npts, cf);
  double wref[3][4];
  stencil_coeffs_double(c->g.npts, wref);
  const bool refd = c->fmt == FMT_F64 && c->weights == STH_WEIGHTS_REFERENCE;
  int ax[12], rr[12];
  const int T = term_order(c->g.npts, ax, rr);
  for (int t = 0; t < 12; t++) {
    a.w[t] = 0.0f;
    a.wd[t] = 0.0;
  }
  for (int t = 0; t < T; t++) {
    const float w = cf.w[ax[t]][rr[t] - 1];
    const double wd = refd ? wref[ax[t]][rr[t] - 1] : (double)w;
    a.w[t] = leap ? 2.0f * w : w;
    a.wd[t] = leap ? 2.0 * wd : wd;
  }
  a.coef0 = cf.coef0;
  a.mA = -cf.A;
  a.gain = lap_gain(c->g.npts);
}


// A one-step tile the context's sweeps can run: a compiled sweep_kernel
// tile, or (FP16, two k points per lane) a packed tile.
bool shape_supported(const sth_ctx *c, int mj, int nw, int vec) {
  if (packed16(c->fmt, MODE_LEAP, vec)) return tile16_supported(mj, nw);
  return tile_supported(mj, nw, vec);
}

Plan make_plan(const sth_ctx *c, const sth_tuning &t) {
  Plan p;
  TileShape ts = default_tile(c->fmt, c->g.ys, c->g.zs, t);
  if (!shape_supported(c, ts.mj, ts.nw, ts.vec))
    ts = default_tile(c->fmt, c->g.ys, c->g.zs, sth_tuning{});
  p.mj = ts.mj;
  p.nw = ts.nw;
  p.vec = ts.vec;
  const int TJ = p.mj * p.nw, TK = 64 * p.vec;
  p.ntk = ((int)c->g.zs + TK - 1) / TK;
  p.ntj = ((int)c->g.ys + TJ - 1) / TJ;
  // A last k-tile with at most 64/f lanes' columns runs folded (f rows per
  // lane group): at 1040 columns and 128-column tiles it holds 16 columns,
  // and unfolded it would cost as much as a full tile.  FP32 folds it onto
  // the tile's own rows; the packed FP16 sweep runs it f tiles' rows at a
  // time (measured at 1040^3, 25 points, 2x8: 1.89 ms per step against
  // 1.93 unfolded).  What the narrow column still costs is mostly its rows'
  // 128-byte lines, which L2 fetches whole: 1032 to 1088 columns take about
  // as long per step, 6 to 10 % more than 1024.
  p.edge = 0;
  const int we = (int)c->g.zs - (p.ntk - 1) * 64 * p.vec;
  if (c->fmt == FMT_F32 && we < 64 * p.vec)
    for (int f : {4, 2, 1})
      if (64 / f >= we && p.mj % f == 0 && (f > 1 || p.vec == 2) &&
          edge_supported(p.mj, p.nw, p.vec, f)) {
        p.edge = f;
        break;
      }
  if (packed16(c->fmt, MODE_LEAP, p.vec) && we < 128)
    for (int f : {8, 4, 2})
      if (128 / f >= we && edge16_supported((int)c->g.npts, p.mj, p.nw, f)) {
        p.edge = f;
        break;
      }
  const int xs = (int)c->g.xs;
  int ch = (int)t.chunks;
  if (ch == 0) {
    // i-chunks of about 64 planes, and at least 4096 workgroups where the
    // grid allows it: many short-lived workgroups drain evenly (no tail of a
    // few long ones); a chunk re-reads 2R planes of u, so none is shorter
    // than max(8, 4R) planes.
    const int tiles = p.ntk * p.ntj;
    ch = std::max((xs + 63) / 64, (4096 + tiles - 1) / tiles);
    ch = std::min(ch, std::max(1, xs / std::max(8, 4 * c->R)));
    ch = std::max(ch, 1);
  }
  ch = std::min(ch, xs);
  p.chunk_len = (xs + ch - 1) / ch;
  if (packed16(c->fmt, MODE_LEAP, p.vec)) {
    // the packed FP16 sweep runs whole turns of its rings: chunks of the
    // nearest multiple of a turn leave idle steps in the last chunk only
    const int u = sweep16_turn(c->R);
    p.chunk_len = std::min(xs, std::max(u, (p.chunk_len + u / 2) / u * u));
  }
  p.chunks = (xs + p.chunk_len - 1) / p.chunk_len;
  // items per chunk: every tile; with the packed sweep's folded edge, the
  // full tiles and one folded tile per p.edge rows of tiles
  const int per_chunk = packed16(c->fmt, MODE_LEAP, p.vec) && p.edge > 0
                            ? p.ntj * (p.ntk - 1) + (p.ntj + p.edge - 1) / p.edge
                            : p.ntk * p.ntj;
  p.grid = (unsigned)(per_chunk * p.chunks);
  p.items = p.grid;
  p.persist = t.launch != 1 && c->tickets_dev != nullptr;
  if (p.persist) {
    // persistent: as many workgroups as stay resident (measured at 1040^3:
    // 10 % faster than one workgroup per tile -- no partly filled last round,
    // and a workgroup's next tile starts without a launch)
    int per_cu = 0;
    const void *fn = packed16(c->fmt, MODE_LEAP, p.vec)
                         ? sweep16_fn((int)c->g.npts, p.mj, p.nw, p.edge)
                         : sweep_fn((int)c->g.npts, MODE_LEAP, c->fmt, p.mj, p.nw, p.vec, 0);
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * p.nw, 0) !=
                   hipSuccess || per_cu < 1) {
      (void)hipGetLastError();
      per_cu = 1;
    }
    p.grid = std::min<unsigned>(p.items, (unsigned)(per_cu * std::max(1, c->cus)));
    // 16-bit storage: bands of about the tiles one XCD holds at once (its
    // blocks), rounded to whole rows of k-tiles, dealt round robin to the
    // XCDs: an XCD's tiles in flight are j-neighbours (their halo rows meet
    // in its L2) and the eight XCDs work on adjacent bands of one i-chunk.
    // FP32: a single queue in item order (measured at 1040^3: FP32 4x8x2
    // 3.26 ms single queue, 3.45 banded; FP16 2x8x2 2.13 banded, 2.17
    // single).  xcd_remap = 0: a single queue for every storage type.
    const int per_xcd = std::max(1, (int)p.grid / kXcds);
    const int rows = std::max(1, (per_xcd + p.ntk / 2) / p.ntk);
    const bool banded = t.xcd_remap && (c->fmt == FMT_F16 || c->fmt == FMT_BF16);
    p.band = banded ? rows * p.ntk : (int)p.items;
  }
  return p;
}

void base_args(const sth_ctx *c, const Plan &p, SweepArgs &a) {
  const size_t o = c->L.origin * c->eb;
  const int vb = 1 - c->ubuf;
  a.u = (const char *)c->base[c->ubuf] + o;
  a.v = (char *)c->base[vb] + o;
  a.s = (const char *)c->base[2] + o;
  a.pitch = (int64_t)c->L.pitch;
  a.plane = (int64_t)c->L.plane;
  a.xs = (int)c->g.xs;
  a.ys = (int)c->g.ys;
  a.zs = (int)c->g.zs;
  a.ntk = p.ntk;
  a.ntj = p.ntj;
  a.chunks = p.chunks;
  a.chunk_len = p.chunk_len;
  a.remap = c->tune.xcd_remap ? 1 : 0;
  a.qs = c->qs_dev;
  a.slot = c->slot;
  a.ubuf = c->ubuf;
  a.vbuf = vb;
  a.c0 = a.c1 = a.c2n = 0.0f;
}

int launch_one(sth_ctx *c, const Plan &p, int mode, SweepArgs &a) {
  a.tickets = nullptr;
  if (p.persist) {
    // counter x advances by its range size + the blocks on XCD x per launch;
    // restart the counts well before one could wrap
    bool wrap = false;
    for (int x = 0; x < kXcds; x++) wrap = wrap || c->tick_next[x] > 0x7fffffffu;
    if (wrap) {
      const hipError_t z = launch_zero_u32(c->tickets_dev, 16 * kXcds, c->stream);
      if (z != hipSuccess) return hip_status(z);
      for (int x = 0; x < kXcds; x++) c->tick_next[x] = 0;
    }
    a.tickets = c->tickets_dev;
    for (int x = 0; x < kXcds; x++) a.tick_base[x] = c->tick_next[x];
    a.items = (int)p.items;
    a.band = p.band;
  }
  const hipError_t e =
      packed16(c->fmt, mode, p.vec)
          ? launch_sweep16((int)c->g.npts, p.mj, p.nw, p.edge, p.grid, c->stream, a)
          : launch_sweep((int)c->g.npts, mode, c->fmt, p.mj, p.nw, p.vec, p.edge, p.grid,
                         c->stream, a);
  if (e != hipSuccess) return hip_status(e);
  if (p.persist && p.band >= (int)p.items) {
    c->tick_next[0] += p.items + p.grid;  // one queue
  } else if (p.persist) {
    for (int x = 0; x < kXcds; x++) {
      const int blocks = ((int)p.grid - x + kXcds - 1) / kXcds;
      c->tick_next[x] += (uint32_t)(xcd_band_items((int)p.items, p.band, x) + std::max(0, blocks));
    }
  }
  if (c->timing) c->launches++;
  return STH_OK;
}

// REF sweeps (iterate_3d, iterators.cpp:130-157): v is updated in place
// from read-only u and s, so each point's sweeps are independent of every
// other point's; a launch applies up to tune.ref_sweeps of them.
int run_ref(sth_ctx *c, const Plan &p, uint32_t iterations) {
  SweepArgs a;
  memset(&a, 0, sizeof(a));
  base_args(c, p, a);
  fill_weights(c, false, a);
  const uint32_t k = std::max<uint32_t>(1, c->tune.ref_sweeps);
  for (uint32_t n = 0; n < iterations;) {
    const uint32_t m = std::min(k, iterations - n);
    a.ref_extra = (int)m - 1;
    int st = launch_one(c, p, MODE_REF, a);
    if (st) return st;
    n += m;
  }
  return STH_OK;
}

// 16-bit passes (sweep2h_kernel.h): the default shape of each pass length
// (rows per wave; 16 waves), and whether the stencil has a pass of that
// length in the storage format (FP16: two steps for every stencil, three
// for 7 and 13 points, four and five for 7; BF16: two and three steps for 7
// points, two for 13).
int q16_pass_mj(int fmt, int npts, int s) {
  if (fmt == FMT_BF16) return npts == 7 && s == 2 ? 4 : 2;
  return s == 2 ? (npts == 25 ? 3 : 4) : s == 3 ? 3 : 2;
}
bool q16_pass_exists(int fmt, int npts, int s) {
  return s >= 2 && s <= 5 && fused2h_supported(fmt, npts, q16_pass_mj(fmt, npts, s), 16, s);
}

// The default 16-bit pass length of a grid (0: the one-step sweep), from
// the number of output tiles of each pass's default shape.  Measured on
// cubic grids (one-step time / pass time per step).  FP16, against the
// packed one-step sweep:
//   7 points, three steps: 0.89-0.92 at 96^3 and 128^3 (3 and 6 output
//     tiles), 2.7 at 192^3 (10), 1.8-2.0 from 256^3 to 640^3; four steps
//     2.0-2.3 from 640^3, 2-6 % faster than three from 768^3 (240 tiles of
//     its 2 x 16 shape) to 1040^3, 1 % slower at 640^3 (175); five steps
//     slower than four at 1040^3 (0.98 against 0.88 ms per step); two steps
//     slower than three everywhere;
//   13 points, two steps 0.87-0.92 up to 512^3 (54 tiles of 4 x 16),
//     1.1-1.25 from 640^3 (77); three steps 3-5 % faster than two from
//     896^3 (230 tiles of 3 x 16) to 1040^3, slower at 768^3 (160);
//   25 points: the two-step pass does 1.47 lane-updates per point-update
//     and is instruction-bound where the sweep is held by the HBM
//     read/write mix (0.63-0.98 from 256^3 to 1040^3), so it stays one-step.
// BF16 (against its one-step sweep):
//   7 points, two steps (4 x 16): 0.67 at 128^3 (6 tiles), 1.20 at 192^3
//     (8), 1.30 at 256^3; three steps (2 x 16): 1.15 at 256^3 (30 tiles of
//     its shape), 2.2 at 384^3 (56), 1.7-1.8 from 512^3 to 1040^3;
//   13 points, two steps (2 x 16): 1.04 at 256^3 (30 tiles), 1.59 at 384^3,
//     1.14-1.21 from 512^3 to 1040^3;
//   25 points: no pass fits 128 VGPRs (see the pass list).
int q16_default_pass(const sth_ctx *c) {
  const int npts = (int)c->g.npts;
  if (npts == 25 || (int)c->g.xs < 16 * c->R) return 0;
  auto tiles = [&](int s) {
    const Fused2Geom g = fused2h_geom(c->fmt, npts, q16_pass_mj(c->fmt, npts, s), 16, s);
    return (((int)c->g.zs + g.ok - 1) / g.ok) * (((int)c->g.ys + g.oj - 1) / g.oj);
  };
  if (c->fmt == FMT_BF16) {
    if (npts == 7) return tiles(3) >= 50 ? 3 : tiles(2) >= 8 ? 2 : 0;
    return tiles(2) >= 30 ? 2 : 0;
  }
  if (npts == 7) return tiles(4) >= 200 ? 4 : tiles(3) >= 10 ? 3 : 0;
  return tiles(3) >= 200 ? 3 : tiles(2) >= 70 ? 2 : 0;
}

// Several LEAPFROG steps per pass (FP32): the default on grids with enough
// output tiles to fill the CUs (the thresholds below), five steps per pass
// for 7 points, three for 13, two for 25; otherwise, or when the tuning asks
// for one step per pass, the one-step sweep.
bool fused_applies(const sth_ctx *c, const sth_tuning &t) {
  if (c->no_fuse) return false;
  if (is16(c)) {
    if (t.steps_per_pass)
      return q16_pass_exists(c->fmt, (int)c->g.npts, (int)t.steps_per_pass);
    return q16_default_pass(c) >= 2;
  }
  if (c->fmt != FMT_F32) return false;
  if (t.steps_per_pass) return t.steps_per_pass >= 2;
  // Measured on cubic grids (FP32 LEAPFROG, one-step time / pass time per
  // step): 7 points, four-step pass 1.14-1.22 from 24^3 (1 tile) to 48^3,
  // 1.21-3.36 from 64^3, five-step pass 1.21-2.58 from 64^3 to 1040^3;
  // 13 points, three-step pass 0.57-0.94 up to 9 tiles, 1.11-2.61 from 12;
  // 25 points, two-step pass 0.41-0.80 up to 42 tiles, 0.98-1.18 from 56.
  // Longer passes were faster than shorter ones nearly everywhere (7 points:
  // five steps 1.03-1.16x four except 0.96 at 256^3, four 1.01-1.36x three;
  // three steps 1.11-1.60x two).  Each chunk also recomputes (s-1)(2R+1)
  // planes, so short grids stay one-step.
  const Plan2 p = make_plan2(c, t);
  const int npts = (int)c->g.npts;
  const int least = npts == 7 ? 1 : npts == 13 ? 12 : 50;
  return p.ntk * p.ntj >= least && (int)c->g.xs >= 16 * c->R;
}

// Steps per pass (the tuning's, where the stencil has that pass; by default
// 5 for 7 points, 3 for 13, 2 for 25) and the fused shape (rows per wave,
// waves): a tuning that names a compiled shape gets it, otherwise the
// default of the stencil and pass length.
Plan2 make_plan2(const sth_ctx *c, const sth_tuning &t) {
  const int npts = (int)c->g.npts;
  Plan2 p;
  Fused2Geom g;
  if (is16(c)) {
    // 16-bit storage (sweep2h_kernel.h): the tuning's pass length where the
    // stencil has it, else the grid's default
    p.s = q16_pass_exists(c->fmt, npts, (int)t.steps_per_pass)
              ? (int)t.steps_per_pass
              : std::max(2, q16_default_pass(c));
    if (fused2h_supported(c->fmt, npts, (int)t.rows_per_wave, (int)t.waves, p.s)) {
      p.mj = (int)t.rows_per_wave;
      p.nw = (int)t.waves;
    } else {
      // two steps: 7 and 13 points 4 x 16 (13 points at 1040^3: 1.42 ms/step
      // against 1.48 for 3 x 16), 25 points 3 x 16 (4 x 16 exceeds 128
      // VGPRs); three steps 3 x 16, four and five 2 x 16 (the LDS planes of
      // the longer passes)
      p.mj = q16_pass_mj(c->fmt, npts, p.s);
      p.nw = 16;
    }
    g = fused2h_geom(c->fmt, npts, p.mj, p.nw, p.s);
  } else {
    if (npts == 25 || t.steps_per_pass == 2)
      p.s = 2;
    else if (npts == 7 && t.steps_per_pass == 4)
      p.s = 4;
    else if (npts == 7 && (t.steps_per_pass == 5 || t.steps_per_pass == 0))
      p.s = 5;
    else
      p.s = 3;
    if (fused2_supported(npts, (int)t.rows_per_wave, (int)t.waves, p.s)) {
      p.mj = (int)t.rows_per_wave;
      p.nw = (int)t.waves;
    } else if (p.s >= 4) {
      p.mj = 2;
      p.nw = 16;
    } else {
      p.mj = npts == 25 || p.s == 3 ? 3 : 4;
      p.nw = 16;
    }
    g = fused2_geom(npts, p.mj, p.nw, p.s);
  }
  p.oj = g.oj;
  p.ok = g.ok;
  p.ntk = ((int)c->g.zs + p.ok - 1) / p.ok;
  p.ntj = ((int)c->g.ys + p.oj - 1) / p.oj;
  const int xs = (int)c->g.xs, tiles = p.ntk * p.ntj, R = c->R;
  // A chunk of n planes takes n + (s-1)(2R+1) steps and the kernel runs
  // them in whole groups of ns = lcm(2, s-1)(R+1) steps: n = -(s-1)(2R+1)
  // mod ns leaves no idle step (except in a shorter last chunk).
  const int ns = ((p.s - 1) % 2 ? 2 : 1) * (p.s - 1) * (R + 1),
            extra = (p.s - 1) * (2 * R + 1);
  auto chunk_len = [&](int ch) {
    const int n = (xs + ch - 1) / ch;
    return std::min(xs, n + ((-extra - n) % ns + ns) % ns);
  };
  int ch = (int)t.chunks;
  if (ch == 0) {
    // One workgroup per CU fits: the pass takes about ceil(workgroups /
    // CUs) rounds of (chunk length + (s-1)(2R+1)) steps.  Take the chunk
    // count with the least such time among those giving at least six rounds
    // (few long rounds balance badly) and chunks of at least 8R planes (each
    // chunk recomputes (s-1)2R planes of the first step and re-reads
    // (s+1)2R of u).
    const int64_t slots = std::max(1, c->cus);
    const int most = std::max(1, std::min(32, xs / (8 * R)));
    const int least = (int)std::min<int64_t>(most, (6 * slots + tiles - 1) / tiles);
    int64_t best = -1;
    for (int k = std::max(1, least); k <= most; k++) {
      const int n = chunk_len(k), chunks = (xs + n - 1) / n;
      const int64_t rounds = ((int64_t)tiles * chunks + slots - 1) / slots;
      const int64_t cost = rounds * (n + extra);
      if (best < 0 || cost < best) {
        best = cost;
        ch = k;
      }
    }
  }
  ch = std::max(1, std::min(ch, xs));
  p.chunk_len = chunk_len(ch);
  p.chunks = (xs + p.chunk_len - 1) / p.chunk_len;
  p.grid = (unsigned)(tiles * p.chunks);
  return p;
}

int ensure_twins(sth_ctx *c) {
  for (int b = 0; b < 2; b++)
    if (!c->twin[b]) {
      if (hipMalloc(&c->twin[b], c->L.bytes) != hipSuccess) {
        (void)hipGetLastError();
        return STH_ERR_NOMEM;
      }
      c->twins_ok = false;
    }
  if (!c->twins_ok) {
    // the twins must carry the halos of a and b (the interiors are rewritten)
    for (int b = 0; b < 2; b++)
      STH_TRY(hipMemcpyAsync(c->twin[b], c->base[b], c->L.bytes, hipMemcpyDeviceToDevice,
                             c->stream));
    c->twins_ok = true;
  }
  return STH_OK;
}

// One pass of p.s steps; c2n holds the source coefficients of its steps.
int run_pass2(sth_ctx *c, const Plan2 &p, const SweepArgs &wts, float c0, float c1,
              const float *c2n) {
  const size_t o = c->L.origin * c->eb;
  const int ub = c->ubuf, vb = 1 - c->ubuf;
  Sweep2Args a;
  memset(&a, 0, sizeof(a));
  a.u = (const char *)c->base[ub] + o;
  a.v = (const char *)c->base[vb] + o;
  a.s = (const char *)c->base[2] + o;
  // The one-step scheme writes u_{n+1} over u_{n-1} (the halo of v),
  // u_{n+2} over u_n (the halo of u), u_{n+3} over u_{n+1}, u_{n+4} over
  // u_{n+2}: the pass writes the two newest fields into the twins of the
  // buffers they belong to.
  if (p.s % 2 == 0) {
    a.o1 = (char *)c->twin[vb] + o;
    a.o2 = (char *)c->twin[ub] + o;
  } else {
    a.o1 = (char *)c->twin[ub] + o;
    a.o2 = (char *)c->twin[vb] + o;
  }
  a.pitch = (int64_t)c->L.pitch;
  a.plane = (int64_t)c->L.plane;
  a.xs = (int)c->g.xs;
  a.ys = (int)c->g.ys;
  a.zs = (int)c->g.zs;
  a.ntk = p.ntk;
  a.ntj = p.ntj;
  a.chunks = p.chunks;
  a.chunk_len = p.chunk_len;
  a.remap = c->tune.xcd_remap ? 1 : 0;
  a.c0 = c0;
  a.c1 = c1;
  a.c2a = c2n ? c2n[0] : 0.0f;
  a.c2b = c2n ? c2n[1] : 0.0f;
  a.c2c = c2n && p.s >= 3 ? c2n[2] : 0.0f;
  a.c2d = c2n && p.s >= 4 ? c2n[3] : 0.0f;
  a.c2e = c2n && p.s >= 5 ? c2n[4] : 0.0f;
  memcpy(a.w, wts.w, sizeof(a.w));
  const bool h16 = is16(c);
  if (h16) {
    a.qs = c->qs_dev;
    a.slot = c->slot;
    a.ubuf = ub;
    a.vbuf = vb;
    a.gain = wts.gain;
  }
  const hipError_t e =
      h16 ? launch_sweep2h(c->fmt, (int)c->g.npts, p.mj, p.nw, p.s, p.grid, c->stream, a)
          : launch_sweep2((int)c->g.npts, p.mj, p.nw, p.s, p.grid, c->stream, a);
  if (e != hipSuccess) return hip_status(e);
  if (c->timing) c->launches++;
  std::swap(c->base[0], c->twin[0]);
  std::swap(c->base[1], c->twin[1]);
  if (p.s % 2) c->ubuf = vb;
  // FP16: a pass takes one slot of the scale ring (it publishes the state
  // after its last step)
  c->slot = (c->slot + (h16 ? 1 : p.s)) % 3;
  return STH_OK;
}

int run_leap(sth_ctx *c, const Plan &p, uint32_t steps, float c0, float c1, const float *c2) {
  int st = push_slot(c);
  if (st) return st;
  SweepArgs a;
  memset(&a, 0, sizeof(a));
  fill_weights(c, true, a);
  uint32_t n = 0;
  if (steps >= 2 && fused_applies(c, c->tune)) {
    st = ensure_twins(c);
    if (st == STH_OK) {
      const Plan2 p2 = make_plan2(c, c->tune);
      for (; n + p2.s <= steps; n += p2.s) {
        st = run_pass2(c, p2, a, c0, c1, c2 ? c2 + n : nullptr);
        if (st) return st;
      }
      // steps left over: the longest shorter pass that fits (down to two
      // steps), then one-step sweeps
      for (int rs = p2.s - 1; rs >= 2; rs--) {
        if (n + (uint32_t)rs > steps) continue;
        if (is16(c) && !q16_pass_exists(c->fmt, (int)c->g.npts, rs)) continue;
        sth_tuning t2 = c->tune;
        t2.steps_per_pass = (uint32_t)rs;
        t2.rows_per_wave = t2.waves = t2.chunks = 0;
        st = run_pass2(c, make_plan2(c, t2), a, c0, c1, c2 ? c2 + n : nullptr);
        if (st) return st;
        n += (uint32_t)rs;
      }
    } else if (c->tune.steps_per_pass >= 2) {
      return st;
    }
  }
  for (; n < steps; n++) {
    base_args(c, p, a);
    a.c0 = c0;
    a.c1 = c1;
    a.c2n = c2 ? c2[n] : 0.0f;
    st = launch_one(c, p, MODE_LEAP, a);
    if (st) return st;
    c->ubuf = 1 - c->ubuf;
    c->slot = (c->slot + 1) % 3;
  }
  return STH_OK;
}

}  // namespace detail
}  // namespace sth

using namespace sth::detail;

extern "C" {

const char *sth_strerror(int status) {
  switch (status) {
    case STH_OK: return "ok";
    case STH_ERR_ARG: return "invalid argument";
    case STH_ERR_HIP: return "HIP runtime error";
    case STH_ERR_NOMEM: return "device out of memory";
    case STH_ERR_STATE: return "invalid state for this call";
    case STH_ERR_NODEV: return "no usable GPU";
    case STH_ERR_UNSUPPORTED: return "unsupported combination";
    default: return "unknown status";
  }
}

int sth_version(void) { return STH_VERSION; }

int sth_layout(const sth_geom *g, sth_layout_info *out) {
  if (!g || !out) return STH_ERR_ARG;
  return compute_layout(*g, *out);
}

int sth_host_elems(const sth_geom *g, const sth_host_layout *hl, uint64_t *elems) {
  if (!g || !hl || !elems) return STH_ERR_ARG;
  *elems = ((uint64_t)g->xs + 2 * hl->xh) * ((uint64_t)g->ys + 2 * hl->yh) *
           ((uint64_t)g->zs + 2 * hl->zh);
  return STH_OK;
}

int sth_device_count(int *n) {
  if (!n) return STH_ERR_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return c > 0 ? STH_OK : STH_ERR_NODEV;
}

int sth_destroy(sth_ctx *c) {
  if (!c) return STH_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int b = 0; b < 3; b++)
    if (c->base[b]) (void)hipFree(c->base[b]);
  for (int b = 0; b < 2; b++)
    if (c->twin[b]) (void)hipFree(c->twin[b]);
  if (c->qs_dev) (void)hipFree(c->qs_dev);
  if (c->maxbits_dev) (void)hipFree(c->maxbits_dev);
  if (c->errors_dev) (void)hipFree(c->errors_dev);
  if (c->partial_dev) (void)hipFree(c->partial_dev);
  if (c->sum_dev) (void)hipFree(c->sum_dev);
  if (c->tickets_dev) (void)hipFree(c->tickets_dev);
  if (c->staging) (void)hipHostFree(c->staging);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return STH_OK;
}

int sth_create(const sth_geom *g, sth_ctx **out) {
  if (!g || !out) return STH_ERR_ARG;
  *out = nullptr;
  sth_layout_info L;
  int st = compute_layout(*g, L);
  if (st) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return STH_ERR_NODEV;
  sth_ctx *c = new (std::nothrow) sth_ctx();
  if (!c) return STH_ERR_NOMEM;
  c->g = *g;
  c->L = L;
  c->R = (int)L.radius;
  c->eb = (int)L.elem_bytes;
  c->fmt = g->dtype == STH_FP32   ? FMT_F32
           : g->dtype == STH_FP64 ? FMT_F64
           : g->dtype == STH_FP16 ? FMT_F16
                                  : FMT_BF16;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  for (int b = 0; b < 3 && e == hipSuccess; b++) e = hipMalloc(&c->base[b], L.bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&c->qs_dev, 3 * sizeof(Q16Slot));
  if (e == hipSuccess) e = hipMalloc((void **)&c->maxbits_dev, 256);
  if (e == hipSuccess) e = hipMalloc((void **)&c->errors_dev, 256);
  if (e == hipSuccess) e = hipMalloc((void **)&c->partial_dev, kReduceBlocks * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&c->sum_dev, 256);
  if (e == hipSuccess) e = hipMalloc((void **)&c->tickets_dev, 64 * kXcds);
  if (e == hipSuccess) e = hipMemsetAsync(c->tickets_dev, 0, 64 * kXcds, c->stream);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipMemsetAsync(c->qs_dev, 0, 3 * sizeof(Q16Slot), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    const int rc = hip_status(e);
    sth_destroy(c);
    return rc;
  }
  c->tune.xcd_remap = 1;
  *out = c;
  return STH_OK;
}

int sth_get_layout(const sth_ctx *c, sth_layout_info *out) {
  if (!c || !out) return STH_ERR_ARG;
  *out = c->L;
  return STH_OK;
}

int sth_init_const(sth_ctx *c, float v0, float s0, float u0) {
  if (!c) return STH_ERR_ARG;
  if (is16(c)) return STH_ERR_UNSUPPORTED;
  const float vals[3] = {v0, s0, u0};
  const sth_field fs[3] = {STH_FIELD_V, STH_FIELD_S, STH_FIELD_U};
  for (int q = 0; q < 3; q++) {
    FillSpec f;
    memset(&f, 0, sizeof(f));
    f.kind = FILL_CONST;
    f.a = vals[q];
    int st = device_fill(c, buf_of(c, fs[q]), f, vals[q] == 0.0f);
    if (st) return st;
  }
  return STH_OK;
}

int sth_init_uniform(sth_ctx *c, sth_field fld, uint64_t seed, int field_id, double lo, double hi,
                     float halo_value) {
  if (!c || fld < 0 || fld > 2) return STH_ERR_ARG;
  FillSpec f;
  memset(&f, 0, sizeof(f));
  f.kind = FILL_UNIFORM;
  f.lo = lo;
  f.hi = hi;
  f.halo = halo_value;
  f.key = field_key(seed, field_id);
  return device_fill(c, buf_of(c, fld), f, halo_value == 0.0f);
}

int sth_init_gaussian(sth_ctx *c, sth_field fld, double amp, double sigma, double noise,
                      uint64_t seed, int field_id) {
  if (!c || fld < 0 || fld > 2 || !(sigma > 0.0)) return STH_ERR_ARG;
  FillSpec f;
  memset(&f, 0, sizeof(f));
  f.kind = FILL_GAUSS;
  f.lo = amp;
  f.hi = sigma;
  f.noise = noise;
  f.key = field_key(seed, field_id);
  return device_fill(c, buf_of(c, fld), f, true);
}

int sth_init_point(sth_ctx *c, sth_field fld, uint32_t i, uint32_t j, uint32_t k, float value) {
  if (!c || fld < 0 || fld > 2) return STH_ERR_ARG;
  if (i >= c->g.xs || j >= c->g.ys || k >= c->g.zs) return STH_ERR_ARG;
  FillSpec f;
  memset(&f, 0, sizeof(f));
  f.kind = FILL_POINT;
  f.a = value;
  f.pi = (int)i;
  f.pj = (int)j;
  f.pk = (int)k;
  return device_fill(c, buf_of(c, fld), f, true);
}

}  // extern "C"

namespace {

// Per-thread partials of a transfer (one cache line each).
struct alignas(64) RowAcc {
  float mx = 0.0f;   // max |value| read
  float qmax = 0.0f; // 16-bit: max |stored value|
  bool hz = true;    // every halo cell read is zero
};

// Host (float or double) dense array -> device field of any storage type.
template <class HT>
int upload_any(sth_ctx *c, sth_field fld, const HT *host, const sth_host_layout *hl) {
  if (!c || !host || fld < 0 || fld > 2) return STH_ERR_ARG;
  int st = check_host_layout(c, hl);
  if (st) return st;
  const HostIdx h = host_idx(c, hl);
  const int64_t R = c->R, zs = c->g.zs, xs = c->g.xs, ys = c->g.ys, n = zs + 2 * R;
  const int buf = buf_of(c, fld);
  const bool q16 = is16(c);
  std::vector<RowAcc> acc(host_threads());
  // a nonzero value at a halo cell the sweep reads clears the halo-zero flag
  auto halo_scan = [&](int64_t i, int64_t j, const HT *src, RowAcc &a) {
    if (!a.hz) return;
    const bool inrow = i >= 0 && i < xs && j >= 0 && j < ys;
    for (int64_t k = 0; k < n; k++)
      if (src[k] != 0 && (!inrow || k < R || k >= zs + R)) {
        a.hz = false;
        return;
      }
  };
  int e = 0;
  if (q16) {
    // 16-bit: the max over the cells the sweep reads sets the exponent
    // before any value is encoded
    parallel_rows(-R, xs + R, R, ys, [&](int64_t i, int64_t j, int t) {
      const HT *src = host + h.ofs + i * h.im + j * h.jm - R;
      RowAcc &a = acc[t];
      float m = a.mx;
      for (int64_t k = 0; k < n; k++) m = std::max(m, fabsf((float)src[k]));
      a.mx = m;
      halo_scan(i, j, src, a);
    });
    float mx = 0.0f;
    for (const RowAcc &a : acc) mx = std::max(mx, a.mx);
    st = fetch_slot(c);
    if (st) return st;
    e = scale_exp((double)mx);
  }
  st = transfer_planes(c, buf, true, [&](int64_t i, int64_t j, char *srow, int t) {
    const HT *src = host + h.ofs + i * h.im + j * h.jm - R;
    RowAcc &a = acc[t];
    if (c->fmt == FMT_F32) {
      float *d = (float *)srow;
      for (int64_t k = 0; k < n; k++) d[k] = (float)src[k];
      halo_scan(i, j, src, a);
    } else if (c->fmt == FMT_F64) {
      double *d = (double *)srow;
      for (int64_t k = 0; k < n; k++) d[k] = (double)src[k];
      halo_scan(i, j, src, a);
    } else {
      uint16_t *d = (uint16_t *)srow;
      float q = a.qmax;
      for (int64_t k = 0; k < n; k++) {
        d[k] = enc16(c->fmt, (float)src[k], e);
        q = std::max(q, fabsf(dec16(c->fmt, d[k], 0)));
      }
      a.qmax = q;
    }
  });
  if (st) return st;
  bool hz = true;
  float qmax = 0.0f;
  for (const RowAcc &a : acc) {
    hz = hz && a.hz;
    qmax = std::max(qmax, a.qmax);
  }
  if (q16) {
    c->host_slot.exp[buf] = e;
    c->host_slot.maxbits[buf] = f2u(qmax);
  }
  c->halo_zero[buf] = hz;
  return STH_OK;
}

// Device field of any storage type -> host (float or double) dense array.
template <class HT>
int download_any(sth_ctx *c, sth_field fld, HT *host, const sth_host_layout *hl) {
  if (!c || !host || fld < 0 || fld > 2) return STH_ERR_ARG;
  int st = check_host_layout(c, hl);
  if (st) return st;
  st = fetch_slot(c);
  if (st) return st;
  const HostIdx h = host_idx(c, hl);
  const int64_t R = c->R, zs = c->g.zs;
  const int buf = buf_of(c, fld);
  const int e = is16(c) ? c->host_slot.exp[buf] : 0;
  return transfer_planes(c, buf, false, [&](int64_t i, int64_t j, char *srow, int) {
    HT *dst = host + h.ofs + i * h.im + j * h.jm - R;
    const int64_t n = zs + 2 * R;
    if (c->fmt == FMT_F32) {
      const float *s = (const float *)srow;
      for (int64_t k = 0; k < n; k++) dst[k] = (HT)s[k];
    } else if (c->fmt == FMT_F64) {
      const double *s = (const double *)srow;
      for (int64_t k = 0; k < n; k++) dst[k] = (HT)s[k];
    } else {
      const uint16_t *s = (const uint16_t *)srow;
      for (int64_t k = 0; k < n; k++) dst[k] = (HT)dec16(c->fmt, s[k], e);
    }
  });
}

}  // namespace

extern "C" {

int sth_upload(sth_ctx *c, sth_field fld, const float *host, const sth_host_layout *hl) {
  return upload_any<float>(c, fld, host, hl);
}

int sth_download(sth_ctx *c, sth_field fld, float *host, const sth_host_layout *hl) {
  return download_any<float>(c, fld, host, hl);
}

int sth_upload_f64(sth_ctx *c, sth_field fld, const double *host, const sth_host_layout *hl) {
  if (c && is16(c)) return STH_ERR_UNSUPPORTED;
  return upload_any<double>(c, fld, host, hl);
}

int sth_download_f64(sth_ctx *c, sth_field fld, double *host, const sth_host_layout *hl) {
  return download_any<double>(c, fld, host, hl);
}

int sth_upload_raw16(sth_ctx *c, sth_field fld, const uint16_t *host, int exp,
                     const sth_host_layout *hl) {
  if (!c || !host || fld < 0 || fld > 2) return STH_ERR_ARG;
  if (!is16(c)) return STH_ERR_UNSUPPORTED;
  int st = check_host_layout(c, hl);
  if (st) return st;
  st = fetch_slot(c);
  if (st) return st;
  const HostIdx h = host_idx(c, hl);
  const int64_t R = c->R, zs = c->g.zs, xs = c->g.xs, ys = c->g.ys;
  const int buf = buf_of(c, fld);
  std::vector<RowAcc> acc(host_threads());
  st = transfer_planes(c, buf, true, [&](int64_t i, int64_t j, char *srow, int t) {
    const uint16_t *src = host + h.ofs + i * h.im + j * h.jm - R;
    memcpy(srow, src, (size_t)(zs + 2 * R) * 2);
    const bool inrow = i >= 0 && i < xs && j >= 0 && j < ys;
    RowAcc &a = acc[t];
    for (int64_t k = 0; k < zs + 2 * R; k++) {
      const float x = dec16(c->fmt, src[k], 0);
      a.qmax = std::max(a.qmax, fabsf(x));
      if (x != 0.0f && (!inrow || k < R || k >= zs + R)) a.hz = false;
    }
  });
  if (st) return st;
  float qmax = 0.0f;
  bool hz = true;
  for (const RowAcc &a : acc) {
    hz = hz && a.hz;
    qmax = std::max(qmax, a.qmax);
  }
  c->host_slot.exp[buf] = exp;
  c->host_slot.maxbits[buf] = f2u(qmax);
  c->halo_zero[buf] = hz;
  return STH_OK;
}

int sth_download_raw16(sth_ctx *c, sth_field fld, uint16_t *host, int *exp, float *maxabs,
                       const sth_host_layout *hl) {
  if (!c || !host || fld < 0 || fld > 2) return STH_ERR_ARG;
  if (!is16(c)) return STH_ERR_UNSUPPORTED;
  int st = check_host_layout(c, hl);
  if (st) return st;
  st = fetch_slot(c);
  if (st) return st;
  const HostIdx h = host_idx(c, hl);
  const int64_t R = c->R, zs = c->g.zs;
  const int buf = buf_of(c, fld);
  if (exp) *exp = c->host_slot.exp[buf];
  if (maxabs) *maxabs = u2f(c->host_slot.maxbits[buf]);
  return transfer_planes(c, buf, false, [&](int64_t i, int64_t j, char *srow, int) {
    memcpy(host + h.ofs + i * h.im + j * h.jm - R, srow, (size_t)(zs + 2 * R) * 2);
  });
}

// ---- field I/O -------------------------------------------------------------
// A directory with manifest.json and one raw file per field (u.raw, v.raw,
// s.raw): the storage elements (float32, float64 or the 16-bit numbers) of the
// dense box with an R-wide halo, C order, k fastest, little-endian.

}  // extern "C"

namespace {

const char *dtype_name(uint32_t d) {
  return d == STH_FP32 ? "fp32" : d == STH_FP64 ? "fp64" : d == STH_FP16 ? "fp16" : "bf16";
}

std::string path_join(const char *dir, const char *f) {
  std::string p(dir);
  if (!p.empty() && p.back() != '/') p += '/';
  return p + f;
}

// Reads an integer value following "key": in a manifest written by sth_save.
bool json_int(const std::string &txt, const char *key, long long &out, size_t from = 0) {
  const std::string k = std::string("\"") + key + "\":";
  const size_t at = txt.find(k, from);
  if (at == std::string::npos) return false;
  return sscanf(txt.c_str() + at + k.size(), " %lld", &out) == 1;
}

template <class T>
int save_field(sth_ctx *c, sth_field f, const std::string &path, int *exp, float *mx) {
  const sth_host_layout hl = {(uint32_t)c->R, (uint32_t)c->R, (uint32_t)c->R};
  uint64_t n = 0;
  sth_host_elems(&c->g, &hl, &n);
  std::vector<T> buf(n);
  int st;
  if constexpr (sizeof(T) == 2)
    st = sth_download_raw16(c, f, (uint16_t *)buf.data(), exp, mx, &hl);
  else
    st = download_any<T>(c, f, buf.data(), &hl);
  if (st) return st;
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) return STH_ERR_ARG;
  const size_t w = fwrite(buf.data(), sizeof(T), buf.size(), fp);
  fclose(fp);
  return w == buf.size() ? STH_OK : STH_ERR_ARG;
}

template <class T>
int load_field(sth_ctx *c, sth_field f, const std::string &path, int exp) {
  const sth_host_layout hl = {(uint32_t)c->R, (uint32_t)c->R, (uint32_t)c->R};
  uint64_t n = 0;
  sth_host_elems(&c->g, &hl, &n);
  std::vector<T> buf(n);
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) return STH_ERR_ARG;
  const size_t r = fread(buf.data(), sizeof(T), buf.size(), fp);
  fclose(fp);
  if (r != buf.size()) return STH_ERR_ARG;
  if constexpr (sizeof(T) == 2)
    return sth_upload_raw16(c, f, (const uint16_t *)buf.data(), exp, &hl);
  else
    return upload_any<T>(c, f, buf.data(), &hl);
}

}  // namespace

extern "C" {

int sth_save(sth_ctx *c, const char *dir) {
  if (!c || !dir) return STH_ERR_ARG;
  static const char *names[3] = {"v", "s", "u"};
  int exps[3] = {0, 0, 0};
  float mxs[3] = {0, 0, 0};
  for (int f = 0; f < 3; f++) {
    const std::string path = path_join(dir, (std::string(names[f]) + ".raw").c_str());
    int st;
    if (c->fmt == FMT_F32) st = save_field<float>(c, (sth_field)f, path, &exps[f], &mxs[f]);
    else if (c->fmt == FMT_F64) st = save_field<double>(c, (sth_field)f, path, &exps[f], &mxs[f]);
    else st = save_field<uint16_t>(c, (sth_field)f, path, &exps[f], &mxs[f]);
    if (st) return st;
  }
  FILE *fp = fopen(path_join(dir, "manifest.json").c_str(), "w");
  if (!fp) return STH_ERR_ARG;
  fprintf(fp,
          "{\n \"format\": \"stencil3d_amd fields\",\n \"version\": 1,\n"
          " \"xs\": %u,\n \"ys\": %u,\n \"zs\": %u,\n \"npts\": %u,\n \"halo\": %d,\n"
          " \"dtype\": \"%s\",\n \"elem_bytes\": %d,\n"
          " \"layout\": \"dense (xs+2*halo) x (ys+2*halo) x (zs+2*halo), C order, k fastest, "
          "little-endian; 16-bit values are q with value = q * 2^exp\",\n"
          " \"fields\": {\n",
          c->g.xs, c->g.ys, c->g.zs, c->g.npts, c->R, dtype_name(c->g.dtype), c->eb);
  for (int f = 0; f < 3; f++)
    fprintf(fp, "  \"%s\": {\"file\": \"%s.raw\", \"exp\": %d, \"maxabs_q\": %.9g}%s\n", names[f],
            names[f], exps[f], (double)mxs[f], f < 2 ? "," : "");
  fprintf(fp, " }\n}\n");
  fclose(fp);
  return STH_OK;
}

int sth_load(sth_ctx *c, const char *dir) {
  if (!c || !dir) return STH_ERR_ARG;
  FILE *fp = fopen(path_join(dir, "manifest.json").c_str(), "r");
  if (!fp) return STH_ERR_ARG;
  std::string txt;
  char chunk[4096];
  size_t r;
  while ((r = fread(chunk, 1, sizeof(chunk), fp)) > 0) txt.append(chunk, r);
  fclose(fp);
  long long xs, ys, zs, npts, halo;
  if (!json_int(txt, "xs", xs) || !json_int(txt, "ys", ys) || !json_int(txt, "zs", zs) ||
      !json_int(txt, "npts", npts) || !json_int(txt, "halo", halo))
    return STH_ERR_ARG;
  const std::string want = std::string("\"dtype\": \"") + dtype_name(c->g.dtype) + "\"";
  if (xs != c->g.xs || ys != c->g.ys || zs != c->g.zs || npts != c->g.npts || halo != c->R ||
      txt.find(want) == std::string::npos)
    return STH_ERR_ARG;
  static const char *names[3] = {"v", "s", "u"};
  for (int f = 0; f < 3; f++) {
    long long e = 0;
    const size_t at = txt.find(std::string("\"") + names[f] + "\": {");
    if (at == std::string::npos || !json_int(txt, "exp", e, at)) return STH_ERR_ARG;
    const std::string path = path_join(dir, (std::string(names[f]) + ".raw").c_str());
    int st;
    if (c->fmt == FMT_F32) st = load_field<float>(c, (sth_field)f, path, 0);
    else if (c->fmt == FMT_F64) st = load_field<double>(c, (sth_field)f, path, 0);
    else st = load_field<uint16_t>(c, (sth_field)f, path, (int)e);
    if (st) return st;
  }
  return STH_OK;
}

int sth_iterate(sth_ctx *c, uint32_t iterations) {
  if (!c) return STH_ERR_ARG;
  if (is16(c)) return STH_ERR_UNSUPPORTED;
  return run_ref(c, make_plan(c, c->tune), iterations);
}

int sth_leapfrog(sth_ctx *c, uint32_t steps, float c0, float c1, const float *c2) {
  if (!c) return STH_ERR_ARG;
  if (is16(c) && !(c->halo_zero[0] && c->halo_zero[1])) return STH_ERR_UNSUPPORTED;
  return run_leap(c, make_plan(c, c->tune), steps, c0, c1, c2);
}

int sth_leapfrog_c0(uint32_t npts, float c1, float *c0) {
  if (!c0 || radius_of(npts) < 0) return STH_ERR_ARG;
  *c0 = (float)(2.0 - lap_gain(npts) * (double)c1);
  return STH_OK;
}

int sth_ricker(uint32_t steps, double dt, double f, double t0, double scale, float *c2) {
  if (!c2 && steps) return STH_ERR_ARG;
  for (uint32_t n = 0; n < steps; n++) {
    const double tau = (double)n * dt - t0;
    const double a = M_PI * M_PI * f * f * tau * tau;
    c2[n] = (float)(scale * ((1.0 - 2.0 * a) * exp(-a)));
  }
  return STH_OK;
}

int sth_sync(sth_ctx *c) {
  if (!c) return STH_ERR_ARG;
  STH_TRY(hipStreamSynchronize(c->stream));
  return STH_OK;
}

int sth_check(sth_ctx *c, float expected, uint64_t *errors) {
  if (!c || !errors) return STH_ERR_ARG;
  int st = fetch_slot(c);
  if (st) return st;
  const int buf = 1 - c->ubuf;
  const int e = is16(c) ? c->host_slot.exp[buf] : 0;
  unsigned long long cnt = 0;
  STH_TRY(hipMemsetAsync(c->errors_dev, 0, 8, c->stream));
  STH_TRY(launch_check(desc_of(c), c->base[buf], c->fmt, e, expected, c->errors_dev, c->stream));
  STH_TRY(hipMemcpyAsync(&cnt, c->errors_dev, 8, hipMemcpyDeviceToHost, c->stream));
  STH_TRY(hipStreamSynchronize(c->stream));
  *errors = cnt;
  return STH_OK;
}

int sth_norms(sth_ctx *c, sth_field fld, double *maxabs, double *l2) {
  if (!c || fld < 0 || fld > 2) return STH_ERR_ARG;
  int st = fetch_slot(c);
  if (st) return st;
  const int buf = buf_of(c, fld);
  const int e = is16(c) ? c->host_slot.exp[buf] : 0;
  uint32_t mb[4] = {0, 0, 0, 0};
  double ss = 0.0;
  STH_TRY(launch_zero_u32(c->maxbits_dev, 4, c->stream));
  STH_TRY(launch_norms(desc_of(c), c->base[buf], c->fmt, e, c->partial_dev, c->maxbits_dev,
                       c->sum_dev, c->stream));
  STH_TRY(hipMemcpyAsync(mb, c->maxbits_dev, 16, hipMemcpyDeviceToHost, c->stream));
  STH_TRY(hipMemcpyAsync(&ss, c->sum_dev, 8, hipMemcpyDeviceToHost, c->stream));
  STH_TRY(hipStreamSynchronize(c->stream));
  double m64;
  memcpy(&m64, &mb[2], 8);
  if (maxabs) *maxabs = m64;
  if (l2) *l2 = sqrt(ss);
  return STH_OK;
}

int sth_energy(sth_ctx *c, float c0, float c1, double *energy) {
  if (!c || !energy) return STH_ERR_ARG;
  int st = fetch_slot(c);
  if (st) return st;
  SweepArgs a{};
  fill_weights(c, true, a);  // the LEAPFROG weights (2w) the context's steps use
  int ax[12], rr[12];
  EnergySpec sp{};
  sp.T = term_order(c->g.npts, ax, rr);
  const int64_t stride[3] = {(int64_t)c->L.plane, (int64_t)c->L.pitch, 1};
  for (int t = 0; t < sp.T; t++) {
    sp.off[t] = stride[ax[t]] * rr[t];
    sp.w[t] = c->fmt == FMT_F64 ? a.wd[t] : (double)a.w[t];
  }
  sp.c0 = (double)c0;
  sp.c1 = (double)c1;
  const int bu = buf_of(c, STH_FIELD_U), bv = buf_of(c, STH_FIELD_V);
  const int eu = is16(c) ? c->host_slot.exp[bu] : 0, ev = is16(c) ? c->host_slot.exp[bv] : 0;
  double q = 0.0;
  STH_TRY(launch_energy(desc_of(c), c->base[bu], c->base[bv], c->fmt, eu, ev, sp, c->partial_dev,
                        c->sum_dev, c->stream));
  STH_TRY(hipMemcpyAsync(&q, c->sum_dev, 8, hipMemcpyDeviceToHost, c->stream));
  STH_TRY(hipStreamSynchronize(c->stream));
  *energy = q;
  return STH_OK;
}

int sth_compare(sth_ctx *a, sth_field fa, sth_ctx *b, sth_field fb, double *max_rel,
                double *l2_rel) {
  if (!a || !b || fa < 0 || fa > 2 || fb < 0 || fb > 2 || a == b) return STH_ERR_ARG;
  if (a->g.xs != b->g.xs || a->g.ys != b->g.ys || a->g.zs != b->g.zs || a->g.npts != b->g.npts ||
      a->device != b->device)
    return STH_ERR_ARG;
  int st = fetch_slot(a);
  if (!st) st = fetch_slot(b);
  if (st) return st;
  const int xa = buf_of(a, fa), xb = buf_of(b, fb);
  const int ea = is16(a) ? a->host_slot.exp[xa] : 0, eb = is16(b) ? b->host_slot.exp[xb] : 0;
  STH_TRY(hipStreamSynchronize(b->stream));  // b's field is complete
  // a's scratch holds the difference's partials, b's those of the reference
  STH_TRY(launch_zero_u32(a->maxbits_dev, 4, a->stream));
  STH_TRY(launch_zero_u32(b->maxbits_dev, 4, a->stream));
  STH_TRY(launch_compare(desc_of(a), a->base[xa], a->fmt, ea, desc_of(b), b->base[xb], b->fmt, eb,
                         a->partial_dev, b->partial_dev, a->maxbits_dev, b->maxbits_dev,
                         a->sum_dev, a->stream));
  uint32_t md[4] = {0, 0, 0, 0}, mr[4] = {0, 0, 0, 0};
  double sums[2] = {0.0, 0.0};
  STH_TRY(hipMemcpyAsync(md, a->maxbits_dev, 16, hipMemcpyDeviceToHost, a->stream));
  STH_TRY(hipMemcpyAsync(mr, b->maxbits_dev, 16, hipMemcpyDeviceToHost, a->stream));
  STH_TRY(hipMemcpyAsync(sums, a->sum_dev, 16, hipMemcpyDeviceToHost, a->stream));
  STH_TRY(hipStreamSynchronize(a->stream));
  double dmax, rmax;
  memcpy(&dmax, &md[2], 8);
  memcpy(&rmax, &mr[2], 8);
  if (max_rel) *max_rel = rmax > 0.0 ? dmax / rmax : dmax;
  if (l2_rel) *l2_rel = sums[1] > 0.0 ? sqrt(sums[0] / sums[1]) : sqrt(sums[0]);
  return STH_OK;
}

int sth_set_weights(sth_ctx *c, int mode) {
  if (!c || (mode != STH_WEIGHTS_FP32 && mode != STH_WEIGHTS_REFERENCE)) return STH_ERR_ARG;
  if (mode == STH_WEIGHTS_REFERENCE && c->fmt != FMT_F64) return STH_ERR_UNSUPPORTED;
  c->weights = mode;
  return STH_OK;
}

int sth_get_weights(const sth_ctx *c, int *mode) {
  if (!c || !mode) return STH_ERR_ARG;
  *mode = c->weights;
  return STH_OK;
}

int sth_set_tuning(sth_ctx *c, const sth_tuning *t) {
  if (!c || !t || t->steps_per_pass > 5 || t->launch > 2 || t->chunks > 64 ||
      t->ref_sweeps > 4096)
    return STH_ERR_ARG;
  if (t->steps_per_pass >= 2 && c->fmt != FMT_F32 &&
      !(is16(c) && q16_pass_exists(c->fmt, (int)c->g.npts, (int)t->steps_per_pass)))
    return STH_ERR_UNSUPPORTED;
  if ((t->steps_per_pass == 3 && c->g.npts == 25) || (t->steps_per_pass >= 4 && c->g.npts != 7))
    return STH_ERR_UNSUPPORTED;
  const bool fused_shape =
      t->steps_per_pass >= 2 &&
      (is16(c)
           ? fused2h_supported(c->fmt, (int)c->g.npts, (int)t->rows_per_wave, (int)t->waves,
                               (int)t->steps_per_pass)
           : fused2_supported((int)c->g.npts, (int)t->rows_per_wave, (int)t->waves,
                              (int)t->steps_per_pass));
  const TileShape ts = default_tile(c->fmt, c->g.ys, c->g.zs, *t);
  if (!fused_shape && !shape_supported(c, ts.mj, ts.nw, ts.vec)) return STH_ERR_ARG;
  c->tune = *t;
  return STH_OK;
}

int sth_get_tuning(const sth_ctx *c, sth_tuning *t) {
  if (!c || !t) return STH_ERR_ARG;
  *t = c->tune;
  if (fused_applies(c, c->tune)) {
    const Plan2 p = make_plan2(c, c->tune);
    t->rows_per_wave = (uint32_t)p.mj;
    t->waves = (uint32_t)p.nw;
    t->vec = 1;
    t->chunks = (uint32_t)p.chunks;
    t->steps_per_pass = (uint32_t)p.s;
    return STH_OK;
  }
  const Plan p = make_plan(c, c->tune);
  t->rows_per_wave = (uint32_t)p.mj;
  t->waves = (uint32_t)p.nw;
  t->vec = (uint32_t)p.vec;
  t->chunks = (uint32_t)p.chunks;
  t->steps_per_pass = 1;
  t->launch = p.persist ? 2 : 1;
  return STH_OK;
}

}  // extern "C"

namespace {

// Tuning cache: one line per measured geometry,
//   sth-tune 2 <npts> <dtype> <leapfrog> <xs> <ys> <zs> <rows> <waves> <vec> <chunks>
//              <steps_per_pass> <ms per step>
// the last matching line wins.
bool tune_cache_lookup(const char *path, const sth_ctx *c, int leap, sth_tuning &t, float &ms) {
  FILE *fp = fopen(path, "r");
  if (!fp) return false;
  char line[256];
  bool found = false;
  while (fgets(line, sizeof(line), fp)) {
    unsigned ver, npts, dt, lf, xs, ys, zs, mj, nw, vec, ch, spp;
    float m;
    if (sscanf(line, "sth-tune %u %u %u %u %u %u %u %u %u %u %u %u %f", &ver, &npts, &dt, &lf,
               &xs, &ys, &zs, &mj, &nw, &vec, &ch, &spp, &m) != 13)
      continue;
    if (ver != 2 || npts != c->g.npts || dt != c->g.dtype || lf != (unsigned)leap ||
        xs != c->g.xs || ys != c->g.ys || zs != c->g.zs || spp < 1 || spp > 5)
      continue;
    const bool ok =
        spp >= 2 ? (!c->no_fuse &&
                    (is16(c)
                         ? fused2h_supported(c->fmt, (int)npts, (int)mj, (int)nw, (int)spp)
                         : c->fmt == FMT_F32 &&
                               fused2_supported((int)npts, (int)mj, (int)nw, (int)spp)))
                 : shape_supported(c, (int)mj, (int)nw, (int)vec);
    if (!ok) continue;
    t.rows_per_wave = mj;
    t.waves = nw;
    t.vec = vec;
    t.chunks = ch;
    t.steps_per_pass = spp;
    ms = m;
    found = true;
  }
  fclose(fp);
  return found;
}

void tune_cache_store(const char *path, const sth_ctx *c, int leap, const sth_tuning &t, float ms) {
  FILE *fp = fopen(path, "a");
  if (!fp) return;
  fprintf(fp, "sth-tune 2 %u %u %d %u %u %u %u %u %u %u %u %.5f\n", c->g.npts, c->g.dtype, leap,
          c->g.xs, c->g.ys, c->g.zs, t.rows_per_wave, t.waves, t.vec, t.chunks, t.steps_per_pass,
          ms);
  fclose(fp);
}

}  // namespace

extern "C" {

int sth_autotune(sth_ctx *c, int leapfrog, uint32_t reps, sth_tuning *best, float *best_ms) {
  if (!c || reps == 0) return STH_ERR_ARG;
  if (!leapfrog && is16(c)) return STH_ERR_UNSUPPORTED;
  leapfrog = leapfrog ? 1 : 0;
  const char *cache = getenv("STH_TUNE_CACHE");
  if (cache && *cache) {
    sth_tuning t = c->tune;
    float ms = 0.0f;
    if (tune_cache_lookup(cache, c, leapfrog, t, ms)) {
      c->tune = t;
      if (best) *best = t;
      if (best_ms) *best_ms = ms;
      return STH_OK;
    }
  }
  static const int tiles[][3] = {{1, 4, 1}, {2, 4, 1}, {4, 4, 1}, {8, 4, 1}, {2, 8, 1},
                                 {4, 8, 1}, {4, 2, 1}, {1, 4, 2}, {2, 4, 2}, {4, 4, 2},
                                 {2, 8, 2}, {4, 8, 2}, {2, 16, 2}, {3, 8, 2}, {3, 16, 2}};
  static const int chunk_opts[] = {0, 8, 16, 32};
  static const int fused_shapes[][3] = {{3, 16, 2}, {4, 16, 2}, {4, 12, 2}, {4, 8, 2},
                                        {3, 8, 2},  {3, 16, 3}, {4, 12, 3}, {4, 10, 4},
                                        {5, 8, 4},  {2, 16, 4}, {2, 16, 5}};
  static const int fused_chunks[] = {0, 2, 4, 8};
  const sth_tuning saved = c->tune;
  std::vector<sth_tuning> cand;
  for (const auto &tl : tiles)
    for (int ch : chunk_opts) {
      sth_tuning t = saved;
      t.rows_per_wave = (uint32_t)tl[0];
      t.waves = (uint32_t)tl[1];
      t.vec = (uint32_t)tl[2];
      t.chunks = (uint32_t)ch;
      t.steps_per_pass = 1;
      if (!shape_supported(c, tl[0], tl[1], tl[2])) continue;
      const Plan p = make_plan(c, t);
      if (ch != 0 && p.chunks != ch) continue;
      cand.push_back(t);
    }
  // 16-bit storage: the passes' shapes
  if (leapfrog && is16(c) && !c->no_fuse)
    for (int sp = 2; sp <= 5; sp++)
      for (int mj : {2, 3, 4})
        for (int ch : {0, 4, 8, 16}) {
          if (!fused2h_supported(c->fmt, (int)c->g.npts, mj, 16, sp)) continue;
          sth_tuning t = saved;
          t.rows_per_wave = (uint32_t)mj;
          t.waves = 16;
          t.vec = 2;
          t.chunks = (uint32_t)ch;
          t.steps_per_pass = (uint32_t)sp;
          const Plan2 p = make_plan2(c, t);
          if (ch != 0 && p.chunks != ch) continue;
          cand.push_back(t);
        }
  if (leapfrog && c->fmt == FMT_F32 && !c->no_fuse)
    for (const auto &fs : fused_shapes) {
      if (!fused2_supported((int)c->g.npts, fs[0], fs[1], fs[2])) continue;
      for (int ch : fused_chunks) {
        sth_tuning t = saved;
        t.rows_per_wave = (uint32_t)fs[0];
        t.waves = (uint32_t)fs[1];
        t.vec = 1;
        t.chunks = (uint32_t)ch;
        t.steps_per_pass = (uint32_t)fs[2];
        const Plan2 p = make_plan2(c, t);
        if (ch != 0 && p.chunks != ch) continue;
        cand.push_back(t);
      }
    }
  const bool saved_timing = c->timing;
  c->timing = false;
  sth_tuning bt = saved;
  float bms = 1e30f;
  // steps per timed run: 2 reps rounded up to a multiple of 60, so passes
  // of two to five steps run whole
  const uint32_t nsteps = (2 * reps + 59) / 60 * 60;
  for (const sth_tuning &t : cand) {
    c->tune = t;
    const Plan p = make_plan(c, t);
    int st = leapfrog ? run_leap(c, p, 12, 1.0f, 0.0f, nullptr) : run_ref(c, p, 1);
    if (st) { c->tune = saved; c->timing = saved_timing; return st; }
    STH_TRY(hipEventRecord(c->ev0, c->stream));
    st = leapfrog ? run_leap(c, p, nsteps, 1.0f, 0.0f, nullptr) : run_ref(c, p, nsteps);
    if (st) { c->tune = saved; c->timing = saved_timing; return st; }
    STH_TRY(hipEventRecord(c->ev1, c->stream));
    STH_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    STH_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ms /= (float)nsteps;
    if (ms < bms) {
      bms = ms;
      bt = t;
      bt.chunks = (uint32_t)(t.steps_per_pass >= 2 ? make_plan2(c, t).chunks : p.chunks);
    }
  }
  c->tune = bt;
  c->timing = saved_timing;
  if (cache && *cache) tune_cache_store(cache, c, leapfrog, bt, bms);
  if (best) *best = bt;
  if (best_ms) *best_ms = bms;
  return STH_OK;
}

int sth_stream_copy(sth_ctx *c, uint64_t bytes, uint32_t reps, double *ms) {
  if (!c || reps == 0 || bytes == 0 || bytes > c->L.bytes || bytes % 16 != 0) return STH_ERR_ARG;
  void *src = c->base[c->ubuf], *dst = c->base[1 - c->ubuf];
  c->twins_ok = false;
  STH_TRY(launch_stream_copy(src, dst, bytes, c->stream));  // warm-up
  STH_TRY(hipEventRecord(c->ev0, c->stream));
  for (uint32_t r = 0; r < reps; r++) STH_TRY(launch_stream_copy(src, dst, bytes, c->stream));
  STH_TRY(hipEventRecord(c->ev1, c->stream));
  STH_TRY(hipEventSynchronize(c->ev1));
  float t = 0.0f;
  STH_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
  if (ms) *ms = (double)t / reps;
  return STH_OK;
}

int sth_bw_probe(sth_ctx *c, uint32_t reads, uint32_t writes, uint32_t lane_bytes,
                 uint64_t bytes, uint32_t grid, uint32_t reps, double *ms) {
  if (!c || reps == 0 || bytes == 0 || reads > 3 || writes > 2 || reads + writes == 0 ||
      (lane_bytes != 4 && lane_bytes != 16) || bytes % 16 != 0 || bytes > c->L.bytes)
    return STH_ERR_ARG;
  if (writes == 2) {
    int st = ensure_twins(c);
    if (st) return st;
  }
  // read u, s, v; write v (in place, as the one-step sweep does) and the
  // twin of u (the fused pass's second output)
  const void *r[3] = {c->base[c->ubuf], c->base[2], c->base[1 - c->ubuf]};
  void *w[2] = {c->base[1 - c->ubuf], writes == 2 ? c->twin[c->ubuf] : nullptr};
  c->twins_ok = false;
  const unsigned g = grid ? grid : (unsigned)std::max(1, c->cus) * 8;
  float *sink = (float *)c->sum_dev;
  STH_TRY(launch_bw_probe((int)reads, (int)writes, (int)lane_bytes, r, w, bytes, sink, g,
                          c->stream));  // warm-up
  STH_TRY(hipEventRecord(c->ev0, c->stream));
  for (uint32_t q = 0; q < reps; q++)
    STH_TRY(launch_bw_probe((int)reads, (int)writes, (int)lane_bytes, r, w, bytes, sink, g,
                            c->stream));
  STH_TRY(hipEventRecord(c->ev1, c->stream));
  STH_TRY(hipEventSynchronize(c->ev1));
  float t = 0.0f;
  STH_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
  if (ms) *ms = (double)t / reps;
  return STH_OK;
}

// ---- placement of the field buffers ------------------------------------------

// The LEAPFROG time of one candidate set of field buffers (a, b, s and, when
// the pass is fused, the twins of a and b): ms per step over four passes
// after one, on whatever the buffers hold.
static int time_fields(sth_ctx *c, void *const set[5], bool twins, float *ms) {
  for (int i = 0; i < 3; i++) c->base[i] = set[i];
  c->twin[0] = twins ? set[3] : nullptr;
  c->twin[1] = twins ? set[4] : nullptr;
  c->ubuf = 0;
  c->slot = 0;
  c->twins_ok = true;  // the halos do not matter to the time
  if (is16(c)) {
    // a scale state for the timing (the work does not depend on the values)
    for (int b = 0; b < 3; b++) {
      c->host_slot.exp[b] = 0;
      c->host_slot.maxbits[b] = 0x3f800000u;  // 1.0
    }
    c->host_dirty = true;
  }
  const float c1 = 0.1f, c0 = (float)(2.0 - lap_gain(c->g.npts) * c1);
  const uint32_t spp = twins ? (uint32_t)make_plan2(c, c->tune).s : 1;
  const Plan p = make_plan(c, c->tune);
  int st = run_leap(c, p, spp, c0, c1, nullptr);
  if (st) return st;
  STH_TRY(hipEventRecord(c->ev0, c->stream));
  st = run_leap(c, p, 4 * spp, c0, c1, nullptr);
  if (st) return st;
  STH_TRY(hipEventRecord(c->ev1, c->stream));
  STH_TRY(hipEventSynchronize(c->ev1));
  float t = 0.0f;
  STH_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / (float)(4 * spp);
  return STH_OK;
}

int sth_place(sth_ctx *c, uint32_t tries, float *ms, uint32_t *chosen) {
  if (!c || tries > 8) return STH_ERR_ARG;
  if (c->no_fuse) return STH_ERR_STATE;  // a batch's shot: its tables hold the buffers
  const bool twins = fused_applies(c, c->tune);
  const int nb = twins ? 5 : 3;
  if (twins) {
    const int st = ensure_twins(c);
    if (st) return st;
  } else {
    // twins left from an earlier tuning are rebuilt when next needed
    for (int b = 0; b < 2; b++)
      if (c->twin[b]) {
        (void)hipFree(c->twin[b]);
        c->twin[b] = nullptr;
      }
  }
  // Candidates only where they fit, leaving a sixteenth of the device (at
  // least 4 GiB) free.
  size_t fr = 0, tot = 0;
  STH_TRY(hipMemGetInfo(&fr, &tot));
  const size_t spare = std::max<size_t>(tot / 16, (size_t)4 << 30), per = nb * c->L.bytes;
  uint32_t n = 0;
  while (n < tries && (size_t)(n + 1) * per + spare <= fr) n++;
  std::vector<std::array<void *, 5>> sets(n + 1);
  sets[0] = {c->base[0], c->base[1], c->base[2], c->twin[0], c->twin[1]};
  // Every candidate stays allocated until the choice is made: a buffer freed
  // and allocated again tends to come back on the same memory.
  uint32_t got = 0;
  for (uint32_t t = 1; t <= n; t++) {
    sets[t] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool ok = true;
    for (int i = 0; i < nb && ok; i++)
      if (hipMalloc(&sets[t][i], c->L.bytes) != hipSuccess) {
        (void)hipGetLastError();
        sets[t][i] = nullptr;
        ok = false;
      }
    if (!ok) {
      for (void *q : sets[t])
        if (q) (void)hipFree(q);
      break;
    }
    got = t;
  }
  const bool timing = c->timing;
  c->timing = false;
  std::vector<float> t_ms(got + 1, 0.0f);
  int st = STH_OK;
  // Three rounds over the sets, the least time of each: clocks ramp up over
  // the first passes after an idle spell and drift under load, which a
  // single pass over the sets in order would charge to the sets.
  for (int round = 0; round < 3 && st == STH_OK; round++)
    for (uint32_t t = 0; t <= got && st == STH_OK; t++) {
      float x = 0.0f;
      st = time_fields(c, sets[t].data(), twins, &x);
      if (st == STH_OK && (round == 0 || x < t_ms[t])) t_ms[t] = x;
    }
  uint32_t best = 0;
  if (st == STH_OK)
    for (uint32_t t = 1; t <= got; t++)
      if (t_ms[t] < t_ms[best]) best = t;
  for (uint32_t t = 0; t <= got; t++)
    if (t != best)
      for (int i = 0; i < nb; i++) (void)hipFree(sets[t][i]);
  for (int i = 0; i < 3; i++) c->base[i] = sets[best][i];
  c->twin[0] = twins ? sets[best][3] : nullptr;
  c->twin[1] = twins ? sets[best][4] : nullptr;
  c->ubuf = 0;
  c->slot = 0;
  c->twins_ok = false;
  c->timing = timing;
  if (is16(c)) {
    // the kept buffers hold what the timing left: no scale state, no zero
    // halo, until the fields are initialised
    c->host_slot = Q16Slot{};
    c->host_dirty = true;
    for (int b = 0; b < 3; b++) c->halo_zero[b] = false;
  }
  if (st) return st;
  if (ms)
    for (uint32_t t = 0; t <= tries; t++) ms[t] = t <= got ? t_ms[t] : 0.0f;
  if (chosen) *chosen = best;
  return STH_OK;
}

int sth_timer_start(sth_ctx *c) {
  if (!c) return STH_ERR_ARG;
  STH_TRY(hipEventRecord(c->ev0, c->stream));
  c->timing = true;
  c->launches = 0;
  return STH_OK;
}

int sth_timer_read(sth_ctx *c, double *kernel_ms, uint64_t *launches) {
  if (!c || !c->timing) return STH_ERR_STATE;
  STH_TRY(hipEventRecord(c->ev1, c->stream));
  STH_TRY(hipEventSynchronize(c->ev1));
  float ms = 0.0f;
  STH_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->timing = false;
  if (kernel_ms) *kernel_ms = (double)ms;
  if (launches) *launches = c->launches;
  return STH_OK;
}

}  // extern "C"

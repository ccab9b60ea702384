// sth_internal.h -- the context object of libstencil_hip and the host-side
// helpers the batched driver (batch.hip) shares with the single-grid API.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fields.h"
#include "stencil_hip.h"
#include "sth_common.h"
#include "sweep_kernel.h"

// One grid on one GPU: three fields, a stream and the reduction scratch.
struct sth_ctx {
  sth_geom g;
  sth_layout_info L;
  int R = 0, eb = 0, fmt = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  void *base[3] = {nullptr, nullptr, nullptr};  // buffers a, b, s
  int ubuf = 0;                                 // buffer holding logical u
  // Two-steps-per-pass sweep: a twin of buffers a and b with the same halo
  // (allocated on first use); a pass writes the twins, then each buffer and
  // its twin swap.  twins_ok: the twins' halos equal a's and b's.
  void *twin[2] = {nullptr, nullptr};
  bool twins_ok = false;
  bool no_fuse = false;  // shots of a batch: their buffers sit in its tables
  // persistent launches: the work-queue counter and the next launch's base
  uint32_t *tickets_dev = nullptr;     // kXcds counters, 64 B apart
  uint32_t tick_next[sth::kXcds] = {};
  int cus = 0;  // compute units of the device
  // 16-bit storage scale state
  sth::Q16Slot *qs_dev = nullptr;
  sth::Q16Slot host_slot{};
  bool host_dirty = true;
  int slot = 0;
  bool halo_zero[3] = {true, true, true};
  // scratch for reductions
  uint32_t *maxbits_dev = nullptr;
  unsigned long long *errors_dev = nullptr;
  double *partial_dev = nullptr;
  double *sum_dev = nullptr;
  // pinned staging for host transfers
  void *staging = nullptr;
  size_t staging_bytes = 0;
  sth_tuning tune{};
  int weights = STH_WEIGHTS_FP32;  // FP64 storage: how its weights are formed
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timing = false;
  uint64_t launches = 0;
};

namespace sth {
namespace detail {

// Default tiles.  A field of at least one tile per CU takes the fastest
// tile measured at 1040^3 on MI355X: FP32 4 rows per wave x 8 waves x 2 k
// points per lane (32 x 128 points per workgroup), 16-bit storage 2 x 8 x 2
// (16 x 128; each lane moves 4 bytes per load as in FP32).  Smaller fields
// (and FP64) take 2 x 4 x 1 (FP32/FP64) or 2 x 8 x 2 (16-bit).  A tuning
// that names some of rows/waves/vec takes the small tile's values for the rest.
constexpr int kDefaultRows = 2;
constexpr int kDefaultWaves = 4;
constexpr int kDefaultVec = 1;
constexpr int kDefaultWaves16 = 8;
constexpr int kDefaultVec16 = 2;
inline int default_waves(int fmt) {
  return (fmt == FMT_F16 || fmt == FMT_BF16) ? kDefaultWaves16 : kDefaultWaves;
}
inline int default_vec(int fmt) {
  return (fmt == FMT_F16 || fmt == FMT_BF16) ? kDefaultVec16 : kDefaultVec;
}
struct TileShape {
  int mj, nw, vec;
};
inline TileShape default_tile(int fmt, uint32_t ys, uint32_t zs, const sth_tuning &t) {
  if (t.rows_per_wave || t.waves || t.vec)
    return {t.rows_per_wave ? (int)t.rows_per_wave : kDefaultRows,
            t.waves ? (int)t.waves : default_waves(fmt), t.vec ? (int)t.vec : default_vec(fmt)};
  if (fmt == FMT_F32 && ((ys + 31) / 32) * ((zs + 127) / 128) >= 256) return {4, 8, 2};
  return {kDefaultRows, default_waves(fmt), default_vec(fmt)};
}

// Launch plan of one sweep: tile (rows per wave x waves), i-chunks, grid.
struct Plan {
  int mj, nw, vec, chunks, chunk_len, ntk, ntj;
  int edge;  // > 0: the last k-tile runs folded, `edge` rows per wave
  bool persist;   // persistent workgroups drawing tiles from a work queue
  unsigned items;  // tiles x chunks
  int band;        // persistent: tiles per band (bands go round robin to the XCDs)
  unsigned grid;
};

// Launch plan of a pass of s (2 to 5) LEAPFROG steps.
struct Plan2 {
  int s, mj, nw, oj, ok, chunks, chunk_len, ntk, ntj;
  unsigned grid;
};

int hip_status(hipError_t e);
Plan make_plan(const sth_ctx *c, const sth_tuning &t);
bool fused_applies(const sth_ctx *c, const sth_tuning &t);
Plan2 make_plan2(const sth_ctx *c, const sth_tuning &t);
// Allocates the twins of u and v on first use and brings their halos in
// step with the fields (stream-ordered on the context's stream).
int ensure_twins(sth_ctx *c);
void base_args(const sth_ctx *c, const Plan &p, SweepArgs &a);
void fill_weights(const sth_ctx *c, bool leap, SweepArgs &a);
int push_slot(sth_ctx *c);
// `steps` LEAPFROG steps with plan p (fused passes where they apply), and
// one launch of plan p (persistent: the ticket base advances on the host).
int run_leap(sth_ctx *c, const Plan &p, uint32_t steps, float c0, float c1, const float *c2);
int launch_one(sth_ctx *c, const Plan &p, int mode, SweepArgs &a);

}  // namespace detail
}  // namespace sth

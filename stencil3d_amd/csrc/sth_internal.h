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
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timing = false;
  uint64_t launches = 0;
};

namespace sth {
namespace detail {

// Default tiles, the fastest measured at 1040^3 on MI355X: FP32/FP64 2 rows
// per wave x 4 waves x 1 k point per lane (8 x 64 points per workgroup);
// 16-bit storage 2 x 8 waves x 2 k points per lane (16 x 128), so each lane
// moves 4 bytes per load as in FP32.
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

// Launch plan of one sweep: tile (rows per wave x waves), i-chunks, grid.
struct Plan {
  int mj, nw, vec, chunks, chunk_len, ntk, ntj;
  unsigned grid;
};

int hip_status(hipError_t e);
Plan make_plan(const sth_ctx *c, const sth_tuning &t);
void base_args(const sth_ctx *c, const Plan &p, SweepArgs &a);
void fill_weights(const sth_ctx *c, bool leap, SweepArgs &a);
int push_slot(sth_ctx *c);

}  // namespace detail
}  // namespace sth

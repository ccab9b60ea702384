// sweep_dispatch.h -- selects the sweep_kernel instance for a run.
#pragma once

#include "sweep16_kernel.h"
#include "sweep_kernel.h"

namespace sth {

// (rows_per_wave, waves, k points per lane) tiles compiled into the library.
#define STH_TILE_LIST(X)                                                                    \
  X(1, 4, 1) X(2, 4, 1) X(4, 4, 1) X(8, 4, 1) X(1, 8, 1) X(2, 8, 1) X(4, 8, 1) X(8, 8, 1) \
  X(2, 2, 1) X(4, 2, 1) X(8, 2, 1) X(1, 4, 2) X(2, 4, 2) X(4, 4, 2) X(2, 8, 2) X(4, 8, 2) \
  X(2, 2, 2) X(4, 2, 2)

// Folded last k-tile (rows_per_wave, waves, k points per lane, fold) variants
// of the default tiles, FP32 only: the 16-bit tiles gain little from the fold
// and its second tile body would cost them occupancy.
#define STH_EDGE_LIST(X) X(4, 8, 2, 1) X(4, 8, 2, 2) X(4, 8, 2, 4) X(2, 8, 2, 1) X(2, 8, 2, 2)

inline bool edge_supported(int mj, int nw, int vec, int ef) {
#define STH_EDGE_OK(MJ, NW, VV, EF) if (mj == MJ && nw == NW && vec == VV && ef == EF) return true;
  STH_EDGE_LIST(STH_EDGE_OK)
#undef STH_EDGE_OK
  return false;
}

inline bool tile_supported(int mj, int nw, int vec) {
#define STH_TILE_OK(MJ, NW, VV) if (mj == MJ && nw == NW && vec == VV) return true;
  STH_TILE_LIST(STH_TILE_OK)
#undef STH_TILE_OK
  return false;
}

template <int NPTS, int MODE, int FMT>
hipError_t launch_tile(int mj, int nw, int vec, int ef, unsigned grid, hipStream_t st,
                       const SweepArgs &a) {
  if constexpr (FMT == FMT_F32) {
#define STH_EDGE_CASE(MJ, NW, VV, EF)                                                 \
  if (ef == EF && mj == MJ && nw == NW && vec == VV) {                                \
    hipLaunchKernelGGL((sweep_kernel<NPTS, MODE, FMT, MJ, NW, VV, EF>), dim3(grid),   \
                       dim3(64 * NW), 0, st, a);                                      \
    return hipGetLastError();                                                         \
  }
    STH_EDGE_LIST(STH_EDGE_CASE)
#undef STH_EDGE_CASE
  }
  if (ef != 0) return hipErrorInvalidValue;
#define STH_TILE_CASE(MJ, NW, VV)                                                     \
  if (mj == MJ && nw == NW && vec == VV) {                                            \
    hipLaunchKernelGGL((sweep_kernel<NPTS, MODE, FMT, MJ, NW, VV>), dim3(grid),       \
                       dim3(64 * NW), 0, st, a);                                      \
    return hipGetLastError();                                                         \
  }
  STH_TILE_LIST(STH_TILE_CASE)
#undef STH_TILE_CASE
  return hipErrorInvalidValue;
}

// The kernel instance a launch would use (for occupancy queries).
template <int NPTS, int MODE, int FMT>
const void *tile_fn(int mj, int nw, int vec, int ef) {
  if constexpr (FMT == FMT_F32) {
#define STH_EDGE_FN(MJ, NW, VV, EF)                        \
  if (ef == EF && mj == MJ && nw == NW && vec == VV)       \
    return (const void *)&sweep_kernel<NPTS, MODE, FMT, MJ, NW, VV, EF>;
    STH_EDGE_LIST(STH_EDGE_FN)
#undef STH_EDGE_FN
  }
  if (ef != 0) return nullptr;
#define STH_TILE_FN(MJ, NW, VV)                            \
  if (mj == MJ && nw == NW && vec == VV)                   \
    return (const void *)&sweep_kernel<NPTS, MODE, FMT, MJ, NW, VV>;
  STH_TILE_LIST(STH_TILE_FN)
#undef STH_TILE_FN
  return nullptr;
}

template <int NPTS>
const void *sweep_fn_n(int mode, int fmt, int mj, int nw, int vec, int ef) {
  if (mode == MODE_REF) {
    if (fmt == FMT_F32) return tile_fn<NPTS, MODE_REF, FMT_F32>(mj, nw, vec, ef);
    if (fmt == FMT_F64) return tile_fn<NPTS, MODE_REF, FMT_F64>(mj, nw, vec, ef);
  }
  if (mode == MODE_LEAP) {
    if (fmt == FMT_F32) return tile_fn<NPTS, MODE_LEAP, FMT_F32>(mj, nw, vec, ef);
    if (fmt == FMT_F64) return tile_fn<NPTS, MODE_LEAP, FMT_F64>(mj, nw, vec, ef);
    if (fmt == FMT_F16) return tile_fn<NPTS, MODE_LEAP, FMT_F16>(mj, nw, vec, ef);
    if (fmt == FMT_BF16) return tile_fn<NPTS, MODE_LEAP, FMT_BF16>(mj, nw, vec, ef);
  }
  return nullptr;
}

template <int NPTS>
hipError_t launch_sweep_n(int mode, int fmt, int mj, int nw, int vec, int ef, unsigned grid,
                          hipStream_t st, const SweepArgs &a) {
  if (mode == MODE_REF) {
    if (fmt == FMT_F32) return launch_tile<NPTS, MODE_REF, FMT_F32>(mj, nw, vec, ef, grid, st, a);
    if (fmt == FMT_F64) return launch_tile<NPTS, MODE_REF, FMT_F64>(mj, nw, vec, ef, grid, st, a);
  }
  if (mode == MODE_LEAP) {
    if (fmt == FMT_F32) return launch_tile<NPTS, MODE_LEAP, FMT_F32>(mj, nw, vec, ef, grid, st, a);
    if (fmt == FMT_F64) return launch_tile<NPTS, MODE_LEAP, FMT_F64>(mj, nw, vec, ef, grid, st, a);
    if (fmt == FMT_F16) return launch_tile<NPTS, MODE_LEAP, FMT_F16>(mj, nw, vec, ef, grid, st, a);
    if (fmt == FMT_BF16) return launch_tile<NPTS, MODE_LEAP, FMT_BF16>(mj, nw, vec, ef, grid, st, a);
  }
  return hipErrorInvalidValue;
}

// Defined in sweep_7.hip, sweep_13.hip, sweep_25.hip (one compile unit each).
hipError_t launch_sweep_7(int mode, int fmt, int mj, int nw, int vec, int ef, unsigned grid,
                          hipStream_t st, const SweepArgs &a);
hipError_t launch_sweep_13(int mode, int fmt, int mj, int nw, int vec, int ef, unsigned grid,
                           hipStream_t st, const SweepArgs &a);
hipError_t launch_sweep_25(int mode, int fmt, int mj, int nw, int vec, int ef, unsigned grid,
                           hipStream_t st, const SweepArgs &a);

const void *sweep_fn_7(int mode, int fmt, int mj, int nw, int vec, int ef);
const void *sweep_fn_13(int mode, int fmt, int mj, int nw, int vec, int ef);
const void *sweep_fn_25(int mode, int fmt, int mj, int nw, int vec, int ef);

// The packed FP16 LEAPFROG sweep (sweep16_kernel.h, sweep16.hip): the tiles
// of two k points per lane listed above.
// ef > 0: the kernel with a folded last k-tile (edge16_supported).
hipError_t launch_sweep16(int npts, int mj, int nw, int ef, unsigned grid, hipStream_t st,
                          const SweepArgs &a);
const void *sweep16_fn(int npts, int mj, int nw, int ef);
bool tile16_supported(int mj, int nw);
bool edge16_supported(int npts, int mj, int nw, int ef);
// Steps per loop turn of the packed sweep: a tile's i-chunk of a multiple of
// it runs no idle step.
inline int sweep16_turn(int R) {
  const int d = Tile16<7, 1, 1>::D;
  return (2 * R + d + 2) / 2 * 2;
}
// An FP16 LEAPFROG sweep of two k points per lane runs the packed kernel.
inline bool packed16(int fmt, int mode, int vec) {
  return fmt == FMT_F16 && mode == MODE_LEAP && vec == 2;
}

inline const void *sweep_fn(int npts, int mode, int fmt, int mj, int nw, int vec, int ef) {
  if (npts == 7) return sweep_fn_7(mode, fmt, mj, nw, vec, ef);
  if (npts == 13) return sweep_fn_13(mode, fmt, mj, nw, vec, ef);
  if (npts == 25) return sweep_fn_25(mode, fmt, mj, nw, vec, ef);
  return nullptr;
}

inline hipError_t launch_sweep(int npts, int mode, int fmt, int mj, int nw, int vec, int ef,
                               unsigned grid, hipStream_t st, const SweepArgs &a) {
  if (npts == 7) return launch_sweep_7(mode, fmt, mj, nw, vec, ef, grid, st, a);
  if (npts == 13) return launch_sweep_13(mode, fmt, mj, nw, vec, ef, grid, st, a);
  if (npts == 25) return launch_sweep_25(mode, fmt, mj, nw, vec, ef, grid, st, a);
  return hipErrorInvalidValue;
}

}  // namespace sth

// sweep16.hip -- instances of the packed FP16 LEAPFROG sweep (sweep16_kernel.h).
#include "sweep16_kernel.h"
#include "sweep_dispatch.h"

namespace sth {

// (rows per wave, waves) of the packed tiles: the two-points-per-lane tiles
// of STH_TILE_LIST and taller ones.
#define STH_TILE16_LIST(X) \
  X(1, 4) X(2, 4) X(4, 4) X(2, 8) X(4, 8) X(2, 2) X(4, 2) X(1, 8) X(3, 8) X(1, 16) X(2, 16) X(3, 16)

// (rows per wave, waves, fold) of the tiles compiled with a folded last
// k-tile: the default tile (measured at 1040^3, 25 points, one box: 2x8
// 1.89 ms per step folded, 1.93 unfolded; 2x16 1.96 folded, 1.91 unfolded,
// so 2x16 stays unfolded).
#define STH_EDGE16_LIST(X) X(2, 8, 2) X(2, 8, 4) X(2, 8, 8)

bool tile16_supported(int mj, int nw) {
#define STH_T16_OK(MJ, NW) \
  if (mj == MJ && nw == NW) return true;
  STH_TILE16_LIST(STH_T16_OK)
#undef STH_T16_OK
  return false;
}

bool edge16_supported(int npts, int mj, int nw, int ef) {
  const int R = npts == 7 ? 1 : npts == 13 ? 2 : 4;
  if (!fold16_fits(R, mj, ef)) return false;
#define STH_E16_OK(MJ, NW, EF) \
  if (mj == MJ && nw == NW && ef == EF) return true;
  STH_EDGE16_LIST(STH_E16_OK)
#undef STH_E16_OK
  return false;
}

template <int NPTS>
static const void *fn16(int mj, int nw, int ef) {
#define STH_E16_FN(MJ, NW, EF)                                               \
  if constexpr (fold16_fits(Terms<NPTS>::R, MJ, EF))                         \
    if (mj == MJ && nw == NW && ef == EF) return (const void *)&sweep16_kernel<NPTS, MJ, NW, EF>;
  STH_EDGE16_LIST(STH_E16_FN)
#undef STH_E16_FN
  if (ef != 0) return nullptr;
#define STH_T16_FN(MJ, NW) \
  if (mj == MJ && nw == NW) return (const void *)&sweep16_kernel<NPTS, MJ, NW>;
  STH_TILE16_LIST(STH_T16_FN)
#undef STH_T16_FN
  return nullptr;
}

template <int NPTS>
static hipError_t launch16(int mj, int nw, int ef, unsigned grid, hipStream_t st,
                           const SweepArgs &a) {
#define STH_E16_CASE(MJ, NW, EF)                                                       \
  if constexpr (fold16_fits(Terms<NPTS>::R, MJ, EF))                                   \
    if (mj == MJ && nw == NW && ef == EF) {                                            \
      hipLaunchKernelGGL((sweep16_kernel<NPTS, MJ, NW, EF>), dim3(grid), dim3(64 * NW), \
                         0, st, a);                                                    \
      return hipGetLastError();                                                        \
    }
  STH_EDGE16_LIST(STH_E16_CASE)
#undef STH_E16_CASE
  if (ef != 0) return hipErrorInvalidValue;
#define STH_T16_CASE(MJ, NW)                                                                \
  if (mj == MJ && nw == NW) {                                                               \
    hipLaunchKernelGGL((sweep16_kernel<NPTS, MJ, NW>), dim3(grid), dim3(64 * NW), 0, st, a); \
    return hipGetLastError();                                                               \
  }
  STH_TILE16_LIST(STH_T16_CASE)
#undef STH_T16_CASE
  return hipErrorInvalidValue;
}

const void *sweep16_fn(int npts, int mj, int nw, int ef) {
  if (npts == 7) return fn16<7>(mj, nw, ef);
  if (npts == 13) return fn16<13>(mj, nw, ef);
  if (npts == 25) return fn16<25>(mj, nw, ef);
  return nullptr;
}

hipError_t launch_sweep16(int npts, int mj, int nw, int ef, unsigned grid, hipStream_t st,
                          const SweepArgs &a) {
  if (npts == 7) return launch16<7>(mj, nw, ef, grid, st, a);
  if (npts == 13) return launch16<13>(mj, nw, ef, grid, st, a);
  if (npts == 25) return launch16<25>(mj, nw, ef, grid, st, a);
  return hipErrorInvalidValue;
}

}  // namespace sth

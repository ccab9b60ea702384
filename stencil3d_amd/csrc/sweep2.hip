// sweep2.hip -- launches of the fused LEAPFROG sweep (FP32, two to five steps per pass).
#include "sweep2_dispatch.h"
#include "sweep2_kernel.h"
#include "sweep2h_kernel.h"

namespace sth {

// (stencil, rows per wave, waves, steps per pass)
#define STH_FUSED2_LIST(X)                                                                \
  X(7, 4, 16, 2) X(7, 4, 8, 2) X(13, 4, 16, 2) X(13, 3, 16, 2) X(13, 4, 8, 2) X(25, 3, 16, 2) \
  X(25, 4, 12, 2) X(25, 3, 8, 2) X(25, 4, 8, 2) X(7, 3, 16, 3) X(7, 4, 12, 3) \
  X(13, 3, 16, 3) X(13, 4, 12, 3) X(7, 4, 10, 4) X(7, 5, 8, 4) X(7, 2, 16, 4) \
  X(7, 2, 16, 5)

bool fused2_supported(int npts, int mj, int nw, int steps) {
#define STH_F2_SUP(N, MJ, NW, S) \
  if (npts == N && mj == MJ && nw == NW && steps == S) return true;
  STH_FUSED2_LIST(STH_F2_SUP)
#undef STH_F2_SUP
  return false;
}

Fused2Geom fused2_geom(int npts, int mj, int nw, int steps) {
#define STH_F2_GEOM(N, MJ, NW, S)                                     \
  if (npts == N && mj == MJ && nw == NW && steps == S)                \
    return {Fused2Shape<N, MJ, NW, S>::OJ, Fused2Shape<N, MJ, NW, S>::OK, 64 * NW};
  STH_FUSED2_LIST(STH_F2_GEOM)
#undef STH_F2_GEOM
  return {0, 0, 0};
}

hipError_t launch_sweep2(int npts, int mj, int nw, int steps, unsigned grid, hipStream_t st,
                         const Sweep2Args &a) {
#define STH_F2_LAUNCH(N, MJ, NW, S)                                                   \
  if (npts == N && mj == MJ && nw == NW && steps == S) {                              \
    hipLaunchKernelGGL((sweep2_kernel<N, MJ, NW, S>), dim3(grid), dim3(64 * NW), 0, st, a); \
    return hipGetLastError();                                                         \
  }
  STH_FUSED2_LIST(STH_F2_LAUNCH)
#undef STH_F2_LAUNCH
  return hipErrorInvalidValue;
}

// FP16 storage, two steps per pass: (stencil, rows per wave, waves)
// (storage format, stencil, rows per wave, waves, steps per pass) of the
// 16-bit passes.  BF16 has no mixed-precision FMA on gfx950: its halves are
// widened by a shift or a mask, which keeps more values live, and only the
// shapes listed fit 128 VGPRs without spilling (25 points, 13 points' three
// steps and 7 points' four and five spill at every shape with 16 waves).
#define STH_FUSED2H_LIST(X)                                                                 \
  X(FMT_F16, 7, 4, 16, 2) X(FMT_F16, 13, 4, 16, 2) X(FMT_F16, 13, 3, 16, 2)                \
  X(FMT_F16, 25, 3, 16, 2) X(FMT_F16, 7, 3, 16, 3) X(FMT_F16, 13, 3, 16, 3)                 \
  X(FMT_F16, 7, 2, 16, 4) X(FMT_F16, 7, 2, 16, 5) X(FMT_BF16, 7, 4, 16, 2)                  \
  X(FMT_BF16, 7, 2, 16, 3) X(FMT_BF16, 13, 2, 16, 2)

bool fused2h_supported(int fmt, int npts, int mj, int nw, int steps) {
#define STH_F2H_SUP(F, N, MJ, NW, S) \
  if (fmt == F && npts == N && mj == MJ && nw == NW && steps == S) return true;
  STH_FUSED2H_LIST(STH_F2H_SUP)
#undef STH_F2H_SUP
  return false;
}

Fused2Geom fused2h_geom(int fmt, int npts, int mj, int nw, int steps) {
#define STH_F2H_GEOM(F, N, MJ, NW, S)                                \
  if (fmt == F && npts == N && mj == MJ && nw == NW && steps == S)   \
    return {Fused2hShape<N, MJ, NW, S>::OJ, Fused2hShape<N, MJ, NW, S>::OK, 64 * NW};
  STH_FUSED2H_LIST(STH_F2H_GEOM)
#undef STH_F2H_GEOM
  return {0, 0, 0};
}

hipError_t launch_sweep2h(int fmt, int npts, int mj, int nw, int steps, unsigned grid,
                          hipStream_t st, const Sweep2Args &a) {
#define STH_F2H_LAUNCH(F, N, MJ, NW, S)                                                       \
  if (fmt == F && npts == N && mj == MJ && nw == NW && steps == S) {                         \
    hipLaunchKernelGGL((sweep2h_kernel<N, MJ, NW, S, F>), dim3(grid), dim3(64 * NW), 0, st, a); \
    return hipGetLastError();                                                                \
  }
  STH_FUSED2H_LIST(STH_F2H_LAUNCH)
#undef STH_F2H_LAUNCH
  return hipErrorInvalidValue;
}

// (stencil, rows per wave, waves, steps per pass)
#define STH_BATCH2_LIST(X) \
  X(7, 4, 16, 2) X(13, 4, 16, 2) X(25, 3, 16, 2) X(7, 3, 16, 3) X(13, 3, 16, 3) X(7, 2, 16, 5)

bool batch2_supported(int npts, int mj, int nw, int steps) {
#define STH_B2_SUP(N, MJ, NW, S) \
  if (npts == N && mj == MJ && nw == NW && steps == S) return true;
  STH_BATCH2_LIST(STH_B2_SUP)
#undef STH_B2_SUP
  return false;
}

const void *batch2_fn(int npts, int mj, int nw, int steps) {
#define STH_B2_FN(N, MJ, NW, S)                                  \
  if (npts == N && mj == MJ && nw == NW && steps == S)           \
    return (const void *)&batch2_kernel<N, MJ, NW, S>;
  STH_BATCH2_LIST(STH_B2_FN)
#undef STH_B2_FN
  return nullptr;
}

hipError_t launch_batch2(int npts, int mj, int nw, int steps, unsigned grid, hipStream_t st,
                         const Batch2Args &b) {
#define STH_B2_LAUNCH(N, MJ, NW, S)                                                   \
  if (npts == N && mj == MJ && nw == NW && steps == S) {                              \
    hipLaunchKernelGGL((batch2_kernel<N, MJ, NW, S>), dim3(grid), dim3(64 * NW), 0, st, b); \
    return hipGetLastError();                                                         \
  }
  STH_BATCH2_LIST(STH_B2_LAUNCH)
#undef STH_B2_LAUNCH
  return hipErrorInvalidValue;
}

}  // namespace sth

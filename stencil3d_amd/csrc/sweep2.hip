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
// (stencil, rows per wave, waves, steps per pass) of the FP16 passes
#define STH_FUSED2H_LIST(X)                                                           \
  X(7, 4, 16, 2) X(13, 4, 16, 2) X(13, 3, 16, 2) X(25, 3, 16, 2) X(7, 3, 16, 3)       \
  X(13, 3, 16, 3) X(7, 2, 16, 4) X(7, 2, 16, 5)

bool fused2h_supported(int npts, int mj, int nw, int steps) {
#define STH_F2H_SUP(N, MJ, NW, S) \
  if (npts == N && mj == MJ && nw == NW && steps == S) return true;
  STH_FUSED2H_LIST(STH_F2H_SUP)
#undef STH_F2H_SUP
  return false;
}

Fused2Geom fused2h_geom(int npts, int mj, int nw, int steps) {
#define STH_F2H_GEOM(N, MJ, NW, S)                        \
  if (npts == N && mj == MJ && nw == NW && steps == S)    \
    return {Fused2hShape<N, MJ, NW, S>::OJ, Fused2hShape<N, MJ, NW, S>::OK, 64 * NW};
  STH_FUSED2H_LIST(STH_F2H_GEOM)
#undef STH_F2H_GEOM
  return {0, 0, 0};
}

hipError_t launch_sweep2h(int npts, int mj, int nw, int steps, unsigned grid, hipStream_t st,
                          const Sweep2Args &a) {
#define STH_F2H_LAUNCH(N, MJ, NW, S)                                                     \
  if (npts == N && mj == MJ && nw == NW && steps == S) {                                 \
    hipLaunchKernelGGL((sweep2h_kernel<N, MJ, NW, S>), dim3(grid), dim3(64 * NW), 0, st, a); \
    return hipGetLastError();                                                            \
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

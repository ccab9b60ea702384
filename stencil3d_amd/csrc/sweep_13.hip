// sweep_13.hip -- the 13-point sweep kernels (one compile unit per stencil size).
#include "sweep_dispatch.h"

namespace sth {

hipError_t launch_sweep_13(int mode, int fmt, int mj, int nw, int vec, int ef, unsigned grid,
                          hipStream_t st, const SweepArgs &a) {
  return launch_sweep_n<13>(mode, fmt, mj, nw, vec, ef, grid, st, a);
}

const void *sweep_fn_13(int mode, int fmt, int mj, int nw, int vec, int ef) {
  return sweep_fn_n<13>(mode, fmt, mj, nw, vec, ef);
}

}  // namespace sth

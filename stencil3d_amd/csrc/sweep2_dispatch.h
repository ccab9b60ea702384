// sweep2_dispatch.h -- host-side entry to the fused sweep (two to five steps per pass).
#pragma once

#include <hip/hip_runtime.h>

namespace sth {

struct Sweep2Args;
struct Batch2Args;

// Output tile of a compiled (stencil, rows per wave, waves, steps per pass)
// shape.
struct Fused2Geom {
  int oj, ok, threads;
};

bool fused2_supported(int npts, int mj, int nw, int steps);
Fused2Geom fused2_geom(int npts, int mj, int nw, int steps);
hipError_t launch_sweep2(int npts, int mj, int nw, int steps, unsigned grid, hipStream_t st,
                         const Sweep2Args &a);
// The 16-bit (FP16, BF16) passes of two to five steps (sweep2h_kernel.h):
// their compiled shapes, their output tiles, and the launch (one workgroup
// per output tile and i-chunk).
bool fused2h_supported(int fmt, int npts, int mj, int nw, int steps);
Fused2Geom fused2h_geom(int fmt, int npts, int mj, int nw, int steps);
hipError_t launch_sweep2h(int fmt, int npts, int mj, int nw, int steps, unsigned grid,
                          hipStream_t st, const Sweep2Args &a);
// The batch form (persistent, items of many shots from a queue) exists for
// each stencil's default shape of each pass length.
bool batch2_supported(int npts, int mj, int nw, int steps);
const void *batch2_fn(int npts, int mj, int nw, int steps);
hipError_t launch_batch2(int npts, int mj, int nw, int steps, unsigned grid, hipStream_t st,
                         const Batch2Args &b);

}  // namespace sth

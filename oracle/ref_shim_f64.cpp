// ref_shim_f64.cpp -- C entry point into the reference's -DUSE_DOUBLE build.
//
// TEST INFRASTRUCTURE ONLY.  Same as ref_shim.cpp with real = double
// (config.h:96-100); used to pin the FP64 REF restatement.
#include "iterators.h"

extern "C" {

int ref_shim_iterate_f64(double *v, const double *s, const double *u, unsigned xs, unsigned ys,
                         unsigned zs, unsigned xh, unsigned yh, unsigned zh, unsigned zb,
                         unsigned iterations) {
  if (iterations == 0) return 0;
  s_grid grid(xs, ys, zs, xh, yh, zh, 1, 1, zb, 12.0, 13, 3);
  ITERATE(v, s, u, grid, 64, 1, iterations, 1);
  return 0;
}

}  // extern "C"

// ref_shim.cpp -- a C entry point into the reference's own sweep, for tests.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's iterators.cpp, grid.cpp, gilbert2d.cpp and aux.cpp, where
// they lie under the reference checkout, into oracle/_ref/libref<N>_<it>.so.
// The tests and tests/golden/gen_golden.py call it to run the reference's
// ITERATE (iterators.cpp:130-157 for iterate_3d, :197-234 for
// iterate_2d_jk_blocked) on seeded fields and to compare the CPU restatement
// (oracle/stencil_oracle.c) against it bit for bit.
#include "iterators.h"

extern "C" {

// Runs ITERATE(v, s, u, grid, ...) as main.cpp:297/307 do, on caller-owned
// arrays in the reference's dense layout (grid.cpp:20-25).  zb is the k-block
// of the blocked iterators (ignored by iterate_3d).
int ref_shim_iterate(float *v, const float *s, const float *u, unsigned xs, unsigned ys,
                     unsigned zs, unsigned xh, unsigned yh, unsigned zh, unsigned zb,
                     unsigned iterations) {
  if (iterations == 0) return 0;
  s_grid grid(xs, ys, zs, xh, yh, zh, 1, 1, zb, 12.0f, 13, 3);
  ITERATE(v, s, u, grid, 64, 1, iterations, 1);
  return 0;
}

int ref_shim_stencil_size(void) { return STENCIL_SIZE; }

}  // extern "C"

// iterators_hip.cpp -- the reference's iterator interface (iterators.h:6-35)
// implemented on libstencil_hip.
//
// A build of the reference that compiles this file in place of iterators.cpp
// (and gilbert2d.cpp stays for grid.cpp) keeps main.cpp -- its CLI, phases,
// checks and report -- unchanged and runs every sweep on the GPU.  This file
// is the reference-side binding: it is compiled against the reference's own
// headers (iterators.h for the signatures, grid.h for s_grid, config.h for
// `real` and STENCIL_SIZE), from the reference tree, by the recipe beside it.
//
// Contract kept from the reference:
//  * ITERATE(v, s, u, grid, prefetch, dups, iterations, cacheiter) takes host
//    arrays in the s_grid layout (grid.cpp:20-25), reads u and s, updates v
//    in place `iterations` times (at least once: the reference's loop is a
//    do-while, iterators.cpp:137-156) and returns with v written.  prefetch,
//    dups and cacheiter have no meaning on the GPU (iterate_3d ignores
//    cacheiter too).  Every iterator variant of config.h:55-65 is the same
//    sweep here: the blocked and Hilbert orders are CPU cache schedules whose
//    results equal iterate_3d's.
//  * init_arrays (iterators.cpp:292-301) and check (iterators.cpp:304-325)
//    act on the host arrays as the reference's do, with the same message.
//
// Each ITERATE call uploads u, s and v, runs the sweeps and downloads v (its
// interior and the radius-wide halo; the rest of v's halo is the caller's and
// is not touched), so its time includes the PCIe transfers of the three
// fields: the caller owns the arrays between calls, as in the reference.
// One device context is kept per grid shape and reused.  `real` = double (the
// reference's -DUSE_DOUBLE build) runs FP64 storage with the weights that
// build forms.  Errors of the library end the process with its message and
// exit code 3 (the reference has no error path here).
//
// Not reproduced: iterate_3d's j loop bound (iterators.cpp:146 runs j to the
// x size); this binding sweeps every row of the grid.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "iterators.h"
#include "stencil_hip.h"

static_assert(sizeof(real) == sizeof(float) || sizeof(real) == sizeof(double),
              "real is float or double");

namespace {

constexpr bool kDouble = sizeof(real) == sizeof(double);

sth_ctx *g_ctx = nullptr;
unsigned g_xs = 0, g_ys = 0, g_zs = 0;

[[noreturn]] void fail(const char *what, int st) {
  fprintf(stderr, "libstencil_hip: %s: %s\n", what, sth_strerror(st));
  exit(3);
}

void release() {
  if (g_ctx) sth_destroy(g_ctx);
  g_ctx = nullptr;
}

sth_ctx *context(const s_grid &grid) {
  if (g_ctx && g_xs == grid.xsize && g_ys == grid.ysize && g_zs == grid.zsize) return g_ctx;
  if (!g_ctx) atexit(release);
  release();
  sth_geom g = {grid.xsize, grid.ysize, grid.zsize, (uint32_t)STENCIL_SIZE,
                kDouble ? (uint32_t)STH_FP64 : (uint32_t)STH_FP32};
  int st = sth_create(&g, &g_ctx);
  if (st) fail("sth_create", st);
  if (kDouble && (st = sth_set_weights(g_ctx, STH_WEIGHTS_REFERENCE))) fail("sth_set_weights", st);
  g_xs = grid.xsize;
  g_ys = grid.ysize;
  g_zs = grid.zsize;
  return g_ctx;
}

int upload(sth_ctx *c, sth_field f, const real *x, const sth_host_layout &hl) {
  if constexpr (kDouble)
    return sth_upload_f64(c, f, (const double *)x, &hl);
  else
    return sth_upload(c, f, (const float *)x, &hl);
}

int download(sth_ctx *c, sth_field f, real *x, const sth_host_layout &hl) {
  if constexpr (kDouble)
    return sth_download_f64(c, f, (double *)x, &hl);
  else
    return sth_download(c, f, (float *)x, &hl);
}

void sweep(real *v, const real *s, const real *u, const s_grid &grid, unsigned iterations) {
  sth_ctx *c = context(grid);
  const sth_host_layout hl = {grid.xhalo, grid.yhalo, grid.zhalo};
  int st;
  if ((st = upload(c, STH_FIELD_U, u, hl))) fail("upload u", st);
  if ((st = upload(c, STH_FIELD_S, s, hl))) fail("upload s", st);
  if ((st = upload(c, STH_FIELD_V, v, hl))) fail("upload v", st);
  if ((st = sth_iterate(c, iterations > 0 ? iterations : 1))) fail("sth_iterate", st);
  if ((st = download(c, STH_FIELD_V, v, hl))) fail("download v", st);
}

}  // namespace

void iterate_3d(real *restrict v, const real *restrict s, const real *restrict u,
                const s_grid &grid, const unsigned int, const unsigned int,
                const unsigned int iterations, const unsigned int) {
  sweep(v, s, u, grid, iterations);
}

void iterate_3d_jk_blocked(real *restrict v, const real *restrict s, const real *restrict u,
                           const s_grid &grid, const unsigned int, const unsigned int,
                           const unsigned int iterations, const unsigned int) {
  sweep(v, s, u, grid, iterations);
}

void iterate_2d_jk_blocked(real *restrict v, const real *restrict s, const real *restrict u,
                           const s_grid &grid, const unsigned int, const unsigned int,
                           const unsigned int iterations, const unsigned int) {
  sweep(v, s, u, grid, iterations);
}

void iterate_2d_hilbert(real *restrict v, const real *restrict s, const real *restrict u,
                        const s_grid &grid, const unsigned int, const unsigned int,
                        const unsigned int iterations, const unsigned int) {
  sweep(v, s, u, grid, iterations);
}

void iterate_2d_checkerboard(real *restrict v, const real *restrict s, const real *restrict u,
                             const s_grid &grid, const unsigned int, const unsigned int,
                             const unsigned int iterations, const unsigned int) {
  sweep(v, s, u, grid, iterations);
}

// v = 0, u = s = 1 over the whole allocation, halos included.
void init_arrays(real *restrict v, real *restrict s, real *restrict u, unsigned int gsize) {
#pragma omp parallel for schedule(static)
  for (unsigned int n = 0; n < gsize; n++) {
    v[n] = 0;
    u[n] = 1;
    s[n] = 1;
  }
}

// Interior points farther than 0.1 from maxiter.
void check(const real *restrict v, const s_grid &grid, const unsigned int maxiter) {
  unsigned int errors = 0;
  const unsigned int xs = grid.xsize, ys = grid.ysize, zs = grid.zsize;
#pragma omp parallel for collapse(2) reduction(+ : errors) schedule(static)
  for (unsigned int i = 0; i < xs; i++)
    for (unsigned int j = 0; j < ys; j++) {
      const real *row = v + grid.ijk_ofs + (size_t)i * grid.imult + (size_t)j * grid.jmult;
      for (unsigned int k = 0; k < zs; k++)
        if (fabs(row[k] - (real)maxiter) > 0.1f) errors++;
    }
  printf("check: %d errors found\n", errors);
}

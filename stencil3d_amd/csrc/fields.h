// fields.h -- descriptors and launchers of the field kernels (fields.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_kernel.h"

namespace sth {

constexpr int kReduceBlocks = 2048;

struct FieldDesc {
  int64_t pitch, plane;  // elements
  uint64_t align;        // elements per 256 B (the front pad of every row)
  uint64_t elems;        // elements allocated
  int xs, ys, zs, R;
};

enum { FILL_CONST = 0, FILL_UNIFORM = 1, FILL_GAUSS = 2, FILL_POINT = 3 };

struct FillSpec {
  int kind;
  float a;       // CONST value / POINT value
  float halo;    // UNIFORM halo value
  double lo, hi; // UNIFORM range; GAUSS: lo = amplitude, hi = sigma
  double noise;  // GAUSS noise half-width
  uint64_t key;  // SplitMix64 stream key of (seed, field_id)
  int pi, pj, pk;
};

// Terms of the LEAPFROG operator M = c0 + c1 L for energy_kernel: element
// offsets of the neighbour pairs and their weights in FP64.
struct EnergySpec {
  int T;
  int64_t off[12];
  double w[12];
  double c0, c1;
};

hipError_t launch_fill(const FieldDesc &d, const FillSpec &f, void *dst, int fmt, int e, int pass,
                       uint32_t *maxbits, hipStream_t st);
hipError_t launch_check(const FieldDesc &d, const void *v, int fmt, int e, float expected,
                        unsigned long long *errors, hipStream_t st);
hipError_t launch_norms(const FieldDesc &d, const void *x, int fmt, int e, double *partial,
                        uint32_t *maxbits, double *out, hipStream_t st);
hipError_t launch_energy(const FieldDesc &d, const void *u, const void *v, int fmt, int eu,
                         int ev, const EnergySpec &sp, double *partial, double *out,
                         hipStream_t st);
hipError_t launch_compare(const FieldDesc &dx, const void *x, int fx, int ex,
                          const FieldDesc &dr, const void *r, int fr, int er, double *pdiff,
                          double *pref, uint32_t *maxdiff, uint32_t *maxref, double *sums,
                          hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t st);
hipError_t launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st);
hipError_t launch_bw_probe(int nr, int nw, int lane_bytes, const void *const *r, void *const *w,
                           uint64_t bytes, float *sink, unsigned grid, hipStream_t st);

}  // namespace sth

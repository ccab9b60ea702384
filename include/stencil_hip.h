/*
 * stencil_hip.h -- C ABI of libstencil_hip, the MI355X (gfx950) sweep library
 * that replaces the reference's time-stepping hot path.
 *
 * Reference interface replaced (files of mathias-puetz/stencil3d):
 *   ITERATE(v,s,u,grid,prefetch,dups,iterations,cacheiter)
 *       iterators.h:6-29, iterate_3d iterators.cpp:130-157,
 *       selected by config.h:55-65                      -> sth_iterate
 *   per-point kernels stencil_{7,13,25}_ref_1
 *       stencil_7.h:31-39, stencil_13.h:50-61, stencil_25.h:75-83
 *                                                       -> device code behind sth_iterate
 *   s_grid geometry (grid.h:7-71, grid.cpp:9-48)        -> sth_geom / sth_layout
 *   array allocation (main.cpp:244-275)                 -> sth_create (hipMalloc)
 *   init_arrays (iterators.h:32, iterators.cpp:292-301) -> sth_init_const
 *   check (iterators.h:35, iterators.cpp:304-325)       -> sth_check
 *   the leapfrog update of README.md:10-24 (no reference code) -> sth_leapfrog
 *
 * Conventions: every call returns an sth_status (0 = OK) and never exits the
 * process.  A context owns its device memory and one HIP stream; calls on a
 * context are stream-ordered and asynchronous until sth_sync / a download /
 * a check.  One host thread per context; contexts are independent.  Axis
 * naming follows the reference: i (x, slowest), j (y), k (z, unit stride).
 */
#ifndef STENCIL_HIP_H
#define STENCIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STH_VERSION 1

typedef enum {
  STH_OK = 0,
  STH_ERR_ARG = 1,    /* invalid argument (the reference exits with 2, main.cpp:149-150) */
  STH_ERR_HIP = 2,    /* HIP runtime error */
  STH_ERR_NOMEM = 3,  /* device allocation failed */
  STH_ERR_STATE = 4,  /* call not valid in the context's current state */
  STH_ERR_NODEV = 5,  /* no usable GPU */
  STH_ERR_UNSUPPORTED = 6
} sth_status;

/* FP64 storage computes in FP64.  Its stencil weights are set per context
 * (sth_set_weights): by default the FP32 constants promoted (the same
 * discrete operator as FP32, the on-GPU error reference of the FP32 and
 * 16-bit forms), or as the reference's -DUSE_DOUBLE build forms them
 * (config.h:96-100). */
typedef enum { STH_FP32 = 0, STH_FP16 = 1, STH_BF16 = 2, STH_FP64 = 3 } sth_dtype;

/* Weights of FP64 storage.  STH_WEIGHTS_FP32: each weight is the FP32
 * constant of stencil_N.h promoted to double.  STH_WEIGHTS_REFERENCE: the
 * reference's -DUSE_DOUBLE build (config.h:96-100): the float literals
 * promoted, cy = cx*1.1f and cz = cx*0.9f evaluated in double
 * (stencil_13.h:13-22, stencil_25.h:19-34); bit-identical REF sweeps to that
 * build. */
typedef enum { STH_WEIGHTS_FP32 = 0, STH_WEIGHTS_REFERENCE = 1 } sth_weights;

/* Logical fields.  In REF mode v is updated in place and u, s are read-only.
 * In LEAPFROG mode u is the newest field and v the previous one; each step
 * writes the new field over v and swaps the two names. */
typedef enum { STH_FIELD_V = 0, STH_FIELD_S = 1, STH_FIELD_U = 2 } sth_field;

/* Geometry of a run (s_grid, grid.h:7-71).  The device keeps a halo of
 * exactly the stencil radius; the caller's halo widths describe its own host
 * buffers (sth_host_layout).  zs, ys, xs: 1..65535, and one device plane
 * ((ys+2R) rows of the padded pitch) below 2 GiB. */
typedef struct {
  uint32_t xs, ys, zs; /* interior extents (--xs --ys --zs) */
  uint32_t npts;       /* 7, 13 or 25 (STENCIL_SIZE, config.h:37-39) */
  uint32_t dtype;      /* sth_dtype: storage type of u, v, s */
} sth_geom;

/* Device layout (host-side computation, no GPU needed).  Element (i,j,k),
 * -R <= i < xs+R etc., lives at element offset
 *   origin + k + j*pitch + i*plane
 * from the start of the field; pitch is a multiple of 256 bytes and
 * origin*elem_bytes is a multiple of 256, so every interior row starts on a
 * 256-byte boundary. */
typedef struct {
  uint32_t radius;     /* R = 1, 2, 4 */
  uint32_t elem_bytes; /* 8, 4 or 2 */
  uint64_t pitch;      /* elements between consecutive j rows */
  uint64_t plane;      /* elements between consecutive i planes */
  uint64_t origin;     /* element offset of interior point (0,0,0) */
  uint64_t elems;      /* elements allocated per field */
  uint64_t bytes;      /* bytes allocated per field */
} sth_layout_info;

/* A caller's dense host array in the reference layout (grid.cpp:20-25):
 * jmult = zs+2zh, imult = jmult*(ys+2yh), element (i,j,k) at
 * zh + yh*jmult + xh*imult + k + j*jmult + i*imult.  Halo widths must be at
 * least the stencil radius. */
typedef struct {
  uint32_t xh, yh, zh;
} sth_host_layout;

/* Tile schedule knobs of the 2.5D sweep (the GPU counterpart of --yb/--zb,
 * iterators.cpp:42-89).  Zero fields select the built-in defaults. */
typedef struct {
  uint32_t rows_per_wave; /* j rows held in registers by one wave: 1, 2, 4, 8
                             (two- or three-step pass: 3 or 4; four-step: 2, 4, 5;
                             five-step: 2) */
  uint32_t waves;         /* waves per workgroup: 2, 4, 8 (two-step pass: 8, 12, 16;
                             three-step pass: 12, 16; four-step: 8, 10, 16;
                             five-step: 16) */
  uint32_t chunks;        /* i-range split: 1..64 chunks per (j,k) tile (0: chosen from
                             the grid and the device's CU count) */
  uint32_t xcd_remap;     /* 1: keep neighbouring tiles on one XCD (default), 0: off */
  uint32_t vec;           /* consecutive k points per lane: 1 or 2 (0: default) */
  uint32_t steps_per_pass; /* LEAPFROG, FP32 and FP16: 2 = two steps per pass over
                              memory (temporal blocking), 3 = three (7 and 13
                              points), 4 or 5 = four or five (7 points); BF16: 2 (7
                              and 13 points) or 3 (7 points).  16-bit: each step's
                              scale from the bound of the previous steps' outputs
                              (rows_per_wave, 16 waves: FP16 4 or 3 for two steps,
                              3 for three, 2 for four and five; BF16 4 for 7
                              points' two steps, else 2); steps left over run as
                              the longest shorter pass that fits, then one-step;
                              1 = one; 0: default */
  uint32_t launch;        /* one-step sweep: 1 = one workgroup per tile, 2 = persistent
                             workgroups drawing tiles from a work queue; 0: default (2) */
  uint32_t ref_sweeps;    /* REF (FP32, FP64): sweeps per launch, 1..4096 (0: one).  u and s
                             are read-only in iterate_3d, so one launch computes each
                             point's stencil once and applies k updates to v in
                             registers: bit-identical to k sweeps, 1/k of the traffic */
} sth_tuning;

typedef struct sth_ctx sth_ctx;

const char *sth_strerror(int status);
int sth_version(void);
/* Host-only: device layout for a geometry (no GPU access). */
int sth_layout(const sth_geom *g, sth_layout_info *out);
/* Host-only: bytes of one field of the caller's dense layout. */
int sth_host_elems(const sth_geom *g, const sth_host_layout *hl, uint64_t *elems);
int sth_device_count(int *n);

/* Creates a context on the current HIP device: one stream and three
 * hipMalloc'd fields (the reference's aligned_alloc, main.cpp:244-255). */
int sth_create(const sth_geom *g, sth_ctx **out);
int sth_destroy(sth_ctx *ctx);
int sth_get_layout(const sth_ctx *ctx, sth_layout_info *out);

/* ---- field initialisation and transfer ----------------------------------- */
/* init_arrays (iterators.cpp:292-301): every element of the allocation,
 * halos included, v = v0, s = s0, u = u0.  FP32 or FP64 storage. */
int sth_init_const(sth_ctx *ctx, float v0, float s0, float u0);
/* Seeded uniform [lo,hi) interior from the counter-based SplitMix64 stream
 * (seed, field_id) -- identical values to the oracle's generator -- and the
 * halo filled with halo_value. */
int sth_init_uniform(sth_ctx *ctx, sth_field f, uint64_t seed, int field_id, double lo,
                     double hi, float halo_value);
/* Gaussian pulse amp*exp(-d^2/(2 sigma^2)) about the interior centre plus
 * seeded uniform noise in [-noise, noise); halo 0. */
int sth_init_gaussian(sth_ctx *ctx, sth_field f, double amp, double sigma, double noise,
                      uint64_t seed, int field_id);
/* Zero field with one interior point set to value (a point source mask). */
int sth_init_point(sth_ctx *ctx, sth_field f, uint32_t i, uint32_t j, uint32_t k, float value);
/* FP32 host array in the caller's dense layout -> device field (interior and
 * the radius-wide halo; wider host halo cells are not read).  16-bit storage
 * fields are quantized with a power-of-two scale chosen from the data's max. */
int sth_upload(sth_ctx *ctx, sth_field f, const float *host, const sth_host_layout *hl);
/* Device field -> FP32 host array (interior and radius-wide halo; other host
 * cells are left untouched).  Synchronises the stream. */
int sth_download(sth_ctx *ctx, sth_field f, float *host, const sth_host_layout *hl);
/* FP64 host arrays: upload into FP32 or FP64 storage (FP32 rounds each
 * value), download from any storage (16-bit values decoded exactly). */
int sth_upload_f64(sth_ctx *ctx, sth_field f, const double *host, const sth_host_layout *hl);
int sth_download_f64(sth_ctx *ctx, sth_field f, double *host, const sth_host_layout *hl);
/* 16-bit storage: raw stored values q and the field's scale exponent
 * (value = q * 2^exp); maxabs = max |q| (upload_raw sets it from the data). */
int sth_upload_raw16(sth_ctx *ctx, sth_field f, const uint16_t *host, int exp,
                     const sth_host_layout *hl);
int sth_download_raw16(sth_ctx *ctx, sth_field f, uint16_t *host, int *exp, float *maxabs,
                       const sth_host_layout *hl);

/* Field I/O (checkpoint / restart; no counterpart in the reference, whose
 * fields live only in memory): writes dir/u.raw, v.raw, s.raw -- the storage
 * elements of the dense box with an R-wide halo, C order, k fastest -- and
 * dir/manifest.json (geometry, dtype, 16-bit exponents).  The directory must
 * exist.  sth_load restores them into a context of the same geometry and
 * dtype (STH_ERR_ARG otherwise). */
int sth_save(sth_ctx *ctx, const char *dir);
int sth_load(sth_ctx *ctx, const char *dir);

/* ---- the sweep -------------------------------------------------------------- */
/* ITERATE replacement (iterate_3d, iterators.cpp:130-157): `iterations` REF
 * sweeps v <- v - A*u + s*(coef0*u + sum c*(u[-r]+u[+r])) in the reference's
 * FMA order (bit-identical to the reference host build).  FP32 or FP64 storage.
 * One kernel launch per sweep; asynchronous. */
int sth_iterate(sth_ctx *ctx, uint32_t iterations);
/* LEAPFROG steps (formula B, README.md:10-24):
 *   v <- c0*u - v + c1*L(u) + c2[n]*s,  L(u) = sum 2*c_axis,r*(u[-r]+u[+r]),
 * then u and v swap.  c2 is a host array of `steps` per-step source
 * coefficients (NULL: all zero).  16-bit storage requires zero halos. */
int sth_leapfrog(sth_ctx *ctx, uint32_t steps, float c0, float c1, const float *c2);
/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  `steps` LEAPFROG steps as sth_leapfrog computes them (the
 * one-step sweep, bit-identical to every other schedule; FP32 or FP64
 * storage, FP64 with the FP32 weights promoted) on fields the caller
 * allocated in the device layout of sth_layout for *g: u, v, s each
 * layout.bytes, element (i,j,k) at origin + k + j*pitch + i*plane, halos
 * included.  Queued on `stream` (a hipStream_t; NULL: the null stream),
 * asynchronous.  u and v swap names every step as in sth_leapfrog: after an
 * odd step count the newest field is in v's buffer.  The library allocates
 * nothing and calls no allocating routine: the only memory it uses beyond
 * the fields is a static device array of STH_NOALLOC_SCRATCH_BYTES per GPU
 * (the work queue's counters, kept until the process ends), which every
 * call clears with a kernel; calls that overlap on one GPU race on it.
 * 16-bit storage: STH_ERR_UNSUPPORTED (its scale state lives in a context). */
#define STH_NOALLOC_SCRATCH_BYTES 512
int sth_leapfrog_noalloc(const sth_geom *g, void *u, void *v, const void *s, uint32_t steps,
                         float c0, float c1, const float *c2, void *stream);
/* Host-only: the centre coefficient that keeps a constant field constant,
 * c0 = 2 - c1 * sum 2*|2 c_axis,r| (the sum is 6 for 7 and 13 points, 12 for
 * 25; for 7 points this is README.md:17-20's c0 = 2 - 6 c1), rounded to FP32. */
int sth_leapfrog_c0(uint32_t npts, float c1, float *c0);
/* Host-only: per-step source coefficients of a Ricker wavelet (no reference
 * counterpart: the reference has no source term),
 *   c2[n] = (float)(scale * (1 - 2a) exp(-a)),  a = pi^2 f^2 (n dt - t0)^2,
 * n = 0 .. steps-1 (scale = dt^2 injects the wavelet as the leapfrog's
 * dt^2 * s term). */
int sth_ricker(uint32_t steps, double dt, double f, double t0, double scale, float *c2);
int sth_sync(sth_ctx *ctx);

/* check (iterators.cpp:304-325): interior points with |v - expected| > 0.1. */
int sth_check(sth_ctx *ctx, float expected, uint64_t *errors);
/* Interior max |x| and L2 norm of a field, accumulated in FP64 on the GPU. */
int sth_norms(sth_ctx *ctx, sth_field f, double *maxabs, double *l2);

/* Discrete energy of the LEAPFROG scheme, accumulated in FP64 on the GPU:
 *   Q = <u,u> + <v,v> - <u, (c0 + c1 L) v>   over the interior,
 * u the newest field, v the one before, L the stencil's neighbour sum with the
 * weights the context's LEAPFROG steps use (2w per pair; FP64 contexts: their
 * sth_set_weights choice), v's halo included.  For steps with s = 0 and zero
 * halos Q is conserved exactly in exact arithmetic (M = c0 + c1 L is
 * symmetric), so its drift measures the storage format's rounding.  No
 * counterpart in the reference (README.md:10-24 states the update only). */
int sth_energy(sth_ctx *ctx, float c0, float c1, double *energy);

/* Error of field fa of context a against field fb of context b (the same
 * geometry and device; storage types may differ, e.g. an FP32 or 16-bit run
 * against its FP64 twin), over the interior, accumulated in FP64 on the GPU:
 *   max_rel = max|a - b| / max|b|,  l2_rel = ||a - b||_2 / ||b||_2.
 * No counterpart in the reference (whose only result test is check());
 * synchronises both streams. */
int sth_compare(sth_ctx *a, sth_field fa, sth_ctx *b, sth_field fb, double *max_rel,
                double *l2_rel);

/* Stencil weights of an FP64-storage context (sth_weights; the default is
 * STH_WEIGHTS_FP32).  STH_WEIGHTS_REFERENCE on other storage types:
 * STH_ERR_UNSUPPORTED.  Replaces the -DUSE_DOUBLE choice of the reference's
 * build (config.h:96-100) with a run-time one. */
int sth_set_weights(sth_ctx *ctx, int mode);
int sth_get_weights(const sth_ctx *ctx, int *mode);

/* ---- schedule and timing ---------------------------------------------------- */
int sth_set_tuning(sth_ctx *ctx, const sth_tuning *t);
int sth_get_tuning(const sth_ctx *ctx, sth_tuning *t);
/* Times the candidate schedules on this context's geometry (REF or LEAPFROG
 * sweeps, 2*reps steps rounded up to a multiple of 60 each, fields are
 * modified; LEAPFROG also tries the passes of two to five steps the stencil
 * has in the storage format), keeps the fastest and returns it with its mean time
 * per step in ms.  If the environment variable
 * STH_TUNE_CACHE names a file, a result stored there for the same (stencil,
 * dtype, mode, xs, ys, zs) is applied without timing, and new results are
 * appended to it. */
int sth_autotune(sth_ctx *ctx, int leapfrog, uint32_t reps, sth_tuning *best, float *best_ms);
/* Kernel timing: HIP events recorded on the context's stream around every
 * sweep launch between timer_start and timer_read.  timer_read synchronises
 * and returns the summed kernel time (ms) and the number of launches. */
int sth_timer_start(sth_ctx *ctx);
int sth_timer_read(sth_ctx *ctx, double *kernel_ms, uint64_t *launches);
/* Diagnostic with no counterpart in the reference: a plain streaming copy of
 * `bytes` (at most one field) from the u field into the v field, `reps`
 * times; returns the mean ms per copy.  Overwrites v.  The copy moves 2*bytes
 * of HBM traffic per launch: the achievable-bandwidth context for the sweep's
 * roofline, and the calibration kernel for the traffic counters. */
int sth_stream_copy(sth_ctx *ctx, uint64_t bytes, uint32_t reps, double *ms);
/* Diagnostic with no counterpart in the reference: a streaming kernel that
 * reads `reads` (0-3) and writes `writes` (0-2) streams of `bytes` each
 * (reads: u, s, v; writes: v in place and the twin of u, allocated if needed;
 * fields are modified), lane_bytes = 4 (the sweep's access width) or 16 per
 * lane access, four accesses per lane in flight, over a grid of `grid`
 * workgroups of 256 (0: 8 per CU); returns the mean ms per launch.  HBM
 * traffic per launch: (reads + writes) * bytes.  The achievable-bandwidth
 * ceilings of the sweep's access mixes (3 reads + 1 write: the one-step
 * sweep; 3 + 2: the fused pass) and the calibration kernels of the traffic
 * counters at either access width. */
int sth_bw_probe(sth_ctx *ctx, uint32_t reads, uint32_t writes, uint32_t lane_bytes,
                 uint64_t bytes, uint32_t grid, uint32_t reps, double *ms);
/* Placement of the field buffers, the counterpart of the reference's shifted
 * array starts (main.cpp:243-276, "to minimize CL bank conflicts").  On
 * MI355X the sweep's time depends on the physical memory behind its fields
 * (a shift of the virtual start changes nothing measurable): the same FP32
 * 25-point pass at 1040^3 runs about 2.62 or about 2.82 ms per step depending
 * on the allocation, with the mixed read/write streaming rate and the memory
 * latency of the slow allocations lower and higher.  sth_place allocates up
 * to `tries` (0-8) further sets of the context's buffers (a, b, s and the
 * twins when the LEAPFROG default is a fused pass), as many as fit while a
 * sixteenth of the device (at least 4 GiB) stays free, all held at once;
 * times the context's LEAPFROG default on each and on the current set; keeps
 * the fastest and frees the others.  ms[0 .. tries] (if not null) receives
 * the ms per step of the current set and of each candidate (0: not
 * allocated), *chosen (if not null) the index kept.  Call it right after
 * sth_create and after sth_set_tuning: field contents (and, for 16-bit
 * storage, the scale state) are undefined afterwards, until the fields are
 * initialised or uploaded.  Any storage type; not for a batch's shots. */
int sth_place(sth_ctx *ctx, uint32_t tries, float *ms, uint32_t *chosen);

/* ---- batching: many independent grids ("shots") advanced together ------ */
/* The reference runs one grid per process; a batch holds nshots grids of
 * uneven size (dims = nshots x {xs, ys, zs}), all with the same stencil and
 * FP32 storage, and advances them together with the LEAPFROG update.
 * sth_batch_shot returns the context of one shot, through which every
 * single-grid call (init, upload, download, norms) works on that shot; the
 * batch orders its stream against the shots' streams. */
typedef struct sth_batch sth_batch;

#define STH_BATCH_PERSISTENT 0 /* one persistent launch per step over all shots */
#define STH_BATCH_SEQUENTIAL 1 /* one plain sweep launch per shot per step */
#define STH_BATCH_TWO_STEP 2   /* one persistent launch per two steps over all shots (two
                                  LEAPFROG steps per pass, bit-identical; an odd last step
                                  runs as STH_BATCH_PERSISTENT) */
#define STH_BATCH_THREE_STEP 3 /* 7 and 13 points: one persistent launch per three steps
                                  (the one or two steps left over run as
                                  STH_BATCH_PERSISTENT); 25 points: STH_ERR_UNSUPPORTED */
#define STH_BATCH_FUSED 4      /* one persistent launch per pass of the stencil's single-grid
                                  default length (7 points: 5 steps, 13: 3, 25: 2); steps
                                  left over run as STH_BATCH_PERSISTENT */
#define STH_BATCH_GRAPH 0x100  /* flag: capture the call's time loop in one hipGraph */

int sth_batch_create(const uint32_t *dims, uint32_t nshots, uint32_t npts, uint32_t dtype,
                     sth_batch **out);
int sth_batch_destroy(sth_batch *b);
sth_ctx *sth_batch_shot(sth_batch *b, uint32_t shot);
uint32_t sth_batch_size(const sth_batch *b);
/* Tile knobs of the persistent method, shared by all shots; planes = i-planes
 * per work item (0: the whole x-range); grid_blocks = persistent grid (0: the
 * resident capacity).  The fused methods (STH_BATCH_TWO_STEP, _THREE_STEP,
 * _FUSED) take t->chunks i-chunks per shot (0: about six rounds of their
 * persistent grid) and ignore planes.  The sequential method launches each
 * shot with the shot's own tuning (sth_set_tuning on sth_batch_shot;
 * default: the single-grid default tile), one workgroup per tile. */
int sth_batch_set_tuning(sth_batch *b, const sth_tuning *t, uint32_t planes,
                         uint32_t grid_blocks);
int sth_batch_get_tuning(const sth_batch *b, sth_tuning *t, uint32_t *planes,
                         uint32_t *grid_blocks);
/* `steps` LEAPFROG steps of every shot (c2: per-step source coefficients,
 * NULL = 0); method = STH_BATCH_PERSISTENT, STH_BATCH_SEQUENTIAL,
 * STH_BATCH_TWO_STEP, STH_BATCH_THREE_STEP or STH_BATCH_FUSED, optionally |
 * STH_BATCH_GRAPH.
 * The fused methods allocate two more buffers per shot on first use (the
 * twins of u and v).  Asynchronous. */
int sth_batch_leapfrog(sth_batch *b, uint32_t steps, float c0, float c1, const float *c2,
                       uint32_t method);
int sth_batch_sync(sth_batch *b);
/* Times tile x planes-per-item candidates of the persistent one-step launch,
 * then, with the fastest of those, the i-chunk counts of STH_BATCH_FUSED
 * (fields are modified); keeps both choices (the chunk count in
 * best->chunks) and returns the fastest time per step of either method.
 * sth_batch_best_method then names the faster method (STH_BATCH_PERSISTENT
 * or STH_BATCH_FUSED). */
int sth_batch_autotune(sth_batch *b, uint32_t reps, sth_tuning *best, uint32_t *planes,
                       float *best_ms);
int sth_batch_best_method(const sth_batch *b, uint32_t *method);
/* Accumulates the GPU time of every sth_batch_leapfrog call (events around
 * the launched work) between start and read. */
int sth_batch_timer_start(sth_batch *b);
int sth_batch_timer_read(sth_batch *b, double *ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif

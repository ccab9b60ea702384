/*
 * stencil_oracle.h -- CPU restatement of the stencil3d time-stepping sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * (libstencil_hip): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time a CPU baseline.
 * Nothing in the product path calls it.
 *
 * Layout: the reference's dense padded layout (grid.cpp:20-25, stencils.h:5):
 *   jmult = zs + 2*zh, imult = jmult*(ys + 2*yh), ofs = zh + yh*jmult + xh*imult,
 *   element (i,j,k) at ofs + k + j*jmult + i*imult, i slowest, k unit stride.
 *   Indices are 64-bit here (the reference uses 32-bit unsigned; identical for
 *   every size below 2^32 elements).
 *
 * Pinning (how each function is checked, tests/test_oracle*.py):
 *   orc_ref_sweep_f32   bit-exact vs golden vectors written by the compiled
 *                       reference (tests/golden/) and vs oracle/_ref when built;
 *                       KAT recurrence (u=s=1, v=0) from iterators.cpp:292-325.
 *   orc_ref_sweep_f64   REF closed form v_N = v_0 + N*(s*P(u) - A*u) (exact in
 *                       rational arithmetic, checked with Python fractions on
 *                       tiny grids) and vs the reference's USE_DOUBLE build.
 *   orc_leapfrog_f32/64 no reference implementation (README.md:10-24 only):
 *                       pinned by brute force (Python fractions) at 5^3, by
 *                       exact discrete sine modes (r=1, zero halo), discrete
 *                       energy conservation and plane-wave dispersion.
 *   orc_leapfrog_q16    fp16/bf16 storage emulation: conversions pinned by
 *                       IEEE special cases, exhaustive round trips and numpy's
 *                       float16 rounding; scaling rule pinned by invariants.
 *   orc_leapfrog_q16_passes (and _pass2, its passes of two)  the same steps
 *                       in passes: pinned by equality with orc_leapfrog_q16
 *                       for passes of one step and wherever the exponent
 *                       rules agree, and by its own invariants (no overflow,
 *                       exponent never below the one-step rule's, the bound
 *                       recursion checked against its closed form).
 *   orc_fill_uniform    SplitMix64 published output vectors.
 */
#ifndef STENCIL_ORACLE_H
#define STENCIL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t xs, ys, zs; /* interior extents along i (slowest), j, k (unit stride) */
  int64_t xh, yh, zh; /* halo widths */
} orc_geom;

int64_t orc_jmult(const orc_geom *g);
int64_t orc_imult(const orc_geom *g);
int64_t orc_ofs(const orc_geom *g);
int64_t orc_gsize(const orc_geom *g);

/* Stencil constants of stencil_7.h:4-8, stencil_13.h:13-22, stencil_25.h:19-34
 * (non-ISOTROPIC), FP32-rounded exactly as the reference computes them.
 * w[axis*4 + (r-1)], axis 0 = i (cx), 1 = j (cy), 2 = k (cz).
 * Returns the radius (1, 2, 4) or -1 for an unsupported point count. */
int orc_coeffs(int npts, float *coef0, float *A, float w[12]);

/* Evaluation order of the neighbour pairs (stencil_7.h:20, stencil_13.h:40-43,
 * stencil_25.h:64-67): fills axis[t], r[t]; returns the number of terms. */
int orc_terms(int npts, int axis[12], int r[12]);

/* REF sweep (formula A), `iterations` in-place sweeps of v (u, s read-only):
 * iterators.cpp:130-157 with the j-bound taken as ys (the reference loops to
 * xs at iterators.cpp:146), per-point update stencil_N_ref_1 in the FMA order
 * the reference host build produces:
 *   p = fmaf(coef0, u0, w_0*pair_0); p = fmaf(w_t, pair_t, p) ...;
 *   v = fmaf(s0, p, fmaf(-A, u0, v0)),  pair = u[-r] + u[+r].
 * nthreads <= 0 keeps the OpenMP default. Returns 0 or -1 (bad args). */
int orc_ref_sweep_f32(int npts, const orc_geom *g, float *v, const float *s,
                      const float *u, int iterations, int nthreads);

/* FP64 REF sweep. coeff_mode 0: the FP32 constants promoted to double (the
 * same discrete operator as the FP32 sweep); 1: the constants as the
 * reference's -DUSE_DOUBLE build forms them (config.h:96-100: float literals
 * promoted, then cy = cx*1.1f and cz = cx*0.9f evaluated in double). */
int orc_ref_sweep_f64(int npts, const orc_geom *g, double *v, const double *s,
                      const double *u, int iterations, int coeff_mode, int nthreads);

/* LEAPFROG sweep (formula B, README.md:10-24) over the same cross stencil:
 *   L   = wh_0*pair_0; L = fmaf(wh_t, pair_t, L) ...   (wh = 2*w, see below)
 *   out = fmaf(c2[n], s0, fmaf(c1, L, fmaf(c0, u0, -v0)))
 * written into the v buffer, then u and v swap roles.  wh_axis,r = 2*w_axis,r
 * so the 7-point weights are exactly 1 and c0 = 2 - 6*c1 is the README's
 * scheme.  a holds u at step 0 and b holds v; c2[n] is the per-step source
 * coefficient.  Returns the index (0 = a, 1 = b) of the buffer holding the
 * newest field after `steps` steps, or -1. */
int orc_leapfrog_f32(int npts, const orc_geom *g, float *a, float *b, const float *s,
                     float c0, float c1, const float *c2, int steps, int nthreads);
int orc_leapfrog_f64(int npts, const orc_geom *g, double *a, double *b, const double *s,
                     double c0, double c1, const double *c2, int steps, int nthreads);

/* Sum over terms of 2*|wh| (the constant-mode gain of L), in double. */
double orc_lap_gain(int npts);

/* ---- 16-bit storage form ------------------------------------------------ */
/* fmt: 1 = IEEE binary16, 2 = bfloat16.  Conversions round to nearest even. */
uint16_t orc_f32_to_16(int fmt, float x);
float orc_16_to_f32(int fmt, uint16_t h);
/* Power-of-two scale exponent for a field whose magnitude is bounded by B:
 * the smallest e with B*2^-e < 2^15 (e from frexp); B <= 0 -> 0. */
int orc_scale_exp(double B);
/* Quantize an FP32 field into 16-bit storage with exponent e
 * (q = rne16(x * 2^-e)) over the whole allocation; returns max |q|. */
float orc_quantize(int fmt, const orc_geom *g, const float *x, uint16_t *q, int e);
void orc_dequantize(int fmt, const orc_geom *g, const uint16_t *q, int e, float *x);
/* LEAPFROG with 16-bit storage, FP32 arithmetic.  A buffer holds numbers q
 * with value q*2^e; exps[0..2] / maxs[0..2] are e and max |q| of buffers a, b
 * and s (in/out).  Before each step the output exponent is orc_scale_exp(B),
 *   B = (|c0| + |c1|*gain)*M_u + M_v + |c2[n]|*M_s,  M = max|q|*2^e (double);
 * the update runs on the stored numbers with the powers of two folded into
 * the coefficients (exact scalings):
 *   L on q_u as in the FP32 order, out = fmaf(k2, q_s, fmaf(k1, L,
 *   fmaf(k0, q_u, -(q_v*kv)))),  k0 = c0 2^(eu-eo), k1 = c1 2^(eu-eo),
 *   k2 = c2[n] 2^(es-eo), kv = 2^(ev-eo);
 * q_out = rne16(out) and the buffer's max is max |q_out| over the interior.
 * Halos must be zero. */
int orc_leapfrog_q16(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                     const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                     const float *c2, int steps, int nthreads);
/* The same update in passes of plen steps (1..8; the GPU's FP16 multi-step
 * passes): step M of a pass takes its output exponent from the bound
 *   B_M = (|c0| + |c1| gain) B_{M-1} + B_{M-2} + |c2[n+M-1]| M_s,
 *   B_0 = M_u, B_{-1} = M_v
 * (B_1: the one-step rule's bound; later bounds stand in for the maxima of
 * outputs the pass does not know before it ends).  The steps left over run
 * as one shorter pass.  plen = 1 is orc_leapfrog_q16. */
int orc_leapfrog_q16_passes(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                            const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                            const float *c2, int steps, int plen, int nthreads);
/* orc_leapfrog_q16_passes with plen = 2 (the GPU's FP16 two-step pass). */
int orc_leapfrog_q16_pass2(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                           const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                           const float *c2, int steps, int nthreads);

/* ---- initialisers, checks, norms ----------------------------------------- */
uint64_t orc_splitmix64_mix(uint64_t z);
/* Counter-based uniform value of field `field` at coordinate (i,j,k) (halo
 * coordinates are negative or >= the extent):
 *   key = mix(seed + (field+1)*G); x = mix(key + (c+1)*G), G = 0x9E3779B97F4A7C15,
 *   c = (i+2^20)<<42 | (j+2^20)<<21 | (k+2^20);  U = (x>>40)*2^-24;
 *   value = (float)(lo + (hi-lo)*U)   (double arithmetic).               */
double orc_uniform01(uint64_t seed, int field, int64_t i, int64_t j, int64_t k);
/* Fill interior with seeded uniform [lo,hi), halo and padding with halo_value. */
void orc_fill_uniform(const orc_geom *g, float *a, uint64_t seed, int field, double lo,
                      double hi, float halo_value);
void orc_fill_uniform_f64(const orc_geom *g, double *a, uint64_t seed, int field, double lo,
                          double hi, double halo_value);
/* Gaussian pulse amp*exp(-d^2/(2 sigma^2)) about the interior centre
 * ((xs-1)/2, (ys-1)/2, (zs-1)/2) plus seeded uniform noise in [-noise, noise);
 * halo = 0.  Double arithmetic; the f32 variant rounds the double value. */
void orc_fill_gaussian(const orc_geom *g, float *a, double amp, double sigma, double noise,
                       uint64_t seed, int field);
void orc_fill_gaussian_f64(const orc_geom *g, double *a, double amp, double sigma,
                           double noise, uint64_t seed, int field);
/* Ricker wavelet (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2). */
double orc_ricker(double f, double tau);

/* init_arrays (iterators.cpp:292-301): v=0, u=1, s=1 over n elements. */
void orc_init_arrays(float *v, float *s, float *u, int64_t n);
/* check (iterators.cpp:304-325): interior points with |v - expected| > 0.1f. */
int64_t orc_check(const orc_geom *g, const float *v, float expected);
/* Interior max |x| and sqrt(sum x^2) in double. */
void orc_norms_f32(const orc_geom *g, const float *x, double *maxabs, double *l2);
void orc_norms_f64(const orc_geom *g, const double *x, double *maxabs, double *l2);
/* Interior error of x against the FP64 reference r:
 * maxrel = max|x-r| / max|r|, l2rel = ||x-r||_2 / ||r||_2. */
void orc_error_f32_vs_f64(const orc_geom *g, const float *x, const double *r, double *maxrel,
                          double *l2rel);

#ifdef __cplusplus
}
#endif
#endif

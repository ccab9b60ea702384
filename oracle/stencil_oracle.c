/*
 * stencil_oracle.c -- CPU restatement of the stencil3d sweep (test
 * infrastructure only; see stencil_oracle.h for what may use it and how each
 * function is pinned).
 *
 * Built with -ffp-contract=off: every fused multiply-add below is an explicit
 * fmaf()/fma() call and every other product is rounded on its own, which is
 * the order the reference host build produces (stencil_N.h compiled with
 * clang++/g++ -O3 and an FMA-capable -march: the first product c*pair is
 * rounded, coef0*u0 and every later term are fused, v0 - A*u0 is fused; the
 * golden vectors of that build and the live comparison with it pin this).
 */
#include "stencil_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

int64_t orc_jmult(const orc_geom *g) { return g->zs + 2 * g->zh; }
int64_t orc_imult(const orc_geom *g) { return orc_jmult(g) * (g->ys + 2 * g->yh); }
int64_t orc_ofs(const orc_geom *g) {
  return g->zh + g->yh * orc_jmult(g) + g->xh * orc_imult(g);
}
int64_t orc_gsize(const orc_geom *g) { return orc_imult(g) * (g->xs + 2 * g->xh); }

/* stencil_7.h:4-8, stencil_13.h:13-22, stencil_25.h:19-34 */
int orc_coeffs(int npts, float *coef0, float *A, float w[12]) {
  for (int t = 0; t < 12; t++) w[t] = 0.0f;
  volatile float f11 = 1.1f, f09 = 0.9f; /* keep the FP32 products rounded */
  float cx[4];
  int R;
  if (npts == 7) {
    *coef0 = 1.5f;
    *A = 3.5f; /* stencil_7.h:38 */
    w[0] = w[4] = w[8] = 0.5f;
    return 1;
  } else if (npts == 13) {
    *coef0 = 1.5f;
    *A = 3.5f; /* stencil_13.h:60 */
    cx[0] = 0.3f;
    cx[1] = 0.2f;
    R = 2;
  } else if (npts == 25) {
    *coef0 = 1.25f;
    *A = 6.25f; /* stencil_25.h:82 */
    cx[0] = 0.4f;
    cx[1] = 0.3f;
    cx[2] = 0.2f;
    cx[3] = 0.1f;
    R = 4;
  } else {
    return -1;
  }
  for (int r = 0; r < R; r++) {
    float y = cx[r] * f11, z = cx[r] * f09;
    w[0 * 4 + r] = cx[r];
    w[1 * 4 + r] = y;
    w[2 * 4 + r] = z;
  }
  return R;
}

int orc_terms(int npts, int axis[12], int r[12]) {
  int T = 0;
  if (npts == 7) { /* coef0*u0 + cx1*(i) + cy1*(j) + cz1*(k): stencil_7.h:20 */
    for (int a = 0; a < 3; a++) {
      axis[T] = a;
      r[T++] = 1;
    }
    return T;
  }
  int R = npts == 13 ? 2 : npts == 25 ? 4 : 0;
  if (!R) return -1;
  /* k (R..1), j (R..1), i (R..1): stencil_13.h:40-43, stencil_25.h:64-67 */
  for (int a = 2; a >= 0; a--)
    for (int q = R; q >= 1; q--) {
      axis[T] = a;
      r[T++] = q;
    }
  return T;
}

typedef struct {
  int T;
  int64_t off[12];
  float w[12];
  double wd[12];
  float coef0, A;
} term_tab;

static int make_tab(int npts, const orc_geom *g, term_tab *tb, int leap, int coeff_mode) {
  float coef0, A, w[12];
  int axis[12], r[12];
  if (orc_coeffs(npts, &coef0, &A, w) < 0) return -1;
  tb->T = orc_terms(npts, axis, r);
  const int64_t stride[3] = {orc_imult(g), orc_jmult(g), 1};
  for (int t = 0; t < tb->T; t++) {
    tb->off[t] = stride[axis[t]] * r[t];
    float wf = w[axis[t] * 4 + r[t] - 1];
    double wd = (double)wf;
    if (coeff_mode == 1 && npts != 7) {
      /* USE_DOUBLE build: cx literal promoted, cy = cx*1.1f, cz = cx*0.9f in double
       * (stencil_13.h:13-22, stencil_25.h:19-34; the 7-point weights are literals) */
      float cxf = w[0 * 4 + r[t] - 1];
      double cx = (double)cxf;
      wd = axis[t] == 0 ? cx : axis[t] == 1 ? cx * (double)1.1f : cx * (double)0.9f;
    }
    tb->w[t] = leap ? 2.0f * wf : wf;
    tb->wd[t] = leap ? 2.0 * wd : wd;
  }
  tb->coef0 = coef0;
  tb->A = A;
  return 0;
}

static void set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

/* ------------------------------------------------------------------------ */
/* REF sweep, FP32.  The loop nest and pragma follow iterators.cpp:137-156;
 * the point update follows stencil_N_ref_1 (stencil_7.h:31-39,
 * stencil_13.h:50-61, stencil_25.h:75-83). */
#define REF_SWEEP_F32(NAME, TT)                                                          \
  static void NAME(const orc_geom *g, const term_tab *tb, float *restrict v,             \
                   const float *restrict s, const float *restrict u, int iterations) {   \
    const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);                  \
    const int64_t xs = g->xs, ys = g->ys, zs = g->zs;                                    \
    float w[TT];                                                                         \
    int64_t off[TT];                                                                     \
    for (int t = 0; t < TT; t++) {                                                       \
      w[t] = tb->w[t];                                                                   \
      off[t] = tb->off[t];                                                               \
    }                                                                                    \
    const float coef0 = tb->coef0, mA = -tb->A;                                          \
    for (int it = 0; it < iterations; it++) {                                            \
      _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < xs; i++) {    \
        for (int64_t j = 0; j < ys; j++) {                                               \
          const int64_t base = o + i * im + j * jm;                                      \
          for (int64_t k = 0; k < zs; k++) {                                             \
            const int64_t c = base + k;                                                  \
            const float u0 = u[c];                                                       \
            float p = fmaf(coef0, u0, w[0] * (u[c - off[0]] + u[c + off[0]]));           \
            _Pragma("GCC unroll 16") for (int t = 1; t < TT; t++) p =                    \
                fmaf(w[t], u[c - off[t]] + u[c + off[t]], p);                            \
            v[c] = fmaf(s[c], p, fmaf(mA, u0, v[c]));                                    \
          }                                                                              \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
  }
REF_SWEEP_F32(ref_f32_3, 3)
REF_SWEEP_F32(ref_f32_6, 6)
REF_SWEEP_F32(ref_f32_12, 12)

int orc_ref_sweep_f32(int npts, const orc_geom *g, float *v, const float *s, const float *u,
                      int iterations, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 0, 0) < 0 || iterations < 0) return -1;
  set_threads(nthreads);
  if (tb.T == 3) ref_f32_3(g, &tb, v, s, u, iterations);
  else if (tb.T == 6) ref_f32_6(g, &tb, v, s, u, iterations);
  else ref_f32_12(g, &tb, v, s, u, iterations);
  return 0;
}

int orc_ref_sweep_f64(int npts, const orc_geom *g, double *v, const double *s,
                      const double *u, int iterations, int coeff_mode, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 0, coeff_mode) < 0 || iterations < 0) return -1;
  set_threads(nthreads);
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  const int T = tb.T;
  const double coef0 = tb.coef0, mA = -(double)tb.A;
  for (int it = 0; it < iterations; it++) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g->xs; i++)
      for (int64_t j = 0; j < g->ys; j++)
        for (int64_t k = 0; k < g->zs; k++) {
          const int64_t c = o + i * im + j * jm + k;
          const double u0 = u[c];
          double p = fma(coef0, u0, tb.wd[0] * (u[c - tb.off[0]] + u[c + tb.off[0]]));
          for (int t = 1; t < T; t++) p = fma(tb.wd[t], u[c - tb.off[t]] + u[c + tb.off[t]], p);
          v[c] = fma(s[c], p, fma(mA, u0, v[c]));
        }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LEAPFROG (formula B of README.md:10-24) in a fixed, documented FMA order. */
static inline float leap_point_f32(const float *u, int64_t c, const term_tab *tb, int T,
                                   float v0, float s0, float c0, float c1, float c2n) {
  float L = tb->w[0] * (u[c - tb->off[0]] + u[c + tb->off[0]]);
  for (int t = 1; t < T; t++) L = fmaf(tb->w[t], u[c - tb->off[t]] + u[c + tb->off[t]], L);
  return fmaf(c2n, s0, fmaf(c1, L, fmaf(c0, u[c], -v0)));
}

int orc_leapfrog_f32(int npts, const orc_geom *g, float *a, float *b, const float *s,
                     float c0, float c1, const float *c2, int steps, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 1, 0) < 0 || steps < 0) return -1;
  set_threads(nthreads);
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  const int T = tb.T;
  float *bufs[2] = {a, b};
  int cur = 0; /* index of the buffer holding u */
  for (int n = 0; n < steps; n++) {
    const float *u = bufs[cur];
    float *v = bufs[1 - cur];
    const float c2n = c2 ? c2[n] : 0.0f;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g->xs; i++)
      for (int64_t j = 0; j < g->ys; j++)
        for (int64_t k = 0; k < g->zs; k++) {
          const int64_t c = o + i * im + j * jm + k;
          v[c] = leap_point_f32(u, c, &tb, T, v[c], s[c], c0, c1, c2n);
        }
    cur = 1 - cur;
  }
  return cur;
}

int orc_leapfrog_f64(int npts, const orc_geom *g, double *a, double *b, const double *s,
                     double c0, double c1, const double *c2, int steps, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 1, 0) < 0 || steps < 0) return -1;
  set_threads(nthreads);
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  const int T = tb.T;
  double *bufs[2] = {a, b};
  int cur = 0;
  for (int n = 0; n < steps; n++) {
    const double *u = bufs[cur];
    double *v = bufs[1 - cur];
    const double c2n = c2 ? c2[n] : 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g->xs; i++)
      for (int64_t j = 0; j < g->ys; j++)
        for (int64_t k = 0; k < g->zs; k++) {
          const int64_t c = o + i * im + j * jm + k;
          double L = tb.wd[0] * (u[c - tb.off[0]] + u[c + tb.off[0]]);
          for (int t = 1; t < T; t++) L = fma(tb.wd[t], u[c - tb.off[t]] + u[c + tb.off[t]], L);
          v[c] = fma(c2n, s[c], fma(c1, L, fma(c0, u[c], -v[c])));
        }
    cur = 1 - cur;
  }
  return cur;
}

double orc_lap_gain(int npts) {
  float coef0, A, w[12];
  int axis[12], r[12];
  if (orc_coeffs(npts, &coef0, &A, w) < 0) return -1.0;
  int T = orc_terms(npts, axis, r);
  double S = 0.0;
  for (int t = 0; t < T; t++) S += 2.0 * fabs(2.0 * (double)w[axis[t] * 4 + r[t] - 1]);
  return S;
}

/* ------------------------------------------------------------------------ */
/* 16-bit storage. */
static inline uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static inline float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static uint16_t f32_to_f16_rne(float f) {
  const uint32_t x = f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu)); /* NaN */
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 (or inf) -> inf */
  if (ax >= 0x38800000u) {                                  /* normal binary16 */
    const uint32_t r = ax + 0xfffu + ((ax >> 13) & 1u);
    return (uint16_t)(sign | ((r - 0x38000000u) >> 13));
  }
  /* subnormal binary16 or zero: round(|f| * 2^24) to nearest even */
  const uint32_t e = ax >> 23;
  if (e == 0) return (uint16_t)sign; /* FP32 subnormals are far below 2^-25 */
  const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
  const uint32_t sh = 126u - e; /* >= 14 */
  if (sh > 24u) return (uint16_t)sign;
  uint32_t q = m >> sh;
  const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
  if (rem > half || (rem == half && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

static float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0x1f) return u2f(sign | 0x7f800000u | (m << 13));
  if (e != 0) return u2f(sign | ((e + 112u) << 23) | (m << 13));
  if (m == 0) return u2f(sign);
  /* subnormal: m * 2^-24, exact in FP32 */
  return sign ? -ldexpf((float)m, -24) : ldexpf((float)m, -24);
}

static uint16_t f32_to_bf16_rne(float f) {
  uint32_t x = f2u(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

uint16_t orc_f32_to_16(int fmt, float x) {
  return fmt == 2 ? f32_to_bf16_rne(x) : f32_to_f16_rne(x);
}
float orc_16_to_f32(int fmt, uint16_t h) {
  return fmt == 2 ? u2f((uint32_t)h << 16) : f16_to_f32(h);
}

int orc_scale_exp(double B) {
  if (!(B > 0.0)) return 0;
  int E;
  frexp(B, &E); /* B = m 2^E, 0.5 <= m < 1  =>  B < 2^E */
  return E - 15;
}

float orc_quantize(int fmt, const orc_geom *g, const float *x, uint16_t *q, int e) {
  const int64_t n = orc_gsize(g);
  float mx = 0.0f;
  for (int64_t t = 0; t < n; t++) {
    q[t] = orc_f32_to_16(fmt, ldexpf(x[t], -e));
    float d = fabsf(orc_16_to_f32(fmt, q[t]));
    if (d > mx) mx = d;
  }
  return mx;
}

void orc_dequantize(int fmt, const orc_geom *g, const uint16_t *q, int e, float *x) {
  const int64_t n = orc_gsize(g);
  for (int64_t t = 0; t < n; t++) x[t] = ldexpf(orc_16_to_f32(fmt, q[t]), e);
}

/* Output exponent of a 16-bit LEAPFROG step: orc_scale_exp of the bound
 * (|c0| + |c1| gain) M_u + M_v + |c2n| M_s with M = max|q| 2^e (double). */
static int q16_out_exp(const int exps[3], const float maxs[3], int cu, int cv, float c0,
                       float c1, float c2n, double gain) {
  const double Mu = ldexp((double)maxs[cu], exps[cu]);
  const double Mv = ldexp((double)maxs[cv], exps[cv]);
  const double Ms = ldexp((double)maxs[2], exps[2]);
  return orc_scale_exp((fabs((double)c0) + fabs((double)c1) * gain) * Mu + Mv +
                       fabs((double)c2n) * Ms);
}

/* One 16-bit LEAPFROG step with output exponent eo: v <- update(u, v, s) on
 * the stored numbers; returns the largest stored magnitude of the interior. */
static float q16_step(const term_tab *tb, int fmt, const orc_geom *g, const uint16_t *u,
                      uint16_t *v, const uint16_t *s, const int exps[3], int cur, int eo,
                      float c0, float c1, float c2n) {
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  const int T = tb->T;
  /* the update on stored numbers: powers of two folded into the coefficients */
  const float k0 = ldexpf(c0, exps[cur] - eo), k1 = ldexpf(c1, exps[cur] - eo);
  const float k2 = ldexpf(c2n, exps[2] - eo), kv = ldexpf(1.0f, exps[1 - cur] - eo);
  float mx = 0.0f;
#pragma omp parallel for schedule(static) reduction(max : mx)
  for (int64_t i = 0; i < g->xs; i++)
    for (int64_t j = 0; j < g->ys; j++)
      for (int64_t k = 0; k < g->zs; k++) {
        const int64_t c = o + i * im + j * jm + k;
        float L = tb->w[0] * (orc_16_to_f32(fmt, u[c - tb->off[0]]) +
                              orc_16_to_f32(fmt, u[c + tb->off[0]]));
        for (int t = 1; t < T; t++)
          L = fmaf(tb->w[t],
                   orc_16_to_f32(fmt, u[c - tb->off[t]]) + orc_16_to_f32(fmt, u[c + tb->off[t]]),
                   L);
        const float u0 = orc_16_to_f32(fmt, u[c]);
        const float v0 = orc_16_to_f32(fmt, v[c]);
        const float s0 = orc_16_to_f32(fmt, s[c]);
        const float out = fmaf(k2, s0, fmaf(k1, L, fmaf(k0, u0, -(v0 * kv))));
        const uint16_t q = orc_f32_to_16(fmt, out);
        v[c] = q;
        const float d = fabsf(orc_16_to_f32(fmt, q));
        if (d > mx) mx = d;
      }
  return mx;
}

int orc_leapfrog_q16(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                     const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                     const float *c2, int steps, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 1, 0) < 0 || steps < 0 || (fmt != 1 && fmt != 2)) return -1;
  set_threads(nthreads);
  const double gain = orc_lap_gain(npts);
  uint16_t *bufs[2] = {a, b};
  int cur = 0;
  for (int n = 0; n < steps; n++) {
    const float c2n = c2 ? c2[n] : 0.0f;
    const int eo = q16_out_exp(exps, maxs, cur, 1 - cur, c0, c1, c2n, gain);
    const float mx = q16_step(&tb, fmt, g, bufs[cur], bufs[1 - cur], s, exps, cur, eo, c0, c1,
                              c2n);
    exps[1 - cur] = eo;
    maxs[1 - cur] = mx;
    cur = 1 - cur;
  }
  return cur;
}

/* The same steps in passes of plen (the GPU's FP16 multi-step passes): the
 * output exponent of a pass's step M is orc_scale_exp of the bound
 *   B_M = (|c0| + |c1| gain) B_{M-1} + B_{M-2} + |c2[n+M-1]| M_s,
 * B_0 = M_u, B_{-1} = M_v (the stored inputs' maxima), so B_1 is the
 * one-step rule's bound and each later bound stands in for the maximum of
 * an output the pass does not know before it ends.  The steps left over run
 * as one shorter pass (a single step: the one-step rule). */
int orc_leapfrog_q16_passes(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                            const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                            const float *c2, int steps, int plen, int nthreads) {
  term_tab tb;
  if (make_tab(npts, g, &tb, 1, 0) < 0 || steps < 0 || plen < 1 || plen > 8 ||
      (fmt != 1 && fmt != 2))
    return -1;
  set_threads(nthreads);
  const double gmul = fabs((double)c0) + fabs((double)c1) * orc_lap_gain(npts);
  uint16_t *bufs[2] = {a, b};
  int cur = 0;
  for (int n = 0; n < steps;) {
    const int len = steps - n < plen ? steps - n : plen;
    const double Ms = ldexp((double)maxs[2], exps[2]);
    double b2 = ldexp((double)maxs[1 - cur], exps[1 - cur]); /* B_{M-2} */
    double b1 = ldexp((double)maxs[cur], exps[cur]);         /* B_{M-1} */
    for (int M = 1; M <= len; M++, n++) {
      const float c2n = c2 ? c2[n] : 0.0f;
      const double B = gmul * b1 + b2 + fabs((double)c2n) * Ms;
      const int eo = orc_scale_exp(B);
      const float mx = q16_step(&tb, fmt, g, bufs[cur], bufs[1 - cur], s, exps, cur, eo, c0, c1,
                                c2n);
      exps[1 - cur] = eo;
      maxs[1 - cur] = mx;
      cur = 1 - cur;
      b2 = b1;
      b1 = B;
    }
  }
  return cur;
}

int orc_leapfrog_q16_pass2(int npts, int fmt, const orc_geom *g, uint16_t *a, uint16_t *b,
                           const uint16_t *s, int exps[3], float maxs[3], float c0, float c1,
                           const float *c2, int steps, int nthreads) {
  return orc_leapfrog_q16_passes(npts, fmt, g, a, b, s, exps, maxs, c0, c1, c2, steps, 2,
                                 nthreads);
}

/* ------------------------------------------------------------------------ */
#define SM_GAMMA 0x9E3779B97F4A7C15ull

uint64_t orc_splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline uint64_t field_key(uint64_t seed, int field) {
  return orc_splitmix64_mix(seed + (uint64_t)(field + 1) * SM_GAMMA);
}

static inline double u01_key(uint64_t key, int64_t i, int64_t j, int64_t k) {
  const uint64_t c = ((uint64_t)(i + (1 << 20)) << 42) | ((uint64_t)(j + (1 << 20)) << 21) |
                     (uint64_t)(k + (1 << 20));
  const uint64_t x = orc_splitmix64_mix(key + (c + 1) * SM_GAMMA);
  return (double)(x >> 40) * 0x1.0p-24;
}

double orc_uniform01(uint64_t seed, int field, int64_t i, int64_t j, int64_t k) {
  return u01_key(field_key(seed, field), i, j, k);
}

#define FILL_UNIFORM(NAME, TYPE)                                                       \
  void NAME(const orc_geom *g, TYPE *a, uint64_t seed, int field, double lo, double hi, \
            TYPE halo_value) {                                                         \
    const int64_t n = orc_gsize(g), im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g); \
    const uint64_t key = field_key(seed, field);                                       \
    for (int64_t t = 0; t < n; t++) a[t] = halo_value;                                 \
    _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < g->xs; i++)   \
        for (int64_t j = 0; j < g->ys; j++) for (int64_t k = 0; k < g->zs; k++)        \
            a[o + i * im + j * jm + k] = (TYPE)(lo + (hi - lo) * u01_key(key, i, j, k)); \
  }
FILL_UNIFORM(orc_fill_uniform, float)
FILL_UNIFORM(orc_fill_uniform_f64, double)

static double gauss_val(const orc_geom *g, int64_t i, int64_t j, int64_t k, double amp,
                        double sigma, double noise, uint64_t key) {
  const double di = (double)i - 0.5 * (double)(g->xs - 1);
  const double dj = (double)j - 0.5 * (double)(g->ys - 1);
  const double dk = (double)k - 0.5 * (double)(g->zs - 1);
  const double d2 = di * di + dj * dj + dk * dk;
  double val = amp * exp(-d2 / (2.0 * sigma * sigma));
  if (noise != 0.0) val += -noise + 2.0 * noise * u01_key(key, i, j, k);
  return val;
}

#define FILL_GAUSS(NAME, TYPE)                                                          \
  void NAME(const orc_geom *g, TYPE *a, double amp, double sigma, double noise,          \
            uint64_t seed, int field) {                                                  \
    const int64_t n = orc_gsize(g), im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g); \
    const uint64_t key = field_key(seed, field);                                        \
    memset(a, 0, (size_t)n * sizeof(TYPE));                                             \
    _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < g->xs; i++)    \
        for (int64_t j = 0; j < g->ys; j++) for (int64_t k = 0; k < g->zs; k++)         \
            a[o + i * im + j * jm + k] = (TYPE)gauss_val(g, i, j, k, amp, sigma, noise, key); \
  }
FILL_GAUSS(orc_fill_gaussian, float)
FILL_GAUSS(orc_fill_gaussian_f64, double)

double orc_ricker(double f, double tau) {
  const double a = M_PI * M_PI * f * f * tau * tau;
  return (1.0 - 2.0 * a) * exp(-a);
}

/* iterators.cpp:292-301 */
void orc_init_arrays(float *v, float *s, float *u, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < n; t++) {
    v[t] = 0.0f;
    u[t] = 1.0f;
    s[t] = 1.0f;
  }
}

/* iterators.cpp:304-325 */
int64_t orc_check(const orc_geom *g, const float *v, float expected) {
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  int64_t err = 0;
#pragma omp parallel for collapse(2) reduction(+ : err) schedule(static)
  for (int64_t i = 0; i < g->xs; i++)
    for (int64_t j = 0; j < g->ys; j++)
      for (int64_t k = 0; k < g->zs; k++)
        if (fabsf(v[o + i * im + j * jm + k] - expected) > 0.1f) err++;
  return err;
}

#define NORMS(NAME, TYPE)                                                                 \
  void NAME(const orc_geom *g, const TYPE *x, double *maxabs, double *l2) {               \
    const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);                   \
    double mx = 0.0, ss = 0.0;                                                            \
    _Pragma("omp parallel for reduction(max : mx) reduction(+ : ss) schedule(static)") for ( \
        int64_t i = 0; i < g->xs; i++) for (int64_t j = 0; j < g->ys; j++) for (int64_t k = 0; \
                                                                                k < g->zs; k++) { \
      const double d = (double)x[o + i * im + j * jm + k];                                \
      if (fabs(d) > mx) mx = fabs(d);                                                     \
      ss += d * d;                                                                        \
    }                                                                                     \
    *maxabs = mx;                                                                         \
    *l2 = sqrt(ss);                                                                       \
  }
NORMS(orc_norms_f32, float)
NORMS(orc_norms_f64, double)

void orc_error_f32_vs_f64(const orc_geom *g, const float *x, const double *r, double *maxrel,
                          double *l2rel) {
  const int64_t im = orc_imult(g), jm = orc_jmult(g), o = orc_ofs(g);
  double me = 0.0, mr = 0.0, se = 0.0, sr = 0.0;
#pragma omp parallel for reduction(max : me, mr) reduction(+ : se, sr) schedule(static)
  for (int64_t i = 0; i < g->xs; i++)
    for (int64_t j = 0; j < g->ys; j++)
      for (int64_t k = 0; k < g->zs; k++) {
        const int64_t c = o + i * im + j * jm + k;
        const double e = (double)x[c] - r[c];
        if (fabs(e) > me) me = fabs(e);
        if (fabs(r[c]) > mr) mr = fabs(r[c]);
        se += e * e;
        sr += r[c] * r[c];
      }
  *maxrel = mr > 0.0 ? me / mr : me;
  *l2rel = sr > 0.0 ? sqrt(se / sr) : sqrt(se);
}

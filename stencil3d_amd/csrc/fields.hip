// fields.hip -- field initialisation, check and norm kernels of libstencil_hip.
//
// These run once per run (untimed in the reference too: init_arrays
// iterators.cpp:292-301, check iterators.cpp:304-325).  Every kernel walks the
// device layout of sth_common.h: element e of a field is the cell
//   t = e - align + R;  i = t / plane - R;  j = (t % plane) / pitch - R;
//   k = (t % plane) % pitch - R,
// a cell with i < xs+R, j < ys+R, k < zs+R; every other element is padding.
#include <hip/hip_runtime.h>

#include "fields.h"

namespace sth {

struct Cell {
  int i, j, k;
  bool cell;      // inside the halo-padded box
  bool interior;  // inside the interior
};

__device__ inline Cell decode(const FieldDesc &d, uint64_t e) {
  Cell c{0, 0, 0, false, false};
  const int64_t t = (int64_t)e - (int64_t)d.align + d.R;
  if (t < 0) return c;
  const int64_t ip = t / d.plane, rem = t % d.plane;
  const int64_t jp = rem / d.pitch, kp = rem % d.pitch;
  c.i = (int)ip - d.R;
  c.j = (int)jp - d.R;
  c.k = (int)kp - d.R;
  c.cell = c.i < d.xs + d.R && c.j < d.ys + d.R && c.k < d.zs + d.R;
  c.interior = c.cell && c.i >= 0 && c.j >= 0 && c.k >= 0 && c.i < d.xs && c.j < d.ys && c.k < d.zs;
  return c;
}

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based SplitMix64 value in [0,1) at coordinate (i,j,k) (the same
// generator as the oracle: value = mix(key + (c+1)*G), top 24 bits).
__device__ inline double u01(uint64_t key, int i, int j, int k) {
  const uint64_t c = ((uint64_t)(i + (1 << 20)) << 42) | ((uint64_t)(j + (1 << 20)) << 21) |
                     (uint64_t)(k + (1 << 20));
  const uint64_t x = mix64(key + (c + 1) * 0x9E3779B97F4A7C15ull);
  return (double)(x >> 40) * 0x1.0p-24;
}

// Field value at a cell in double; FP32 and 16-bit storage round it to FP32.
__device__ inline double gen_value(const FillSpec &f, const FieldDesc &d, const Cell &c) {
  if (!c.cell) return f.kind == FILL_CONST ? f.a : 0.0;
  switch (f.kind) {
    case FILL_CONST:
      return f.a;
    case FILL_UNIFORM:
      if (!c.interior) return f.halo;
      return f.lo + (f.hi - f.lo) * u01(f.key, c.i, c.j, c.k);
    case FILL_GAUSS: {
      if (!c.interior) return 0.0;
      const double di = (double)c.i - 0.5 * (double)(d.xs - 1);
      const double dj = (double)c.j - 0.5 * (double)(d.ys - 1);
      const double dk = (double)c.k - 0.5 * (double)(d.zs - 1);
      const double d2 = di * di + dj * dj + dk * dk;
      double val = f.lo * exp(-d2 / (2.0 * f.hi * f.hi));
      if (f.noise != 0.0) val += -f.noise + 2.0 * f.noise * u01(f.key, c.i, c.j, c.k);
      return val;
    }
    case FILL_POINT:
      return (c.interior && c.i == f.pi && c.j == f.pj && c.k == f.pk) ? f.a : 0.0;
  }
  return 0.0;
}

// pass 0: max |value| over cells (atomicMax of FP32 bits); pass 1: store.
__global__ void fill_kernel(FieldDesc d, FillSpec f, void *dst, int fmt, int e, int pass,
                            uint32_t *maxbits) {
  const uint64_t n = d.elems;
  float mx = 0.0f;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const Cell c = decode(d, idx);
    const double xd = gen_value(f, d, c);
    const float x = (float)xd;
    if (pass == 0) {
      if (c.cell) mx = fmaxf(mx, fabsf(x));
    } else if (fmt == FMT_F64) {
      ((double *)dst)[idx] = xd;
    } else if (fmt == FMT_F32) {
      ((float *)dst)[idx] = x;
    } else if (fmt == FMT_F16) {
      ((uint16_t *)dst)[idx] = Codec<FMT_F16>::enc(ldexpf(x, -e));
    } else {
      ((uint16_t *)dst)[idx] = Codec<FMT_BF16>::enc(ldexpf(x, -e));
    }
  }
  if (pass == 0) {
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(mx));
  }
}

__device__ inline double load_dec(const void *p, uint64_t idx, int fmt, int e) {
  if (fmt == FMT_F64) return ((const double *)p)[idx];
  if (fmt == FMT_F32) return ((const float *)p)[idx];
  if (fmt == FMT_F16) return ldexpf(Codec<FMT_F16>::dec(((const uint16_t *)p)[idx]), e);
  return ldexpf(Codec<FMT_BF16>::dec(((const uint16_t *)p)[idx]), e);
}

// check(): interior cells with |v - expected| > 0.1f (iterators.cpp:319).
__global__ void check_kernel(FieldDesc d, const void *v, int fmt, int e, float expected,
                             unsigned long long *errors) {
  unsigned long long cnt = 0;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < d.elems;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const Cell c = decode(d, idx);
    if (c.interior) {
      // FP32: the reference's float test; FP64: the same test in double
      const double x = load_dec(v, idx, fmt, e);
      const bool bad = fmt == FMT_F64 ? fabs(x - (double)expected) > (double)0.1f
                                      : fabsf((float)x - expected) > 0.1f;
      if (bad) cnt++;
    }
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(errors, cnt);
}

// Largest of non-negative doubles by their bit patterns (which order like
// the values), and its FP32 rounding (rounding is monotone) for the callers
// that keep FP32 maxima.
__device__ inline void max_both(uint32_t *maxbits, double m) {
  atomicMax((unsigned long long *)(maxbits + 2), (unsigned long long)__double_as_longlong(m));
  atomicMax(maxbits, __float_as_uint((float)m));
}

// Interior max |x| and per-block sums of x^2 in FP64 (summed in block order
// by norms_finish, so the L2 norm is deterministic).  maxbits[0]: the max as
// FP32 bits; maxbits[2..3]: as FP64 bits.
__global__ void norms_kernel(FieldDesc d, const void *x, int fmt, int e, double *partial,
                             uint32_t *maxbits) {
  __shared__ double ss[16], sm[16];
  double s = 0.0, mx = 0.0;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < d.elems;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const Cell c = decode(d, idx);
    if (c.interior) {
      const double val = load_dec(x, idx, fmt, e);
      s += val * val;
      mx = fmax(mx, fabs(val));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    mx = fmax(mx, __shfl_xor(mx, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ss[w] = s;
    sm[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0, m = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      t += ss[q];
      m = fmax(m, sm[q]);
    }
    partial[blockIdx.x] = t;
    max_both(maxbits, m);
  }
}

// Interior difference of two fields of one geometry in two layouts (storage
// types may differ): per-block sums of (x - r)^2 and r^2 in FP64 and the
// maxima of |x - r| and |r| (FP32 bits).
__global__ void compare_kernel(FieldDesc dx, const void *x, int fx, int ex, FieldDesc dr,
                               const void *r, int fr, int er, double *pdiff, double *pref,
                               uint32_t *maxdiff, uint32_t *maxref) {
  __shared__ double sd[16], sr[16], md[16], mr[16];
  double s1 = 0.0, s2 = 0.0, m1 = 0.0, m2 = 0.0;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < dx.elems;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const Cell c = decode(dx, idx);
    if (c.interior) {
      const uint64_t jdx = dr.align - dr.R + (uint64_t)(c.i + dr.R) * dr.plane +
                           (uint64_t)(c.j + dr.R) * dr.pitch + (uint64_t)(c.k + dr.R);
      const double a = load_dec(x, idx, fx, ex), b = load_dec(r, jdx, fr, er);
      const double d = a - b;
      s1 += d * d;
      s2 += b * b;
      m1 = fmax(m1, fabs(d));
      m2 = fmax(m2, fabs(b));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
    m1 = fmax(m1, __shfl_xor(m1, o, 64));
    m2 = fmax(m2, __shfl_xor(m2, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sd[w] = s1;
    sr[w] = s2;
    md[w] = m1;
    mr[w] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t1 = 0.0, t2 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      t1 += sd[q];
      t2 += sr[q];
      a1 = fmax(a1, md[q]);
      a2 = fmax(a2, mr[q]);
    }
    pdiff[blockIdx.x] = t1;
    pref[blockIdx.x] = t2;
    max_both(maxdiff, a1);
    max_both(maxref, a2);
  }
}

// Discrete energy of the LEAPFROG scheme (s = 0, zero halos): with
// M = c0 + c1 L symmetric, Q = <u,u> + <v,v> - <u, M v> is the same after
// every step u' = M u - v, v' = u.  Per-block FP64 sums over the interior of
// u^2 + v^2 - u (c0 v + c1 L v), L v = sum_t w_t (v[c - o_t] + v[c + o_t]),
// the neighbours read from v's halo-padded box.
__global__ void energy_kernel(FieldDesc d, const void *u, const void *v, int fmt, int eu, int ev,
                              EnergySpec sp, double *partial) {
  __shared__ double ss[16];
  double s = 0.0;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < d.elems;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const Cell c = decode(d, idx);
    if (c.interior) {
      const double x = load_dec(u, idx, fmt, eu), y = load_dec(v, idx, fmt, ev);
      double L = 0.0;
      for (int t = 0; t < sp.T; t++)
        L += sp.w[t] * (load_dec(v, idx - (uint64_t)sp.off[t], fmt, ev) +
                        load_dec(v, idx + (uint64_t)sp.off[t], fmt, ev));
      s += x * x + y * y - x * (sp.c0 * y + sp.c1 * L);
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) ss[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += ss[q];
    partial[blockIdx.x] = t;
  }
}

__global__ void norms_finish(const double *partial, int n, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < n; b++) t += partial[b];
    out[0] = t;
  }
}

// Source-free zero fill of a small device buffer (graph-safe replacement for
// hipMemsetAsync).
__global__ void zero_u32(uint32_t *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

// Plain streaming copy (16-byte vectors, grid-stride): the reference point
// for the HBM bandwidth a simple kernel reaches, and the calibration kernel
// for the FETCH_SIZE / WRITE_SIZE counters.
__global__ void stream_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                   uint64_t n4) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4;
       t += (uint64_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

// Bandwidth probe: NR streams read and NW streams written, element by element
// (LB bytes per lane: 4 = the sweep's access width, 16 = float4), four
// elements per thread in flight per grid-stride step.  Read-only probes keep
// their sum live with a store no input can trigger.
template <int NR, int NW, class T>
__global__ __launch_bounds__(256) void bw_probe_kernel(const T *__restrict__ r0,
                                                       const T *__restrict__ r1,
                                                       const T *__restrict__ r2,
                                                       T *__restrict__ w0, T *__restrict__ w1,
                                                       uint64_t n, float *sink) {
  constexpr int U = 4;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  T acc{};
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += U * stride) {
    T x[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint64_t e = t + q * stride;
      x[q] = T{};
      if (e < n) {
        if constexpr (NR >= 1) x[q] = r0[e];
        if constexpr (NR >= 2) x[q] = x[q] + r1[e];
        if constexpr (NR >= 3) x[q] = x[q] + r2[e];
      }
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint64_t e = t + q * stride;
      if (e < n) {
        if constexpr (NW >= 1) w0[e] = x[q];
        if constexpr (NW >= 2) w1[e] = x[q];
        if constexpr (NW == 0) acc = acc + x[q];
      }
    }
  }
  if constexpr (NW == 0) {
    float a;
    if constexpr (sizeof(T) == 4)
      a = acc;
    else
      a = acc.x + acc.y + acc.z + acc.w;
    if (a == -1.2345e38f) sink[0] = a;
  }
}

template <class T>
hipError_t launch_probe_t(int nr, int nw, const void *const *r, void *const *w, uint64_t n,
                          float *sink, unsigned grid, hipStream_t st) {
  const T *r0 = (const T *)r[0], *r1 = (const T *)r[1], *r2 = (const T *)r[2];
  T *w0 = (T *)w[0], *w1 = (T *)w[1];
#define STH_PROBE(NR, NW)                                                                    if (nr == NR && nw == NW) {                                                                  hipLaunchKernelGGL((bw_probe_kernel<NR, NW, T>), dim3(grid), dim3(256), 0, st, r0, r1, r2,                        w0, w1, n, sink);                                                       return hipGetLastError();                                                                }
  STH_PROBE(1, 0) STH_PROBE(2, 0) STH_PROBE(3, 0) STH_PROBE(0, 1) STH_PROBE(0, 2)
  STH_PROBE(1, 1) STH_PROBE(2, 1) STH_PROBE(3, 1) STH_PROBE(1, 2) STH_PROBE(2, 2) STH_PROBE(3, 2)
#undef STH_PROBE
  return hipErrorInvalidValue;
}

hipError_t launch_bw_probe(int nr, int nw, int lane_bytes, const void *const *r, void *const *w,
                           uint64_t bytes, float *sink, unsigned grid, hipStream_t st) {
  if (lane_bytes == 4)
    return launch_probe_t<float>(nr, nw, r, w, bytes / 4, sink, grid, st);
  if (lane_bytes == 16)
    return launch_probe_t<float4>(nr, nw, r, w, bytes / 16, sink, grid, st);
  return hipErrorInvalidValue;
}

hipError_t launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st) {
  const uint64_t n4 = bytes / 16;
  hipLaunchKernelGGL(stream_copy_kernel, dim3(8192), dim3(256), 0, st, (const float4 *)src,
                     (float4 *)dst, n4);
  return hipGetLastError();
}

hipError_t launch_fill(const FieldDesc &d, const FillSpec &f, void *dst, int fmt, int e, int pass,
                       uint32_t *maxbits, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3(kReduceBlocks), dim3(256), 0, st, d, f, dst, fmt, e, pass,
                     maxbits);
  return hipGetLastError();
}

hipError_t launch_check(const FieldDesc &d, const void *v, int fmt, int e, float expected,
                        unsigned long long *errors, hipStream_t st) {
  hipLaunchKernelGGL(check_kernel, dim3(kReduceBlocks), dim3(256), 0, st, d, v, fmt, e, expected,
                     errors);
  return hipGetLastError();
}

hipError_t launch_norms(const FieldDesc &d, const void *x, int fmt, int e, double *partial,
                        uint32_t *maxbits, double *out, hipStream_t st) {
  hipLaunchKernelGGL(norms_kernel, dim3(kReduceBlocks), dim3(256), 0, st, d, x, fmt, e, partial,
                     maxbits);
  hipLaunchKernelGGL(norms_finish, dim3(1), dim3(64), 0, st, (const double *)partial,
                     kReduceBlocks, out);
  return hipGetLastError();
}

hipError_t launch_energy(const FieldDesc &d, const void *u, const void *v, int fmt, int eu,
                         int ev, const EnergySpec &sp, double *partial, double *out,
                         hipStream_t st) {
  hipLaunchKernelGGL(energy_kernel, dim3(kReduceBlocks), dim3(256), 0, st, d, u, v, fmt, eu, ev,
                     sp, partial);
  hipLaunchKernelGGL(norms_finish, dim3(1), dim3(64), 0, st, (const double *)partial,
                     kReduceBlocks, out);
  return hipGetLastError();
}

hipError_t launch_compare(const FieldDesc &dx, const void *x, int fx, int ex,
                          const FieldDesc &dr, const void *r, int fr, int er, double *pdiff,
                          double *pref, uint32_t *maxdiff, uint32_t *maxref, double *sums,
                          hipStream_t st) {
  hipLaunchKernelGGL(compare_kernel, dim3(kReduceBlocks), dim3(256), 0, st, dx, x, fx, ex, dr, r,
                     fr, er, pdiff, pref, maxdiff, maxref);
  hipLaunchKernelGGL(norms_finish, dim3(1), dim3(64), 0, st, (const double *)pdiff,
                     kReduceBlocks, sums);
  hipLaunchKernelGGL(norms_finish, dim3(1), dim3(64), 0, st, (const double *)pref,
                     kReduceBlocks, sums + 1);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t st) {
  hipLaunchKernelGGL(zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace sth

// sweep_kernel.h -- the 2.5D cross-stencil sweep for gfx950.
//
// One workgroup owns a (j,k) tile of TJ x 64 points and streams it along i
// (the reference's slowest axis x, grid.cpp:20-22), replacing the OpenMP
// loop nest of iterate_3d (iterators.cpp:137-156) and the CPU cache blocking
// of iterate_{2d,3d}_jk_blocked / Hilbert / checkerboard (iterators.cpp:
// 159-289), which exist only to reuse u from cache.
//
//  * k (unit stride): one wave covers 64 consecutive k of a row, so every
//    global access of a wave is one 256-byte (FP32) row segment.
//  * j: each wave holds MJ consecutive rows in registers; NW waves stack
//    to TJ = NW*MJ rows.  j-neighbours inside a wave's rows come from
//    registers, the others from LDS.
//  * i: a register queue of 2R+1 planes per point; each step loads only the
//    leading plane i+R of u (and s, v at plane i), one step ahead.
//  * LDS: the centre plane of the tile plus an R-wide j/k halo, double
//    buffered, one barrier per plane.
//
// Per point the arithmetic follows the reference order exactly (REF) or the
// documented LEAPFROG order; see sweep_point_* below.  Compiled with
// -ffp-contract=off: every product that is not an explicit fmaf is rounded.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "sth_common.h"

namespace sth {

enum { MODE_REF = 0, MODE_LEAP = 1 };
enum { FMT_F32 = 0, FMT_F16 = 1, FMT_BF16 = 2, FMT_F64 = 3 };

// Per-step scale state of the 16-bit storage form (a ring of 3 slots: step n
// reads slot n%3, publishes slot (n+1)%3 and clears the max of slot (n+2)%3).
// A buffer holds values q * 2^exp; maxbits is max |q| (FP32 bits).
struct Q16Slot {
  uint32_t maxbits[3];  // max |stored value| of buffers a, b, s as FP32 bits
  int32_t exp[3];       // power-of-two exponents of buffers a, b, s
};

struct SweepArgs {
  const void *u;  // field origins: element (0,0,0) of the interior
  void *v;
  const void *s;
  int64_t pitch, plane;  // elements
  int xs, ys, zs;
  int ntk, ntj, chunks, chunk_len;
  int remap;
  float coef0, mA;      // REF: coef0 and -A
  int ref_extra;        // REF: updates of v after the first (u, s are read-only)
  float c0, c1, c2n;    // LEAPFROG coefficients (c2n: this step's source coefficient)
  float w[12];          // per-term weight, evaluation order (LEAPFROG: 2*w)
  double wd[12];        // FP64 storage: the same weights as FP64 numbers
  Q16Slot *qs;          // 16-bit storage: the 3-slot ring
  int slot;             // slot read by this step
  int ubuf, vbuf;       // ring indices (0 = a, 1 = b) of u and v
  double gain;          // sum over terms of 2*|w| (LEAPFROG weights)
  // Persistent launch (tickets != null): the tiles, in item order, form
  // bands of `band` tiles; band q belongs to XCD q % kXcds (blocks are
  // dispatched to XCDs round robin, so block b runs on XCD b % kXcds).  The
  // blocks of XCD x draw its bands' tiles from counter x (tickets[16 x])
  // until they draw one past them; a launch consumes exactly (tiles of XCD x
  // + blocks on XCD x) tickets of counter x, so the launches of a sequence
  // start at tick_base[x].
  unsigned *tickets;
  unsigned tick_base[kXcds];
  int items, band;
};

// Calls f(integral_constant<int, 0>) .. f(integral_constant<int, N-1>).
template <class Fn, int... K>
__device__ __forceinline__ void static_for_seq(Fn &&f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <int FMT>
struct Codec;
// Storage codecs.  16-bit values are the stored numbers q themselves (exact
// in FP32); a field's power-of-two scale 2^e is folded into the update's
// coefficients instead of into every load and store.
template <>
struct Codec<FMT_F32> {
  using T = float;
  using A = float;
  __device__ static inline float dec(T x) { return x; }
  __device__ static inline T enc(float x) { return x; }
};
template <>
struct Codec<FMT_F16> {
  using T = uint16_t;
  using A = float;
  __device__ static inline float dec(T q) { return (float)__builtin_bit_cast(_Float16, q); }
  __device__ static inline T enc(float x) {
    // The update rounds to FP32 first, then to binary16 (two roundings, as
    // the oracle does).  The empty asm keeps the compiler from fusing the
    // last fma with the conversion into v_fma_mix*_f16 (one rounding).
    __asm__ volatile("" : "+v"(x));
    return __builtin_bit_cast(uint16_t, (_Float16)x);  // v_cvt_f16_f32, RNE
  }
};
template <>
struct Codec<FMT_BF16> {
  using T = uint16_t;
  using A = float;
  __device__ static inline float dec(T q) { return __uint_as_float((uint32_t)q << 16); }
  __device__ static inline T enc(float x) {
    uint32_t b = __float_as_uint(x);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (T)((b >> 16) | 0x40u);
    b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
    return (T)(b >> 16);
  }
};

template <>
struct Codec<FMT_F64> {
  using T = double;
  using A = double;
  __device__ static inline double dec(T x) { return x; }
  __device__ static inline T enc(double x) { return x; }
};

__device__ inline float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Smallest e with B*2^-e < 2^15 (frexp exponent - 15); B <= 0 -> 0.
__host__ __device__ inline int scale_exp(double B) {
  if (!(B > 0.0)) return 0;
  int E;
  frexp(B, &E);
  return E - 15;
}

// Exponent of the field a LEAPFROG step writes: from the bound
// |out| <= (|c0| + |c1| gain) M_u + M_v + |c2n| M_s, with M = max|q| * 2^e.
__host__ __device__ inline int q16_out_exp(const Q16Slot &cur, int ubuf, int vbuf, float c0,
                                           float c1, float c2n, double gain) {
  auto M = [&](int b) {
    float m;
    memcpy(&m, &cur.maxbits[b], 4);
    return ldexp((double)m, cur.exp[b]);
  };
  const double gmul = fabs((double)c0) + fabs((double)c1) * gain;
  return scale_exp(gmul * M(ubuf) + M(vbuf) + fabs((double)c2n) * M(2));
}

// Tiles owned by XCD x when `items` tiles form bands of `band`, band q
// going to XCD q % kXcds (only the last band may be partial).
__host__ __device__ inline int xcd_band_items(int items, int band, int x) {
  const int nb = (items + band - 1) / band;
  if (x >= nb) return 0;
  const int owned = (nb - 1 - x) / kXcds + 1;
  const bool has_last = (nb - 1) % kXcds == x;
  return has_last ? (owned - 1) * band + (items - (nb - 1) * band) : owned * band;
}

// Blocks are dispatched round-robin over the 8 XCDs; this bijection gives
// each XCD a contiguous run of tiles so the j-halo rows two neighbouring tiles
// share are served from that XCD's L2.
__device__ inline int xcd_remap(int bid, int nb) {
  const int per = (nb + kXcds - 1) / kXcds;
  const int tall = nb - (per - 1) * kXcds;
  const int x = bid % kXcds, l = bid / kXcds;
  return x < tall ? x * per + l : tall * per + (x - tall) * (per - 1) + l;
}

// V consecutive storage elements of one lane (one aligned vector load).
template <class ST, int V>
struct VecT;
template <class ST>
struct VecT<ST, 1> {
  using type = ST;
  __device__ static inline ST get(const type &x, int) { return x; }
};
template <class ST>
struct VecT<ST, 2> {
  using type = ST __attribute__((ext_vector_type(2)));
  __device__ static inline ST get(const type &x, int c) { return c ? x.y : x.x; }
};

// Buffer loads and stores: a plane's base lives in a scalar buffer resource
// (range 4 GiB, checked against the 32-bit lane offset), so per access the
// lane supplies only its 32-bit byte offset -- no 64-bit address arithmetic
// in vector registers.  AUX = 2 marks the access nontemporal.
using BufRes = __amdgpu_buffer_rsrc_t;
template <class Ptr>
__device__ __forceinline__ BufRes buf_res(Ptr base) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)base, (short)0, (int)0xffffffff,
                                           0x00020000);
}
// A resource whose range is `bytes`: a load at an offset at or past it
// returns 0 without a memory access.
template <class Ptr>
__device__ __forceinline__ BufRes buf_res_n(Ptr base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)base, (short)0, (int)bytes,
                                           0x00020000);
}
template <class T, int AUX>
__device__ __forceinline__ T buf_load(BufRes r, uint32_t off) {
  if constexpr (sizeof(T) == 2) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, AUX));
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
  } else if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
  } else {
    static_assert(sizeof(T) == 16, "2, 4, 8 or 16 bytes");
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX));
  }
}
template <class T, int AUX>
__device__ __forceinline__ void buf_store(T x, BufRes r, uint32_t off) {
  using I2 = int __attribute__((ext_vector_type(2)));
  using I4 = int __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(T) == 2) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(short, x), r, off, 0, AUX);
  } else if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, x), r, off, 0, AUX);
  } else if constexpr (sizeof(T) == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(I2, x), r, off, 0, AUX);
  } else {
    static_assert(sizeof(T) == 16, "2, 4, 8 or 16 bytes");
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(I4, x), r, off, 0, AUX);
  }
}

// One register-queue entry: a lane's V values of one plane.  With V = 2 the
// pair is one 2-element vector, so shifting the queue moves 64-bit registers.
template <class A, int V>
struct QCell {
  A e[1];
  __device__ A get(int) const { return e[0]; }
  __device__ void set(int, A x) { e[0] = x; }
};
template <class A>
struct QCell<A, 2> {
  A __attribute__((ext_vector_type(2))) e;
  __device__ A get(int c) const { return c ? e.y : e.x; }
  __device__ void set(int c, A x) {
    if (c)
      e.y = x;
    else
      e.x = x;
  }
};

// LDS plane of a tile: the centre plane plus an R-wide halo.  Column CO
// holds the tile's first k; with two k points per lane CO is even, so every
// lane's pair sits on an 8-byte boundary and LDS traffic moves as 2-element
// vectors (no 2-way bank conflicts from lanes 8 bytes apart).
template <int NPTS, int MJ, int NW, int V, int F = 1>
struct TileLds {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int CO = V == 2 ? 2 * ((R + 1) / 2) : R;
  static constexpr int LW = 64 / F * V + 2 * CO;
  static constexpr int SIZE = (NW * MJ * F + 2 * R) * LW;
};

// Tile engine shared by the single-grid and the batched kernels: a
// (TJ = NW*MJ*F) x (TK = 64/F*V) tile of one grid at (j0, k0), planes
// [ib, ie).  A wave's 64 lanes form F rows of LL = 64/F lanes; lane l of wave
// w holds k = k0 + V*(l % LL) .. + V-1 of rows j0 + (w*MJ + m)*F + l / LL,
// m < MJ.  F = 1 is the regular tile (one row per lane group of 64); F > 1
// folds a narrow tile (the last k-tile of a row whose length is not a
// multiple of TK) onto whole waves.
template <int NPTS, int MODE, int FMT, int MJ, int NW, int V, int F = 1>
__device__ __forceinline__ void sweep_tile(
    const SweepArgs &a, int k0, int j0, int ib, int ie,
    typename Codec<FMT>::A (*lds)[TileLds<NPTS, MJ, NW, V, F>::SIZE],
    float *red) {
  using Tm = Terms<NPTS>;
  constexpr int R = Tm::R, T = Tm::T;
  using CD = Codec<FMT>;
  using ST = typename CD::T;
  using AT = typename CD::A;  // arithmetic type: FP32, or FP64 for FP64 storage
  using VV = VecT<ST, V>;
  using VT = typename VV::type;
  // global (address space 1) views: plane bases are byte pointers
  using GB = __attribute__((address_space(1))) char;
  using GPtr = GB *;
  constexpr int LL = 64 / F;  // lanes per row
  constexpr int TJ = NW * MJ * F, TK = LL * V;
  constexpr int CO = TileLds<NPTS, MJ, NW, V, F>::CO, LW = TileLds<NPTS, MJ, NW, V, F>::LW;
  constexpr int HJ = (2 * R + NW * F - 1) / (NW * F);  // j-halo row loads per wave
  static_assert(MJ * F * 2 * R <= 64, "k-halo needs at most one entry per thread");
  static_assert(V == 1 || V == 2, "1 or 2 k points per lane");
  static_assert(F == 1 || V == 1, "folded tiles have one k point per lane");
  static_assert(V - 1 <= R, "a lane's vector may reach past the last halo cell by V-1");

  const int lane = threadIdx.x & 63;
  const int lc = lane % LL, lf = lane / LL;  // column lane, row in the fold
  const int w = threadIdx.x >> 6;
  const int tid = threadIdx.x;
  const int64_t P = a.plane, Pj = a.pitch;

  // LEAPFROG coefficients.  16-bit storage works on the stored numbers q
  // (value = q 2^e): with the output exponent eo the update becomes
  //   q_out = c2n 2^(es-eo) q_s + c1 2^(eu-eo) L(q_u) + c0 2^(eu-eo) q_u - 2^(ev-eo) q_v,
  // the same FP32 operations on values scaled by exact powers of two.
  AT kc0 = (AT)a.c0, kc1 = (AT)a.c1, kc2 = (AT)a.c2n, kv = (AT)1;
  if constexpr (FMT == FMT_F16 || FMT == FMT_BF16) {
    const Q16Slot &cur = a.qs[a.slot];
    const int eu = cur.exp[a.ubuf], ev = cur.exp[a.vbuf], es = cur.exp[2];
    const int eo = q16_out_exp(cur, a.ubuf, a.vbuf, a.c0, a.c1, a.c2n, a.gain);
    kc0 = ldexpf(a.c0, eu - eo);
    kc1 = ldexpf(a.c1, eu - eo);
    kc2 = ldexpf(a.c2n, es - eo);
    kv = ldexpf(1.0f, ev - eo);
  }

  // Tile base: element (i = 0, j0 - R, k0 - R); all in-tile offsets >= 0.
  const ST *Ub = (const ST *)a.u + ((int64_t)(j0 - R) * Pj + (k0 - R));
  ST *Vb = (ST *)a.v + ((int64_t)(j0 - R) * Pj + (k0 - R));
  const ST *Sb = (const ST *)a.s + ((int64_t)(j0 - R) * Pj + (k0 - R));

  // Clamp reads to the halo-padded extent: lanes/rows past the interior still
  // load the halo values their neighbours need, and never leave the field.
  // A lane's vector starts at or before the last readable column, so it ends
  // at most V-1 elements past it: inside the row padding (pitch >= zs + 2R,
  // and the field has 256 B of slack after its last row).
  const int jmax = a.ys + R - 1 - j0;  // last readable tile row
  const int kmax = a.zs + R - 1 - k0;  // last readable tile column
  const int kc = min(lc * V, (kmax / V) * V);
  uint32_t off[MJ];
  bool valid[MJ];    // the lane's first point is inside the interior
  bool valid2[MJ];   // ... and its last one too
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int row = (w * MJ + m) * F + lf;
    off[m] = (uint32_t)(((int64_t)(min(row, jmax) + R) * Pj + (kc + R)) * sizeof(ST));
    valid[m] = (j0 + row < a.ys) && (k0 + lc * V < a.zs);
    valid2[m] = (j0 + row < a.ys) && (k0 + lc * V + V - 1 < a.zs);
  }
  uint32_t hoff[HJ];
  int hrow[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) {
    const int hh = (w + NW * h) * F + lf;  // halo row id 0..2R-1
    const int trow = hh < R ? hh - R : TJ + hh - R;
    hrow[h] = hh < 2 * R ? (hh < R ? hh : TJ + hh) : -1;  // LDS row, -1: none
    hoff[h] = (uint32_t)(((int64_t)(min(trow, jmax) + R) * Pj + (kc + R)) * sizeof(ST));
  }
  const bool has_kh = tid < TJ * 2 * R;
  uint32_t khoff = 0;
  int khpos = 0;
  if (has_kh) {
    const int row = tid / (2 * R), c = tid % (2 * R);
    const int tcol = c < R ? c - R : TK + c - R;
    khoff = (uint32_t)(((int64_t)(min(row, jmax) + R) * Pj + (min(tcol, kmax) + R)) * sizeof(ST));
    khpos = (row + R) * LW + (c < R ? CO - R + c : CO + TK + c - R);
  }

  // In-tile offsets are 32-bit byte offsets from a workgroup-uniform plane
  // base, so loads and stores take the SGPR-base + VGPR-offset form.  The
  // base goes through readfirstlane each plane: it stays a scalar, and the
  // compiler does not turn base + offset into per-lane 64-bit pointers
  // stepped by the plane stride.
  auto uni = [](const void *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (GPtr)(((uint64_t)hi << 32) | lo);
  };
  auto vload = [](BufRes p, uint32_t ob) { return buf_load<VT, 0>(p, ob); };
  auto vload_nt = [](BufRes p, uint32_t ob) { return buf_load<VT, 2>(p, ob); };
  auto sload = [](BufRes p, uint32_t ob) { return buf_load<ST, 0>(p, ob); };

  // Per-term weight in the arithmetic type: FP64 storage takes its own FP64
  // weights (the FP32 constants promoted, or the reference's -DUSE_DOUBLE
  // constants), FP32 and 16-bit storage the FP32 ones.
  auto wt = [&](int t) -> AT {
    if constexpr (FMT == FMT_F64)
      return a.wd[t];
    else
      return a.w[t];
  };

  // Register queue: q[m][t].get(c) = u at plane i - R + t.
  QCell<AT, V> q[MJ][2 * R + 1];
#pragma unroll
  for (int t = 0; t < 2 * R; t++) {
    const BufRes Up = buf_res(uni(Ub + (int64_t)(ib - R + t) * P));
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      const VT x = vload(Up, off[m]);
#pragma unroll
      for (int c = 0; c < V; c++) q[m][t].set(c, CD::dec(VV::get(x, c)));
    }
  }
  // Every u value of a plane -- the tile's own rows, its j-halo rows and its
  // k-halo columns -- is loaded in the same step, when the plane enters as
  // the leading plane i+R; the halo values then wait R steps in register
  // queues (hq, kq: planes i..i+R).  Neighbouring workgroups run in step, so
  // each line is fetched once and the other tiles' reads of it hit in L2
  // (loading the halo when the plane becomes the centre, R steps later, finds
  // it evicted).
  VT hq[HJ][R + 1];
  ST kq[R + 1];
#pragma unroll
  for (int t = 0; t < R; t++) {
    const BufRes Up = buf_res(uni(Ub + (int64_t)(ib + t) * P));
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][t] = hrow[h] >= 0 ? vload(Up, hoff[h]) : VT(0);
    kq[t] = has_kh ? sload(Up, khoff) : ST(0);
  }
  // Loads run two planes ahead: two sets of staging registers, used in
  // turn by the even and odd steps of a plane pair, so a set is refilled
  // only after it has been consumed (no copy of registers still in flight).
  struct Staged {
    VT u[MJ], s[MJ], v[MJ], hj[HJ];
    ST hk;
  };
  Staged PA, PB;
  // s and v are touched once per sweep: nontemporal loads and stores keep
  // them from displacing u's lines, which neighbouring tiles re-read.
  auto prefetch = [&](int i, Staged &n) __attribute__((always_inline)) {
    const BufRes Up = buf_res(uni(Ub + (int64_t)(i + R) * P));
    const BufRes Sp = buf_res(uni(Sb + (int64_t)i * P));
    const BufRes Vp = buf_res(uni(Vb + (int64_t)i * P));
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      // masked lanes read their clamped (in-bounds) cells: no branches
      n.u[m] = vload(Up, off[m]);
      n.s[m] = vload_nt(Sp, off[m]);
      n.v[m] = vload_nt(Vp, off[m]);
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) n.hj[h] = hrow[h] >= 0 ? vload(Up, hoff[h]) : VT(0);
    n.hk = has_kh ? sload(Up, khoff) : ST(0);
  };
  prefetch(ib, PA);
  if (ib + 1 < ie) prefetch(ib + 1, PB);

  float mx = 0.0f;
  auto step = [&](int i, Staged &n) __attribute__((always_inline)) {
    AT *L = lds[(i - ib) & 1];
    VT cs[MJ], cv[MJ];
#pragma unroll
    for (int m = 0; m < MJ; m++) {
#pragma unroll
      for (int c = 0; c < V; c++) q[m][2 * R].set(c, CD::dec(VV::get(n.u[m], c)));
      cs[m] = n.s[m];
      cv[m] = n.v[m];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R] = n.hj[h];
    kq[R] = n.hk;
    if (i + 2 < ie) prefetch(i + 2, n);

#pragma unroll
    for (int m = 0; m < MJ; m++)
#pragma unroll
      for (int c = 0; c < V; c++)
        L[((w * MJ + m) * F + lf + R) * LW + lc * V + c + CO] = q[m][R].get(c);
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0)
#pragma unroll
        for (int c = 0; c < V; c++)
          L[hrow[h] * LW + lc * V + c + CO] = CD::dec(VV::get(hq[h][0], c));
    if (has_kh) L[khpos] = CD::dec(kq[0]);
    __syncthreads();

    const BufRes Vp = buf_res(uni(Vb + (int64_t)i * P));
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      const int lr = (w * MJ + m) * F + lf + R, col = lc * V + CO;
      // Two k points per lane: the k- and j-neighbours come from LDS as
      // aligned 2-element vectors, read once for both points.
      // kn[t] = column col - 2 KH + t; jl[r] / jh[r] = rows lr -/+ r.
      using AT2 = AT __attribute__((ext_vector_type(2)));
      constexpr int KH = (R + 1) / 2;
      AT kn[V == 2 ? 4 * KH + 2 : 1];
      AT jl[R + 1][V], jh[R + 1][V];
      if constexpr (V == 2) {
#pragma unroll
        for (int p = -KH; p <= KH; p++) {
          if (p == 0) {
            kn[2 * KH] = q[m][R].get(0);
            kn[2 * KH + 1] = q[m][R].get(1);
          } else {
            const AT2 x = *(const AT2 *)&L[lr * LW + col + 2 * p];
            kn[2 * (p + KH)] = x.x;
            kn[2 * (p + KH) + 1] = x.y;
          }
        }
#pragma unroll
        for (int r = 1; r <= R; r++) {
          if (m - r < 0) {
            const AT2 x = *(const AT2 *)&L[(lr - r) * LW + col];
            jl[r][0] = x.x;
            jl[r][1] = x.y;
          }
          if (m + r >= MJ) {
            const AT2 x = *(const AT2 *)&L[(lr + r) * LW + col];
            jh[r][0] = x.x;
            jh[r][1] = x.y;
          }
        }
      }
      AT res[V];
#pragma unroll
      for (int c = 0; c < V; c++) {
        auto pair = [&](int ax, int r) -> AT {
          if (ax == 0) return q[m][R - r].get(c) + q[m][R + r].get(c);
          if constexpr (V == 2) {
            if (ax == 2) return kn[2 * KH + c - r] + kn[2 * KH + c + r];
            const AT lo = (m - r >= 0) ? q[(m - r) >= 0 ? m - r : 0][R].get(c) : jl[r][c];
            const AT hi = (m + r < MJ) ? q[(m + r) < MJ ? m + r : 0][R].get(c) : jh[r][c];
            return lo + hi;
          }
          if (ax == 2) {
            // k-neighbours inside the lane's own vector come from registers
            const AT lo = (c - r >= 0) ? q[m][R].get((c - r) >= 0 ? c - r : 0)
                                       : L[lr * LW + col + c - r];
            const AT hi = (c + r < V) ? q[m][R].get((c + r) < V ? c + r : 0)
                                      : L[lr * LW + col + c + r];
            return lo + hi;
          }
          // rows lr -/+ r are this lane's own rows m -/+ r/F when F divides r
          const bool lreg = r % F == 0 && m - r / F >= 0;
          const bool hreg = r % F == 0 && m + r / F < MJ;
          const AT lo = lreg ? q[lreg ? m - r / F : 0][R].get(c) : L[(lr - r) * LW + col + c];
          const AT hi = hreg ? q[hreg ? m + r / F : 0][R].get(c) : L[(lr + r) * LW + col + c];
          return lo + hi;
        };
        const AT u0 = q[m][R].get(c);
        const AT s0 = CD::dec(VV::get(cs[m], c));
        const AT v0 = CD::dec(VV::get(cv[m], c));
        AT out;
        if constexpr (MODE == MODE_REF) {
          // stencil_N_ref_1: p = coef0*u0 + sum c*pair (the first product
          // rounded, the rest fused), then v0 - A*u0 + s0*p.
          AT p = fmaT((AT)a.coef0, u0, wt(0) * pair(Tm::axis[0], Tm::r[0]));
#pragma unroll
          for (int t = 1; t < T; t++) p = fmaT(wt(t), pair(Tm::axis[t], Tm::r[t]), p);
          out = fmaT(s0, p, fmaT((AT)a.mA, u0, v0));
          for (int r = 0; r < a.ref_extra; r++) out = fmaT(s0, p, fmaT((AT)a.mA, u0, out));
        } else {
          AT Lp = wt(0) * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
          for (int t = 1; t < T; t++) Lp = fmaT(wt(t), pair(Tm::axis[t], Tm::r[t]), Lp);
          if constexpr (FMT == FMT_F16 || FMT == FMT_BF16)
            out = fmaT(kc2, s0, fmaT(kc1, Lp, fmaT(kc0, u0, -(v0 * kv))));
          else
            out = fmaT(kc2, s0, fmaT(kc1, Lp, fmaT(kc0, u0, -v0)));
        }
        res[c] = out;
      }
      ST qo[V];
#pragma unroll
      for (int c = 0; c < V; c++) qo[c] = CD::enc(res[c]);
      if constexpr (V == 1) {
        if (valid[m]) buf_store<ST, 2>(qo[0], Vp, off[m]);
      } else {
        if (valid2[m]) {
          VT o;
          o.x = qo[0];
          o.y = qo[1];
          buf_store<VT, 2>(o, Vp, off[m]);
        } else if (valid[m]) {
          buf_store<ST, 2>(qo[0], Vp, off[m]);
        }
      }
      if constexpr (FMT == FMT_F16 || FMT == FMT_BF16) {
        // max |stored value| = round(max |FP32 result|): rounding to the
        // 16-bit format is odd and monotone, so it commutes with max |.|;
        // the rounding is applied once, after the reduction
        mx = fmaxf(mx, valid[m] ? fabsf((float)res[0]) : 0.0f);
        if constexpr (V == 2) mx = fmaxf(mx, valid2[m] ? fabsf((float)res[1]) : 0.0f);
      }
    }
#pragma unroll
    for (int m = 0; m < MJ; m++)
#pragma unroll
      for (int t = 0; t < 2 * R; t++) q[m][t] = q[m][t + 1];
#pragma unroll
    for (int t = 0; t < R; t++) {
#pragma unroll
      for (int h = 0; h < HJ; h++) hq[h][t] = hq[h][t + 1];
      kq[t] = kq[t + 1];
    }
  };
  for (int i = ib; i < ie; i += 2) {
    step(i, PA);
    if (i + 1 < ie) step(i + 1, PB);
  }

  if constexpr (FMT == FMT_F16 || FMT == FMT_BF16) {
    mx = CD::dec(CD::enc(mx));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[w] = mx;
    __syncthreads();
    if (tid == 0) {
      float b = red[0];
#pragma unroll
      for (int t = 1; t < NW; t++) b = fmaxf(b, red[t]);
      atomicMax(&a.qs[(a.slot + 1) % 3].maxbits[a.vbuf], __float_as_uint(b));
    }
  }
}

// Publishes the next slot of the 16-bit scale ring (block 0, thread 0).
template <int FMT>
__device__ inline void q16_publish(const SweepArgs &a) {
  if constexpr (FMT == FMT_F16 || FMT == FMT_BF16) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const Q16Slot &cur = a.qs[a.slot];
      Q16Slot &nxt = a.qs[(a.slot + 1) % 3];
      Q16Slot &clr = a.qs[(a.slot + 2) % 3];
      nxt.exp[a.vbuf] = q16_out_exp(cur, a.ubuf, a.vbuf, a.c0, a.c1, a.c2n, a.gain);
      nxt.exp[a.ubuf] = cur.exp[a.ubuf];
      nxt.exp[2] = cur.exp[2];
      nxt.maxbits[a.ubuf] = cur.maxbits[a.ubuf];
      nxt.maxbits[2] = cur.maxbits[2];
      clr.maxbits[0] = clr.maxbits[1] = clr.maxbits[2] = 0u;
    }
  }
}

// Runs tile(item) for the work items of one launch: one item per workgroup
// (through the XCD-contiguous remap when a.remap), or, when a.tickets is
// set, persistent workgroups drawing items from the ticket counters; within
// an XCD's bands the tiles go in item order (k fastest, then j, then the
// i-chunk).  One call site of the tile body for both launch forms keeps the
// kernel's register budget that of one tile.  sh_item: a __shared__ int.
template <class TileFn>
__device__ __forceinline__ void for_items(const SweepArgs &a, int &sh_item, TileFn &&tile) {
  const bool persistent = a.tickets != nullptr;
  const bool one = a.band >= a.items;  // a single queue (counter 0)
  const int x = one ? 0 : (int)(blockIdx.x % kXcds);
  const int n = persistent ? (one ? a.items : xcd_band_items(a.items, a.band, x)) : 0;
  unsigned *cnt = persistent ? a.tickets + 16 * x : nullptr;
  for (;;) {
    int item;
    if (persistent) {
      if (threadIdx.x == 0) sh_item = (int)(atomicAdd(cnt, 1u) - a.tick_base[x]);
      __syncthreads();
      const int t = sh_item;
      __syncthreads();
      if (t < 0 || t >= n) break;
      item = one ? t : (x + kXcds * (t / a.band)) * a.band + t % a.band;
    } else {
      item = a.remap ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    }
    tile(item);
    if (!persistent) break;
    __syncthreads();  // LDS is reused by the next tile
  }
}

// EF > 0: the last k-tile of every row is the narrow remainder of the row
// (at most 64/EF points) and runs folded (EF rows per wave, one k point per
// lane) instead of as a full, mostly masked tile.
template <int NPTS, int MODE, int FMT, int MJ, int NW, int V, int EF = 0>
__global__ __launch_bounds__(64 * NW) void sweep_kernel(const SweepArgs a) {
  using AT = typename Codec<FMT>::A;
  __shared__ __attribute__((aligned(16))) AT lds[2][TileLds<NPTS, MJ, NW, V>::SIZE];
  __shared__ float red[NW];
  __shared__ int sh_item;
  q16_publish<FMT>(a);
  // One tile: item = tk + ntk * (tj + ntj * chunk).
  auto tile = [&](int item) __attribute__((always_inline)) {
    const int tk = item % a.ntk;
    const int t = item / a.ntk;
    const int tj = t % a.ntj;
    const int ch = t / a.ntj;
    const int ib = ch * a.chunk_len;
    const int ie = min(a.xs, ib + a.chunk_len);
    if (ib >= ie) return;
    const int k0 = tk * 64 * V, j0 = tj * NW * MJ;
    if constexpr (EF > 0) {
      static_assert(MJ % EF == 0, "a folded tile keeps the tile's rows");
      using E = TileLds<NPTS, MJ / EF, NW, 1, EF>;
      static_assert(E::SIZE <= TileLds<NPTS, MJ, NW, V>::SIZE, "");
      if (tk == a.ntk - 1) {
        sweep_tile<NPTS, MODE, FMT, MJ / EF, NW, 1, EF>(a, k0, j0, ib, ie,
                                                         (AT(*)[E::SIZE])&lds[0][0], red);
        return;
      }
    }
    sweep_tile<NPTS, MODE, FMT, MJ, NW, V>(a, k0, j0, ib, ie, lds, red);
  };
  for_items(a, sh_item, tile);
}

}  // namespace sth

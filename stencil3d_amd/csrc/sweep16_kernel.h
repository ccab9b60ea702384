// sweep16_kernel.h -- the one-step LEAPFROG sweep on FP16 storage.
//
// The schedule is sweep_kernel's: a workgroup streams a (j, k) tile along i
// with a register queue of 2R+1 planes of u, the centre plane plus an R-wide
// halo in LDS (double buffered, one barrier per plane), and persistent
// workgroups draw tiles from ticket counters.  What differs is that the data
// stays in its stored binary16 form throughout:
//
//  * a lane holds two consecutive k points as one packed 32-bit word -- in
//    the register queue, the staging registers and LDS -- so a tile row is
//    128 points (256 B, one 32-bit load per lane) and the queue costs one
//    register per plane and row;
//  * the arithmetic reads halves where they lie: v_fma_mix_f32 selects the
//    low or high half of a word and widens it to FP32 exactly inside the
//    instruction, so a neighbour pair u[c-r] + u[c+r] is one instruction
//    (x * 1.0 + y, one rounding: the FP32 sum of the two widened values),
//    and the centre, v and s terms take their halves inside the fused
//    multiply-adds; no instruction is spent on converting an input;
//  * a lane's two results round to binary16 in one instruction
//    (v_cvt_pk_f16_f32, round to nearest even) and leave as one 32-bit store;
//  * loads run D = 3 planes ahead, straight into ring slots (u, s, v and the
//    halo cells), and the loop is unrolled over one turn of the rings, so
//    every ring position is a compile-time constant: the rings rotate by
//    renaming, with no register moves and no copy of a staged load; every
//    load is unconditional (plane pointers stop at the chunk's last plane),
//    so the compiler never waits on a load that a later step does not need.
//
// Measured at 1040^3 (25 points): 31 VALU instructions per point in the
// inner loop against 48 for sweep_kernel's 16-bit path, 1.90-1.92 ms per
// step against 2.0; the time is then held by the HBM read/write mix, not by
// the instruction stream.
//
// The arithmetic is the one-step sweep's LEAPFROG expression on the stored
// numbers q (value = q 2^e) with the coefficients scaled by exact powers of
// two, in the same order and with the same roundings (FP32 operations on
// exact widenings of the stored halves): the results are bit-identical to
// sweep_kernel's FP16 path and to the oracle's emulation of it.
#pragma once

#include <utility>

#include "sweep_kernel.h"

namespace sth {

// LDS plane of a packed tile, in 32-bit words: TJ + 2R rows of LL words
// (the lanes' pairs) plus KH = ceil(R/2) words of k-halo on each side.  F
// is the fold: a wave's 64 lanes form F rows of LL = 64/F lanes, so the tile
// is TJ = NW MJ F rows by 2 LL points (F = 1: the regular tile, one row of
// 128 points per lane group).
template <int NPTS, int MJ, int NW, int F = 1>
struct Tile16 {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int KH = (R + 1) / 2;
  static constexpr int LL = 64 / F;
  static constexpr int LWW = LL + 2 * KH;
  static constexpr int TJ = NW * MJ * F;
  static constexpr int SIZE = (TJ + 2 * R) * LWW;
  static constexpr int D = 3;  // planes loaded ahead
};

// A folded tile (F > 1) loads its k-halo as one 32-bit word per thread.
__host__ __device__ constexpr bool fold16_fits(int R, int mj, int f) {
  return f == 2 || f == 4 || f == 8 ? mj * f * 2 * ((R + 1) / 2) <= 64 : false;
}

// Slots of the s and v rings: the smallest length of at least d+1 that
// divides the u ring's length qu.
__host__ __device__ constexpr int sv_ring(int qu, int d) {
  int s = d + 1;
  while (qu % s) s++;
  return s;
}

// FP32 sum of half H of x and half H of y (one rounding), as
// v_fma_mix_f32 x, 1.0, y: each half is widened exactly inside the
// instruction, and x * 1.0 + y rounds once, like the sum of the widened
// values (signed zeros included).
template <int H>
__device__ __forceinline__ float hadd(uint32_t x, uint32_t y) {
  float r;
  if constexpr (H == 0)
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(x), "v"(y));
  else
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(x), "v"(y));
  return r;
}

// -(half H of x * k) as one instruction: x * (-k) + (-0).  The product of a
// binary16 number and a power of two is exact in FP32, and adding -0 keeps
// it (and its sign) unchanged.
template <int H>
__device__ __forceinline__ float hnegmul(uint32_t x, float k) {
  float r;
  if constexpr (H == 0)
    asm("v_fma_mix_f32 %0, %1, -%2, neg(0) op_sel_hi:[1,0,0]" : "=v"(r) : "v"(x), "v"(k));
  else
    asm("v_fma_mix_f32 %0, %1, -%2, neg(0) op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=v"(r)
        : "v"(x), "v"(k));
  return r;
}

// Half H of x widened to FP32 (the compiler folds it into v_fma_mix_f32
// operands).
template <int H>
__device__ __forceinline__ float hwide(uint32_t x) {
  using H2 = _Float16 __attribute__((ext_vector_type(2)));
  const H2 v = __builtin_bit_cast(H2, x);
  return (float)(H ? v.y : v.x);
}

// The update's coefficients on the stored numbers (see sweep_tile).
struct Coef16 {
  float kc0, kc1, kc2, kv;
};

// One tile: lane l of wave w holds the pair at k = k0 + 2 (l % LL) of rows
// j0 + (w MJ + m) F + l / LL, m < MJ.  F > 1 folds a narrow tile (the last
// k-tile of rows whose length is not a multiple of 128) onto whole waves: a
// lane's j-neighbours then come from LDS, and its rows differ across the
// wave, so every result is stored under a per-lane mask.
template <int NPTS, int MJ, int NW, int F = 1>
__device__ __forceinline__ void sweep16_tile(const SweepArgs &a, const Coef16 &K, int k0, int j0,
                                             int ib, int ie, uint32_t *lds, float *red) {
  using Tm = Terms<NPTS>;
  using G = Tile16<NPTS, MJ, NW, F>;
  constexpr int R = Tm::R, T = Tm::T, KH = G::KH, LWW = G::LWW, LL = G::LL;
  constexpr int TJ = G::TJ, TK = 2 * LL;
  constexpr int HJ = (2 * R + NW * F - 1) / (NW * F);  // j-halo rows loaded per wave
  // Loads run D planes ahead.  u ring: the 2R+1 planes of the stencil and
  // the D in flight, rounded up to an even length (the LDS buffers
  // alternate); s / v rings: the smallest length of at least D+1 slots that
  // divides it.  One loop iteration is one turn of the rings.
  constexpr int D = Tile16<NPTS, MJ, NW>::D;
  constexpr int QU = (2 * R + D + 2) / 2 * 2;
  constexpr int SV = sv_ring(QU, D);
  constexpr int UL = QU;
  static_assert(F == 1 ? TJ * 2 * R <= 64 * NW : fold16_fits(R, MJ, F), "one k-halo cell per thread");
  using GB = __attribute__((address_space(1))) char;
  using LW32 = __attribute__((address_space(3))) uint32_t;
  using LW16 = __attribute__((address_space(3))) uint16_t;

  const int lane = threadIdx.x & 63;
  const int lc = F == 1 ? lane : lane % LL, lf = F == 1 ? 0 : lane / LL;  // pair, row in the fold
  // the wave index as a scalar: conditions on rows are wave-uniform branches
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = threadIdx.x;
  const int64_t P = a.plane, Pj = a.pitch;

  // Tile base: element (i = 0, j0 - R, k0 - R); in-tile byte offsets >= 0.
  const char *Ub = (const char *)a.u + ((int64_t)(j0 - R) * Pj + (k0 - R)) * 2;
  char *Vb = (char *)a.v + ((int64_t)(j0 - R) * Pj + (k0 - R)) * 2;
  const char *Sb = (const char *)a.s + ((int64_t)(j0 - R) * Pj + (k0 - R)) * 2;

  // Reads clamp to the halo-padded box.  A lane's pair starts at an even
  // column at or before the last readable one (kmax), so it ends at most one
  // element past it, inside the row padding (pitch >= zs + 2R, and the field
  // has 256 B of slack after its last row).  Interior k = 0 of every row is
  // 256-B aligned and k0 and the pair's column are even: every pair is one
  // aligned 32-bit word.
  const int jmax = a.ys + R - 1 - j0;
  const int kmax = a.zs + R - 1 - k0;
  const int kc = min(2 * lc, (kmax / 2) * 2);
  // every lane's pair is inside (a wave-uniform branch; folded tiles always
  // store under the per-lane mask)
  const bool kfull = F == 1 && __builtin_amdgcn_readfirstlane((int)(k0 + TK <= a.zs));
  uint32_t off[MJ];
  bool val0[MJ], val1[MJ];  // the lane's first / second point is inside
  int rowin[MJ];            // the row is inside the interior (uniform; folded: 1)
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int row = (w * MJ + m) * F + lf;
    off[m] = (uint32_t)(((int64_t)(min(row, jmax) + R) * Pj + (kc + R)) * 2);
    rowin[m] = F == 1 ? __builtin_amdgcn_readfirstlane((int)(j0 + row < a.ys)) : 1;
    val0[m] = (j0 + row < a.ys) && (k0 + 2 * lc < a.zs);
    val1[m] = (j0 + row < a.ys) && (k0 + 2 * lc + 1 < a.zs);
  }
  uint32_t hoff[HJ];
  int hrow[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) {
    const int hh = (w + NW * h) * F + lf;  // halo row id 0 .. 2R-1
    const int trow = hh < R ? hh - R : TJ + hh - R;
    hrow[h] = hh < 2 * R ? (hh < R ? hh : TJ + hh) : -1;  // LDS row, -1: none
    hoff[h] = (uint32_t)(((int64_t)(min(trow, jmax) + R) * Pj + (kc + R)) * 2);
  }
  // k-halo.  F = 1: thread tid < TJ 2R loads one 16-bit cell (row tid / 2R,
  // R columns on either side of the tile).  F > 1: thread tid < TJ 2KH loads
  // one 32-bit word (row tid / 2KH, KH words on either side), clamped like
  // the lanes' pairs: a word past the last readable column is needed by no
  // point inside; for odd R the outer half of a left word (column k0 - R - 1)
  // is never read, and lies inside the allocation (row >= 0 has R rows of
  // halo above it).
  bool has_kh;
  uint32_t khoff = 0, khpos = 0;
  if constexpr (F == 1) {
    has_kh = tid < TJ * 2 * R;
    if (has_kh) {
      const int row = tid / (2 * R), c = tid % (2 * R);
      const int tcol = c < R ? c - R : TK + c - R;
      khoff = (uint32_t)(((int64_t)(min(row, jmax) + R) * Pj + (min(tcol, kmax) + R)) * 2);
      khpos = (uint32_t)((row + R) * 2 * LWW + 2 * KH + tcol);  // half index
    }
  } else {
    has_kh = tid < TJ * 2 * KH;
    if (has_kh) {
      const int row = tid / (2 * KH), c = tid % (2 * KH);
      const int wc = c < KH ? c - KH : LL + c - KH;  // word column from the tile's first
      const int tcol = min(2 * wc, (kmax / 2) * 2);
      khoff = (uint32_t)(((int64_t)(min(row, jmax) + R) * Pj + (tcol + R)) * 2);
      khpos = (uint32_t)(((row + R) * LWW + KH + wc) * 2);  // half index
    }
  }

  auto uni = [](const void *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (GB *)(((uint64_t)hi << 32) | lo);
  };
  auto pl = [&](const char *b, int i) { return uni(b + (int64_t)i * P * 2); };
  // Every load is unconditional and goes through a resource whose range is
  // one plane (under 2 GiB): lanes with nothing to load (no halo row or
  // column) pass kNone and get 0 without a memory access.
  const uint32_t plane_bytes = (uint32_t)(P * 2);
  constexpr uint32_t kNone = 0x80000000u;
  auto ld32 = [&](GB *b, uint32_t o) { return buf_load<uint32_t, 0>(buf_res_n(b, plane_bytes), o); };
  auto ld32nt = [&](GB *b, uint32_t o) {
    return buf_load<uint32_t, 2>(buf_res_n(b, plane_bytes), o);
  };
  auto ld16 = [&](GB *b, uint32_t o) {
    return (uint32_t)buf_load<uint16_t, 0>(buf_res_n(b, plane_bytes), o);
  };
  auto ldkh = [&](GB *b, uint32_t o) {  // a k-halo cell
    if constexpr (F == 1)
      return ld16(b, o);
    else
      return ld32(b, o);
  };
#pragma unroll
  for (int h = 0; h < HJ; h++)
    if (hrow[h] < 0) hoff[h] = kNone;
  if (!has_kh) khoff = kNone;

  // LDS: per-lane byte bases; every access adds a compile-time offset.  The
  // word base passes through an empty asm statement at each step so the
  // compiler does not keep one register per distinct address.
  const uint32_t l0 = (uint32_t)(uintptr_t)(LW32 *)&lds[0];
  uint32_t lb = l0 + (uint32_t)(((w * MJ * F + lf) * LWW + lc) * 4);
  uint32_t hadr[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) hadr[h] = l0 + (uint32_t)((max(hrow[h], 0) * LWW + KH + lc) * 4);
  const uint32_t khadr = l0 + khpos * 2;
  auto w32 = [](uint32_t adr) -> LW32 & { return *(LW32 *)(uintptr_t)adr; };
  auto w16 = [](uint32_t adr) -> LW16 & { return *(LW16 *)(uintptr_t)adr; };
  constexpr uint32_t BUFB = G::SIZE * 4;

  // Rings.  u: plane i - R + t of step k sits in slot (t + k) % QU for the
  // lanes' rows (q) and the halo cells (hq, kq); t = 0 .. 2R are read by the
  // stencil, t = 2R+1 .. 2R+D are in flight: the load of plane i+D+R issued
  // at step k lands in slot (2R+D+k) % QU, free since plane i-R-1 left.  s
  // and v: plane i+D, loaded at step k, lands in slot (k+D) % SV and is read
  // D steps later.  Loads go straight into their ring slots, so no register
  // is copied.
  uint32_t q[MJ][QU], hq[HJ][QU], kq[QU], sq[MJ][SV], vq[MJ][SV];
#pragma unroll
  for (int t = 0; t < 2 * R; t++) {
    GB *Up = pl(Ub, ib - R + t);
#pragma unroll
    for (int m = 0; m < MJ; m++) q[m][t] = ld32(Up, off[m]);
  }
#pragma unroll
  for (int t = 0; t < R; t++) {
    GB *Up = pl(Ub, ib + t);
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R + t] = ld32(Up, hoff[h]);
    kq[R + t] = ldkh(Up, khoff);
  }
  // u(i+R) with its halo, s(i), v(i) into the slots of step k's plane + d
  // (d = 0, 1: the first two planes; d = 2: the prefetch of step k).  s and
  // v are touched once per sweep: nontemporal loads keep them from
  // displacing u's lines, which neighbouring tiles re-read.
  auto load_plane = [&](GB *Upl, GB *Spl, GB *Vpl, auto kd) __attribute__((always_inline)) {
    constexpr int kk = decltype(kd)::value;  // step k + d
    constexpr int su = (2 * R + kk) % QU, ss = kk % SV;
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      q[m][su] = ld32(Upl, off[m]);
      sq[m][ss] = ld32nt(Spl, off[m]);
      vq[m][ss] = ld32nt(Vpl, off[m]);
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][su] = ld32(Upl, hoff[h]);
    kq[su] = ldkh(Upl, khoff);
  };
  // Plane pointers of the prefetch (u(i+2+R), s(i+2), v(i+2)), one plane
  // further at every step while the plane is inside the chunk (they stay on
  // the chunk's last plane, whose lines the workgroup has just read), and
  // the store pointer v(i).  They pass through an empty asm statement at
  // each step, so the compiler keeps four scalar pointers instead of one per
  // plane of the unrolled loop.
  const int ilast = ie - 1;
  GB *Un = pl(Ub, min(ib + D, ilast) + R), *Sn = pl(Sb, min(ib + D, ilast)),
     *Vn = pl(Vb, min(ib + D, ilast)), *Vo = pl(Vb, ib);
  static_for<D>([&](auto dc) __attribute__((always_inline)) {
    const int p = min(ib + decltype(dc)::value, ilast);
    load_plane(pl(Ub, p + R), pl(Sb, p), pl(Vb, p), dc);
  });
  const int64_t PLB = P * 2;

  float mx = 0.0f;
  auto step = [&](int i, auto kcst) __attribute__((always_inline)) {
    constexpr int k = decltype(kcst)::value;
    constexpr int sC = (R + k) % QU;  // slot of the centre plane i
    constexpr int sv = k % SV;        // slot of s(i), v(i)
    constexpr uint32_t BO = (k & 1) * BUFB;
    load_plane(Un, Sn, Vn, std::integral_constant<int, k + D>{});
    const BufRes Vp = buf_res(Vo);
    const bool more = i + D + 1 <= ilast;  // the next prefetch's plane is inside
    Un = more ? Un + PLB : Un;
    Sn = more ? Sn + PLB : Sn;
    Vn = more ? Vn + PLB : Vn;
    Vo += PLB;
    asm volatile("" : "+s"(Un), "+s"(Sn), "+s"(Vn), "+s"(Vo));

    asm volatile("" : "+v"(lb));
#pragma unroll
    for (int m = 0; m < MJ; m++) w32(lb + BO + (uint32_t)(((m * F + R) * LWW + KH) * 4)) = q[m][sC];
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0) w32(hadr[h] + BO) = hq[h][sC];
    if (has_kh) {
      if constexpr (F == 1)
        w16(khadr + BO) = (uint16_t)kq[sC];
      else
        w32(khadr + BO) = kq[sC];
    }
    __syncthreads();

    // Steps past the chunk (in the last turn of the rings) keep the barrier
    // count but compute and store nothing.
    const bool live = i < ie;
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      if (!live || !rowin[m]) continue;
      const uint32_t cw = q[m][sC];
      // k-neighbour words: kw[KH + p] = the word p to the right (p = -KH .. KH)
      uint32_t kw[2 * KH + 1];
#pragma unroll
      for (int p = -KH; p <= KH; p++)
        kw[KH + p] = p == 0 ? cw : w32(lb + BO + (uint32_t)(((m * F + R) * LWW + KH + p) * 4));
      // j-neighbour words: the lane's own rows m -/+ r from registers, the
      // rest (every row of a folded tile) from LDS
      uint32_t jl[R + 1], jh[R + 1];
#pragma unroll
      for (int r = 1; r <= R; r++) {
        jl[r] = F == 1 && m - r >= 0 ? q[m - r >= 0 ? m - r : 0][sC]
                                     : w32(lb + BO + (uint32_t)(((m * F + R - r) * LWW + KH) * 4));
        jh[r] = F == 1 && m + r < MJ ? q[m + r < MJ ? m + r : 0][sC]
                                     : w32(lb + BO + (uint32_t)(((m * F + R + r) * LWW + KH) * 4));
      }
      float res[2];
      static_for<2>([&](auto cc) __attribute__((always_inline)) {
        constexpr int c = decltype(cc)::value;
        // pair of term t: the two neighbours at distance r along its axis
        auto pair = [&](auto tc) __attribute__((always_inline)) -> float {
          constexpr int t = decltype(tc)::value;
          constexpr int ax = Tm::axis[t], r = Tm::r[t];
          if constexpr (ax == 0) {
            return hadd<c>(q[m][(sC + QU - r) % QU], q[m][(sC + r) % QU]);
          } else if constexpr (ax == 1) {
            return hadd<c>(jl[r], jh[r]);
          } else {
            // halves 2KH + c -/+ r of the words kw: same parity on both sides
            constexpr int il = 2 * KH + c - r, ih = 2 * KH + c + r;
            return hadd<il & 1>(kw[il >> 1], kw[ih >> 1]);
          }
        };
        float Lp = a.w[0] * pair(std::integral_constant<int, 0>{});
        static_for<T - 1>([&](auto tc) __attribute__((always_inline)) {
          constexpr int t = decltype(tc)::value + 1;
          Lp = __builtin_fmaf(a.w[t], pair(std::integral_constant<int, t>{}), Lp);
        });
        const float t0 = __builtin_fmaf(K.kc0, hwide<c>(cw), hnegmul<c>(vq[m][sv], K.kv));
        float o = __builtin_fmaf(K.kc2, hwide<c>(sq[m][sv]), __builtin_fmaf(K.kc1, Lp, t0));
        // keep the last fma apart from the conversion below (two roundings,
        // FP32 then binary16, as the oracle rounds)
        __asm__ volatile("" : "+v"(o));
        res[c] = o;
      });
      using H2 = _Float16 __attribute__((ext_vector_type(2)));
      H2 o2;
      o2.x = (_Float16)res[0];
      o2.y = (_Float16)res[1];
      const uint32_t ow = __builtin_bit_cast(uint32_t, o2);
      // max |stored value| = round(max |FP32 result|): rounding to binary16
      // is odd and monotone, so it commutes with max |.|; it is applied
      // once, after the reduction
      if (kfull) {
        buf_store<uint32_t, 2>(ow, Vp, off[m]);
        mx = fmaxf(mx, fmaxf(fabsf(res[0]), fabsf(res[1])));
      } else {
        if (val1[m])
          buf_store<uint32_t, 2>(ow, Vp, off[m]);
        else if (val0[m])
          buf_store<uint16_t, 2>((uint16_t)ow, Vp, off[m]);
        mx = fmaxf(mx, fmaxf(val0[m] ? fabsf(res[0]) : 0.0f, val1[m] ? fabsf(res[1]) : 0.0f));
      }
    }
  };
  for (int i0 = ib; i0 < ie; i0 += UL) {
    static_for<UL>([&](auto kcst) __attribute__((always_inline)) {
      step(i0 + decltype(kcst)::value, kcst);
    });
  }

  mx = (float)(_Float16)mx;
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

// EF > 0: the last k-tile of every row of tiles holds at most 128/EF points
// and runs folded (EF rows per lane group), EF tiles' rows at a time.  The
// items of a chunk come in groups of EF rows of tiles: the rows' ntk - 1
// full tiles in row order, then one folded tile over the group's rows, so
// the narrow column costs 1/EF of its tiles and no item is empty (a
// persistent workgroup's band of items is all work).
template <int NPTS, int MJ, int NW, int EF = 0>
__global__ __launch_bounds__(64 * NW) void sweep16_kernel(const SweepArgs a) {
  using G = Tile16<NPTS, MJ, NW>;
  using GE = Tile16<NPTS, MJ, NW, (EF > 0 ? EF : 1)>;
  constexpr int LS = GE::SIZE > G::SIZE ? GE::SIZE : G::SIZE;
  __shared__ __attribute__((aligned(16))) uint32_t lds[2 * LS];
  __shared__ float red[NW];
  __shared__ int sh_item;
  q16_publish<FMT_F16>(a);
  // LEAPFROG on the stored numbers: with the output exponent eo the update is
  //   q_out = c2n 2^(es-eo) q_s + c1 2^(eu-eo) L(q_u) + c0 2^(eu-eo) q_u - 2^(ev-eo) q_v.
  Coef16 K;
  {
    const Q16Slot &cur = a.qs[a.slot];
    const int eu = cur.exp[a.ubuf], ev = cur.exp[a.vbuf], es = cur.exp[2];
    const int eo = q16_out_exp(cur, a.ubuf, a.vbuf, a.c0, a.c1, a.c2n, a.gain);
    K.kc0 = ldexpf(a.c0, eu - eo);
    K.kc1 = ldexpf(a.c1, eu - eo);
    K.kc2 = ldexpf(a.c2n, es - eo);
    K.kv = ldexpf(1.0f, ev - eo);
  }
  auto tile = [&](int item_) __attribute__((always_inline)) {
    // the item is the same in every lane: keep it and its bounds scalar
    const int item = __builtin_amdgcn_readfirstlane(item_);
    int tk, tj, ch;
    if constexpr (EF > 0) {
      const int m = a.ntk - 1;                              // full tiles per row
      const int ni = a.ntj * m + (a.ntj + EF - 1) / EF;     // items per chunk
      ch = item / ni;
      const int r = item - ch * ni, gs = EF * m + 1;        // group size
      const int g = r / gs, o = r - g * gs;
      if (o < min(EF, a.ntj - g * EF) * m) {
        tj = g * EF + o / m;
        tk = o % m;
      } else {
        tj = g * EF;
        tk = m;
      }
    } else {
      tk = item % a.ntk;
      const int t = item / a.ntk;
      tj = t % a.ntj;
      ch = t / a.ntj;
    }
    const int ib = ch * a.chunk_len;
    const int ie = min(a.xs, ib + a.chunk_len);
    if (ib >= ie) return;
    if constexpr (EF > 0) {
      if (tk == a.ntk - 1) {
        sweep16_tile<NPTS, MJ, NW, EF>(a, K, tk * 128, tj * NW * MJ, ib, ie, lds, red);
        return;
      }
    }
    sweep16_tile<NPTS, MJ, NW>(a, K, tk * 128, tj * NW * MJ, ib, ie, lds, red);
  };
  for_items(a, sh_item, tile);
}

}  // namespace sth

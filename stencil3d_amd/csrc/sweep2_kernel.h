// sweep2_kernel.h -- two LEAPFROG steps per pass over memory (FP32).
//
// The one-step sweep moves 16 B per point-update (read u, v, s; write v)
// and on MI355X sits at the read/write mix bandwidth of HBM.  A pass of this
// kernel reads u_n (with its halo), u_{n-1} and s once and writes u_{n+1}
// and u_{n+2}: 20 B per two point-updates.
//
//  * A workgroup owns an output tile of OJ x OK (j, k) points.  Step one
//    (u_{n+1}) runs on the tile widened by R on every side (E1: EJ rows x 64
//    columns, one wave row of 64 lanes), so step two (u_{n+2}) finds every
//    neighbour on chip.  The widened rows and columns are recomputed by the
//    neighbouring tiles too (overlapped tiling).
//  * Both steps stream along i: step one at plane p (a register queue of u
//    planes p-R-1 .. p+R, the centre plane plus an R halo in LDS, double
//    buffered, as in the one-step sweep), step two at plane p-R-1, whose
//    u_{n+1} planes p-2R-1 .. p-1 are held in an LDS ring of 2R+2 planes.
//    One barrier per plane covers both.
//  * Both steps evaluate the one-step sweep's expression in its order, so a
//    pass is bit-identical to two one-step sweeps.  Outside the interior
//    u_{n+1} takes the value the v field holds there: the one-step scheme
//    writes u_{n+1} into the old v buffer, whose halo no sweep writes.
//  * The four buffers: u_{n+1} goes to a buffer whose halo equals v's,
//    u_{n+2} to one whose halo equals u's (the host keeps a halo-identical
//    twin of each of the two field buffers and swaps them after a pass).
#pragma once

#include "sweep_kernel.h"

namespace sth {

struct Sweep2Args {
  const void *u, *v, *s;  // u_n, u_{n-1}, s: element (0,0,0) of the interior
  void *o1, *o2;          // outputs u_{n+1}, u_{n+2}
  int64_t pitch, plane;   // elements
  int xs, ys, zs;
  int ntk, ntj, chunks, chunk_len, remap;
  float c0, c1;
  float c2a, c2b;  // source coefficients of steps n and n+1
  float w[12];     // LEAPFROG weights (2 * stencil weight), evaluation order
};

template <int NPTS, int MJ, int NW>
struct Fused2Shape {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int EJ = NW * MJ, EK = 64;  // step-one tile
  // Output tile.  For R <= 2 its rows are 48 points (192 B): every tile
  // writes whole 64-byte segments that no other tile touches; rows of
  // 64 - 2R points would share partly written segments with the next tile,
  // which at 1040^3 cost 10-30 % and made the time depend on where the
  // allocator placed the fields.  For R = 4 the narrower row would need 16 %
  // more tiles of the (compute-bound) 25-point pass, so it keeps 56.
  static constexpr int OJ = EJ - 2 * R;
  static constexpr int OK = R <= 2 ? ((EK - 2 * R) / 16) * 16 : EK - 2 * R;
  static constexpr int LW = EK + 2 * R;                   // u plane + R halo
  static constexpr int LU = (EJ + 2 * R) * LW;
  static constexpr int NS = 2 * R + 2;  // ring planes of u_{n+1}
  static constexpr int LR = EJ * EK;
  static constexpr int LDS_BYTES = (2 * LU + NS * LR) * 4;
};

template <int NPTS, int MJ, int NW>
__global__ __launch_bounds__(64 * NW) void sweep2_kernel(const Sweep2Args a) {
  using Tm = Terms<NPTS>;
  using F = Fused2Shape<NPTS, MJ, NW>;
  constexpr int R = Tm::R, T = Tm::T;
  constexpr int EJ = F::EJ, LW = F::LW, NS = F::NS, OJ = F::OJ, OK = F::OK;
  constexpr int HJ = (2 * R + NW - 1) / NW;  // u halo rows per wave
  constexpr int Q = 2 * R + 2;               // u queue: planes p-R-1 .. p+R
  static_assert(MJ * 2 * R <= 64, "k-halo needs at most one entry per thread");
  static_assert(F::LDS_BYTES <= 160 * 1024, "LDS");
  __shared__ float LU[2][F::LU];
  __shared__ float RING[NS][F::LR];

  int bid = blockIdx.x;
  if (a.remap) bid = xcd_remap(bid, (int)gridDim.x);
  const int tk = bid % a.ntk;
  const int tt = bid / a.ntk;
  const int tj = tt % a.ntj;
  const int ch = tt / a.ntj;
  const int ib = ch * a.chunk_len;
  const int ie = min(a.xs, ib + a.chunk_len);
  if (ib >= ie) return;

  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int tid = threadIdx.x;
  const int je = tj * OJ - R, ke = tk * OK - R;  // E1 origin (global j, k)
  const int64_t P = a.plane, Pj = a.pitch;

  // Byte offsets from the corner (i, -R, -R) of a plane.  Reads clamp to the
  // allocated box [-R, n+R): E1 points outside the interior take v there
  // (not a stencil value), so what their clamped neighbours hold never
  // reaches an output.
  auto boff = [&](int j, int k) {
    const int jc = min(max(j, -R), a.ys + R - 1) + R;
    const int kc = min(max(k, -R), a.zs + R - 1) + R;
    return (uint32_t)(((int64_t)jc * Pj + kc) * 4);
  };
  using GB = __attribute__((address_space(1))) char;
  using GF = __attribute__((address_space(1))) float;
  auto uni = [](const void *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (GB *)(((uint64_t)hi << 32) | lo);
  };
  auto pbase = [&](const void *f, int i) {
    const int ic = min(max(i, -R), a.xs + R - 1);
    return uni((const char *)f + ((int64_t)ic * P - R * Pj - R) * 4);
  };
  auto ld = [](GB *p, uint32_t ob) { return *(const GF *)(p + ob); };

  const int kg = ke + lane;
  uint32_t off[MJ];
  bool inner[MJ], outp[MJ];
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int er = wv * MJ + m, jg = je + er;
    off[m] = boff(jg, kg);
    inner[m] = jg >= 0 && jg < a.ys && kg >= 0 && kg < a.zs;
    outp[m] = inner[m] && er >= R && er < R + OJ && lane >= R && lane < R + OK;
  }
  uint32_t hoff[HJ];
  int hrow[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) {
    const int hh = wv + NW * h;  // halo row id 0..2R-1
    const int trow = hh < R ? hh - R : EJ + hh - R;
    hrow[h] = hh < 2 * R ? (hh < R ? hh : EJ + hh) : -1;
    hoff[h] = boff(je + trow, kg);
  }
  const bool has_kh = tid < EJ * 2 * R;
  uint32_t khoff = 0;
  int khpos = 0;
  if (has_kh) {
    const int row = tid / (2 * R), c = tid % (2 * R);
    const int tcol = c < R ? c - R : 64 + c - R;
    khoff = boff(je + row, ke + tcol);
    khpos = (row + R) * LW + (c < R ? c : 64 + c);
  }

  // u queue q[m][t] = u(p - R - 1 + t); s queue sq[m][t] = s(p - R - 1 + t).
  float q[MJ][Q], sq[MJ][R + 2];
  float hq[HJ][R + 1], kq[R + 1];
  const int p0 = ib - R, pe = ie + R;  // step one at p0 .. pe-1, step two at p0+R+1 .. pe
#pragma unroll
  for (int t = 1; t <= 2 * R; t++) {
    GB *Up = pbase(a.u, p0 - R - 1 + t);
#pragma unroll
    for (int m = 0; m < MJ; m++) q[m][t] = ld(Up, off[m]);
  }
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    q[m][0] = 0.0f;
#pragma unroll
    for (int t = 0; t <= R; t++) sq[m][t] = 0.0f;
  }
#pragma unroll
  for (int t = 0; t < R; t++) {
    GB *Up = pbase(a.u, p0 + t);
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][t] = hrow[h] >= 0 ? ld(Up, hoff[h]) : 0.0f;
    kq[t] = has_kh ? ld(Up, khoff) : 0.0f;
  }

  // Staged loads of iteration p: u(p+R) with its halo, v(p), s(p); two
  // sets used in turn, filled two iterations ahead.
  struct Staged {
    float u[MJ], v[MJ], s[MJ], hj[HJ];
    float hk;
  };
  Staged SA, SB;
  auto prefetch = [&](int p, Staged &n) __attribute__((always_inline)) {
    GB *Up = pbase(a.u, p + R);
    GB *Vp = pbase(a.v, p);
    GB *Sp = pbase(a.s, p);
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      n.u[m] = ld(Up, off[m]);
      n.v[m] = ld(Vp, off[m]);
      n.s[m] = ld(Sp, off[m]);
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) n.hj[h] = hrow[h] >= 0 ? ld(Up, hoff[h]) : 0.0f;
    n.hk = has_kh ? ld(Up, khoff) : 0.0f;
  };
  prefetch(p0, SA);
  prefetch(p0 + 1, SB);

  auto slot = [](int x) { return (x + 4 * NS) % NS; };

  auto step = [&](int p, Staged &n) __attribute__((always_inline)) {
    float vv[MJ];
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      q[m][Q - 1] = n.u[m];
      sq[m][R + 1] = n.s[m];
      vv[m] = n.v[m];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R] = n.hj[h];
    kq[R] = n.hk;
    if (p + 2 <= pe) prefetch(p + 2, n);

    // u(p) into LDS: the E1 rows, the halo rows, the k-halo columns.
    float *L = LU[p & 1];
#pragma unroll
    for (int m = 0; m < MJ; m++) L[(wv * MJ + m + R) * LW + lane + R] = q[m][R + 1];
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0) L[hrow[h] * LW + lane + R] = hq[h][0];
    if (has_kh) L[khpos] = kq[0];
    __syncthreads();

    // Step one at plane p: u_{n+1} on E1.
    if (p < pe) {
      float *W = RING[slot(p)];
      const bool pin = p >= 0 && p < a.xs;
#pragma unroll
      for (int m = 0; m < MJ; m++) {
        const int lr = wv * MJ + m + R;
        auto pair = [&](int ax, int r) -> float {
          if (ax == 0) return q[m][R + 1 - r] + q[m][R + 1 + r];
          if (ax == 2) return L[lr * LW + lane + R - r] + L[lr * LW + lane + R + r];
          const float lo = (m - r >= 0) ? q[(m - r) >= 0 ? m - r : 0][R + 1]
                                        : L[(lr - r) * LW + lane + R];
          const float hi = (m + r < MJ) ? q[(m + r) < MJ ? m + r : 0][R + 1]
                                        : L[(lr + r) * LW + lane + R];
          return lo + hi;
        };
        float Lp = a.w[0] * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
        for (int t = 1; t < T; t++) Lp = __builtin_fmaf(a.w[t], pair(Tm::axis[t], Tm::r[t]), Lp);
        const float u0 = q[m][R + 1];
        const float w1 =
            __builtin_fmaf(a.c2a, sq[m][R + 1], __builtin_fmaf(a.c1, Lp, __builtin_fmaf(a.c0, u0, -vv[m])));
        W[(wv * MJ + m) * 64 + lane] = (inner[m] && pin) ? w1 : vv[m];
      }
    }

    // Step two at plane i = p-R-1: u_{n+2} on the output tile; the ring
    // holds u_{n+1} of planes i-R .. i+R.
    const int i = p - R - 1;
    if (i >= ib) {
      const float *W0 = RING[slot(i)];
      GB *Wp = pbase(a.o1, i);
      GB *Zp = pbase(a.o2, i);
#pragma unroll
      for (int m = 0; m < MJ; m++) {
        if (!outp[m]) continue;
        const int idx = (wv * MJ + m) * 64 + lane;
        auto pair = [&](int ax, int r) -> float {
          if (ax == 0) return RING[slot(i - r)][idx] + RING[slot(i + r)][idx];
          if (ax == 1) return W0[idx - r * 64] + W0[idx + r * 64];
          return W0[idx - r] + W0[idx + r];
        };
        float Lp = a.w[0] * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
        for (int t = 1; t < T; t++) Lp = __builtin_fmaf(a.w[t], pair(Tm::axis[t], Tm::r[t]), Lp);
        const float w0 = W0[idx];
        const float z =
            __builtin_fmaf(a.c2b, sq[m][0], __builtin_fmaf(a.c1, Lp, __builtin_fmaf(a.c0, w0, -q[m][0])));
        __builtin_nontemporal_store(w0, (GF *)(Wp + off[m]));
        __builtin_nontemporal_store(z, (GF *)(Zp + off[m]));
      }
    }

#pragma unroll
    for (int m = 0; m < MJ; m++) {
#pragma unroll
      for (int t = 0; t < Q - 1; t++) q[m][t] = q[m][t + 1];
#pragma unroll
      for (int t = 0; t < R + 1; t++) sq[m][t] = sq[m][t + 1];
    }
#pragma unroll
    for (int t = 0; t < R; t++) {
#pragma unroll
      for (int h = 0; h < HJ; h++) hq[h][t] = hq[h][t + 1];
      kq[t] = kq[t + 1];
    }
  };
  for (int p = p0; p <= pe; p += 2) {
    step(p, SA);
    if (p + 1 <= pe) step(p + 1, SB);
  }
}

}  // namespace sth

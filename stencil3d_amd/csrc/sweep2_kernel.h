// sweep2_kernel.h -- two to five LEAPFROG steps per pass over memory (FP32).
//
// The one-step sweep moves 16 B per point-update (read u, v, s; write v)
// and on MI355X sits at the read/write mix bandwidth of HBM.  A pass of this
// kernel reads u_n (with its halo), u_{n-1} and s once and writes the two
// newest fields: 20 B per S point-updates (S = 2 .. 5).  The two-step form
// is described here; each longer pass adds one stage, described at
// sweep2_tile.
//
//  * A workgroup owns an output tile of OJ x OK (j, k) points.  Step one
//    (u_{n+1}) runs on the tile widened by R on every side (E1: EJ rows x 64
//    columns, one wave row of 64 lanes), so step two (u_{n+2}) finds every
//    neighbour on chip.  The widened rows and columns are recomputed by the
//    neighbouring tiles too (overlapped tiling).
//  * Both steps stream along i: step one at plane p (a register queue of u
//    planes p-R-1 .. p+R, the centre plane plus an R halo in LDS, double
//    buffered, as in the one-step sweep), step two at plane i = p-R-1.  Each
//    lane keeps its own u_{n+1} values of planes i-R .. i+R in a register
//    queue (the i neighbours of step two, and its j neighbours inside the
//    wave's rows); plane i of u_{n+1} also goes to LDS, double buffered, for
//    the other j and k neighbours.  One barrier per plane covers both steps.
//  * Both steps evaluate the one-step sweep's expression in its order, so a
//    pass is bit-identical to two one-step sweeps.  Outside the interior
//    u_{n+1} takes the value the v field holds there: the one-step scheme
//    writes u_{n+1} into the old v buffer, whose halo no sweep writes.
//  * The four buffers: u_{n+1} goes to a buffer whose halo equals v's,
//    u_{n+2} to one whose halo equals u's (the host keeps a halo-identical
//    twin of each of the two field buffers and swaps them after a pass).
#pragma once

#include <type_traits>
#include <utility>

#include "sweep_kernel.h"

namespace sth {

struct Sweep2Args {
  const void *u, *v, *s;  // u_n, u_{n-1}, s: element (0,0,0) of the interior
  void *o1, *o2;          // outputs u_{n+1}, u_{n+2}
  int64_t pitch, plane;   // elements
  int xs, ys, zs;
  int ntk, ntj, chunks, chunk_len, remap;
  float c0, c1;
  float c2a, c2b, c2c, c2d, c2e;  // source coefficients of steps n .. n+4
  float w[12];     // LEAPFROG weights (2 * stencil weight), evaluation order
  // FP16 storage (sweep2h_kernel.h): the scale ring, the slot this pass
  // reads, the ring indices of u and v, and the sum of 2|w| over the terms
  Q16Slot *qs;
  int slot, ubuf, vbuf;
  double gain;
};

// S steps per pass (2, 3 or 4): step one runs on the output tile widened by
// WD = (S-1)R on every side.
template <int NPTS, int MJ, int NW, int S = 2>
struct Fused2Shape {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int WD = (S - 1) * R;
  static constexpr int EJ = NW * MJ, EK = 64;  // rows x columns the lanes hold
  // Output tile.  Its rows are 48 points (192 B): every tile writes whole
  // 64-byte segments that no other tile touches.  Rows of 64 - 2R points
  // would share partly written segments with the next tile, which at 1040^3
  // cost 10-30 % (R <= 2) and 15 % (R = 4, although 48-point rows need 16 %
  // more tiles) and made the time depend on where the allocator placed the
  // fields.
  static constexpr int OJ = EJ - 2 * WD;
  static constexpr int OK = 48;
  // The 64 lanes of a row hold the output columns centred, KO = 8 columns
  // on each side: step one's columns (the output widened by WD) and their
  // R-wide k-halo all lie inside the lanes' own 64 columns, so the u plane
  // needs no separate k-halo (OK + 2 (WD + R) <= 64, i.e. S R <= 8).
  static constexpr int KO = (EK - OK) / 2;
  static constexpr int LW = EK;  // LDS u plane row: the lanes' 64 columns
  static constexpr int LU = (EJ + 2 * R) * LW;  // plus the R-wide j-halo rows
  static constexpr int LR = EJ * EK;  // one plane of u_{n+1}, u_{n+2} or s
  static constexpr int SR = (S - 1) * (R + 1);  // s ring slots
  static constexpr int LDS_BYTES = (2 * LU + (2 * (S - 1) + SR) * LR) * 4;
};

// Length of the u_{n+j} register queue of an S-step pass, and its entry
// at the plane of step j+1.
template <int S, int R>
__host__ __device__ constexpr int glen(int j) {
  return j == S - 1 ? 2 * R + 2 : 2 * R + 3;
}
template <int S, int R>
__host__ __device__ constexpr int gcen(int j) {
  return glen<S, R>(j) - R - 2;
}

// One work item of a pass: output tile (tk, tj), i-chunk ch.  Called by the
// single-grid kernel (one item per workgroup) and the batch kernel (items
// of many grids drawn from a queue; LDS is reused between items).
//
// A pass of S steps: step one runs on the output tile widened by (S-1)R
// (E1) at plane p; step M (2 <= M <= S) on the tile widened by (S-M)R at
// plane p-(M-1)(R+1), from the register queue of step M-1's values (its i
// neighbours and in-wave j neighbours) and an LDS plane of them (the other
// j and k neighbours), into a register queue of its own; the last step
// writes u_{n+S-1} and u_{n+S} of the output tile.
template <int NPTS, int MJ, int NW, int S = 2>
__device__ __forceinline__ void sweep2_tile(const Sweep2Args &a, const int tk, const int tj,
                                            const int ch) {
  using Tm = Terms<NPTS>;
  using F = Fused2Shape<NPTS, MJ, NW, S>;
  constexpr int R = Tm::R, T = Tm::T;
  constexpr int EJ = F::EJ, LW = F::LW, OJ = F::OJ, OK = F::OK, KO = F::KO, WD = F::WD,
                SR = F::SR;
  constexpr int HJ = (2 * R + NW - 1) / NW;  // u halo rows per wave
  constexpr int Q = 2 * R + 2;               // u queue: planes p-R-1 .. p+R
  // Queue j (1 <= j < S) holds the lane's u_{n+j} values: 2R+2 entries for
  // the last one, one more for the others (step j+2 takes its -u_{n+j} term
  // from the oldest entry).  GL: the array length.
  constexpr int GL = 2 * R + 3;
  // Steps per loop iteration: a multiple of the u queue length and of the
  // s ring ((S-1)(R+1) slots).
  constexpr int U = (S - 1) % 2 == 0 ? (S - 1) * (R + 1) : 2 * (S - 1) * (R + 1);
  static_assert(S >= 2 && S <= 5, "two to five steps per pass");
  static_assert(U % SR == 0 && U % 2 == 0, "ring slots and buffers repeat with the loop");
  static_assert(KO >= WD + R, "step one and its k-halo inside the lanes' columns");
  static_assert(F::LDS_BYTES <= 160 * 1024, "LDS");
  // One LDS array: the two u planes (LU), two planes each of u_{n+1} ..
  // u_{n+S-1} (step j+1's j and k neighbours), then SR planes of s, each
  // lane's own values (step m needs s(p-(m-1)(R+1))).
  __shared__ float SM[2 * F::LU + (2 * (S - 1) + SR) * F::LR];

  const int ib = ch * a.chunk_len;
  const int ie = min(a.xs, ib + a.chunk_len);
  if (ib >= ie) return;

  const int lane = threadIdx.x & 63;
  // The wave index as a scalar: every condition on rows is wave-uniform and
  // compiles to a branch rather than to exec-mask updates.
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = threadIdx.x;
  // E1's first row; the lanes' first column (the output's first is lane KO)
  const int je = tj * OJ - WD, ke = tk * OK - KO;
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
  // Loads and stores go through a buffer resource per plane (a scalar base,
  // 32-bit lane offsets): no 64-bit address arithmetic per access.  The
  // load range is one plane (under 2 GiB), so lanes that have nothing to
  // load pass kNone and get 0 without touching memory.
  const uint32_t plane_bytes = (uint32_t)(P * 4);
  constexpr uint32_t kNone = 0x80000000u;
  auto ld = [&](GB *p, uint32_t ob) { return buf_load<float, 0>(buf_res_n(p, plane_bytes), ob); };

  // inner[m]: E1 point inside the interior.  A point is an output when its
  // row is (rowout, uniform) and its column is (kout, per lane).
  const int kg = ke + lane;
  const bool kin = kg >= 0 && kg < a.zs;
  const bool kout = kin && lane >= KO && lane < KO + OK;
  uint32_t off[MJ];
  bool inner[MJ];
  // rowst[M][m]: the row is one of step M's (2 <= M < S: the tile widened by
  // (S-M)R, E1 less (M-1)R on each side)
  int rowout[MJ], rowst[S][MJ];
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int er = wv * MJ + m, jg = je + er;
    const bool jin = jg >= 0 && jg < a.ys;
    off[m] = boff(jg, kg);
    inner[m] = jin && kin;
    rowout[m] = __builtin_amdgcn_readfirstlane((int)(jin && er >= WD && er < WD + OJ));
#pragma unroll
    for (int M = 0; M < S; M++)
      rowst[M][m] =
          __builtin_amdgcn_readfirstlane((int)(er >= (M - 1) * R && er < EJ - (M - 1) * R));
  }
  uint32_t hoff[HJ];
  int hrow[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) {
    const int hh = wv + NW * h;  // halo row id 0..2R-1
    const int trow = hh < R ? hh - R : EJ + hh - R;
    hrow[h] = hh < 2 * R ? (hh < R ? hh : EJ + hh) : -1;
    hoff[h] = hrow[h] >= 0 ? boff(je + trow, kg) : kNone;
  }

  // LDS addressing.  Every LDS access of a step is a per-lane base plus a
  // compile-time offset below 64 KiB.  The bases pass through an empty asm
  // statement at each step, so the compiler cannot hoist one register per
  // distinct address out of the loop.
  //   lbL: LU[0] at row wv*MJ, column lane (offsets: buffer, row m+R+dr,
  //        column dc; the k-neighbour reads have dr = 0, so m+R >= 1 rows
  //        keep row*64+dc >= 0);
  //   lbW: W[0] at E1 row wv*MJ-R, column lane-R (offsets: buffer, row
  //        m+R+dr, column R+dc);
  //   lbS: s slot 0 at E1 row wv*MJ, column lane (offsets: slot, row m).
  using LF = __attribute__((address_space(3))) float;
  constexpr uint32_t LUB = F::LU * 4, LRB = F::LR * 4;
  static_assert(2 * LUB <= 65536 && (2 * S - 3) * LRB + (2 * R + MJ) * 256 <= 65536 &&
                    (SR - 1) * LRB + MJ * 256 <= 65536,
                "offsets");
  const uint32_t lu0 = (uint32_t)(uintptr_t)(LF *)&SM[0];
  const uint32_t wbase = lu0 + 2 * LUB, sbase = wbase + 2 * (S - 1) * LRB;
  uint32_t lbL = lu0 + (uint32_t)((wv * MJ * LW + lane) * 4);
  uint32_t lbW = wbase + (uint32_t)(((wv * MJ - R) * 64 + lane - R) * 4);
  uint32_t lbS = sbase + (uint32_t)((wv * MJ * 64 + lane) * 4);
  uint32_t hadr[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) hadr[h] = lu0 + (uint32_t)((max(hrow[h], 0) * LW + lane) * 4);
  auto lds = [](uint32_t adr) -> LF & { return *(LF *)(uintptr_t)adr; };

  // Register queues of the lane's points:
  //   q[m][t]  = u_n(p - R - 1 + t),      t = 0 .. 2R+1;
  //   G[j][m][t] = u_{n+j}(p_j - L_j + 1 + t), t = 0 .. L_j-2, with p_j =
  //              p-(j-1)(R+1) the plane of step j and L_j = glen(j) (t = L_j-1
  //              takes step j's result at p_j before the shift).  Step j+1
  //              reads entries gcen(j) -+ R (its plane p_{j+1} and i
  //              neighbours), step j+2 entry 0 (u_{n+j} at p_{j+2}).
  float q[MJ][Q], G[S][MJ][GL];
  float hq[HJ][R + 1];
  // Step one from p0 (its planes past ie-1+WD are never used), the last
  // step (output plane p - (S-1)(R+1)) up to plast.
  const int p0 = ib - WD, plast = ie + (S - 1) * (R + 1) - 1;
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
    for (int j = 1; j < S; j++)
#pragma unroll
      for (int t = 0; t < glen<S, R>(j); t++) G[j][m][t] = 0.0f;
  }
#pragma unroll
  for (int t = 0; t < R; t++) {
    GB *Up = pbase(a.u, p0 + t);
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][t] = ld(Up, hoff[h]);
  }

  // Staged loads of step p: u(p+R) with its halo, v(p), s(p); two sets used
  // in turn, filled two steps ahead.  The loads are unconditional (waves
  // without a halo entry load at kNone), so no staged value has to be
  // carried over a skipped load.
  struct Staged {
    float u[MJ], v[MJ], s[MJ], hj[HJ];
  };
  Staged SA, SB;
  auto prefetch = [&](GB *Up, GB *Vp, GB *Sp, Staged &n) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      n.u[m] = ld(Up, off[m]);
      n.v[m] = ld(Vp, off[m]);
      n.s[m] = ld(Sp, off[m]);
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) n.hj[h] = ld(Up, hoff[h]);
  };
  prefetch(pbase(a.u, p0 + R), pbase(a.v, p0), pbase(a.s, p0), SA);
  prefetch(pbase(a.u, p0 + 1 + R), pbase(a.v, p0 + 1), pbase(a.s, p0 + 1), SB);

  // Plane pointers advanced by one plane per step, each held inside
  // [-R, xs+R): the prefetch of step p reads u(p+2+R), v(p+2), s(p+2); the
  // last step stores plane p-(S-1)(R+1), from plane ib on.
  const int64_t PB = P * 4;
  const int last = a.xs + R - 1;
  GB *Un = pbase(a.u, p0 + 2 + R);
  GB *Vn = pbase(a.v, p0 + 2);
  GB *Sn = pbase(a.s, p0 + 2);
  GB *Wo = pbase(a.o1, ib);
  GB *Zo = pbase(a.o2, ib);

  // Step of plane p = p0 + U*it + kc: the LDS buffers and the staging set
  // alternate with kc and every queue position is a compile-time constant.
  // The loop runs whole iterations of U steps; steps past plast compute
  // values nobody stores (their loads stay on the last planes), so no
  // register state depends on a skipped step.
  auto step = [&](int p, Staged &n, auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    float vv[MJ], s1[MJ], sc[S + 1][MJ];
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      q[m][Q - 1] = n.u[m];
      s1[m] = n.s[m];
      vv[m] = n.v[m];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R] = n.hj[h];
    prefetch(Un, Vn, Sn, n);
    // Each pointer stays at its next plane clamped to [-R, last]: it moves
    // on only while that plane is inside.  The first step of a chunk is WD
    // planes below it, so the lower bound only binds for WD > R+2 (v, s)
    // and WD > 2R+2 (u); elsewhere the test folds away.
    constexpr bool low_v = WD > R + 2, low_u = WD > 2 * R + 2;
    Un = ((!low_u || p + 2 + R >= -R) && p + 3 + R <= last) ? Un + PB : Un;
    Vn = ((!low_v || p + 2 >= -R) && p + 3 <= last) ? Vn + PB : Vn;
    Sn = ((!low_v || p + 2 >= -R) && p + 3 <= last) ? Sn + PB : Sn;

    asm volatile("" : "+v"(lbL), "+v"(lbW), "+v"(lbS));
    // The s ring of (S-1)(R+1) slots (the lane's own words, so no barrier is
    // involved): step M of plane p-(M-1)(R+1) finds its s in slot
    // k + (S-M)(R+1); the last step's s and s(p) share slot k, the old value
    // read before the new one is stored.
#pragma unroll
    for (int m = 0; m < MJ; m++) {
#pragma unroll
      for (int M = S; M >= 2; M--)
        sc[M][m] = lds(lbS + (uint32_t)(((k + (S - M) * (R + 1)) % SR) * LRB) + m * 256);
      lds(lbS + (uint32_t)((k % SR) * LRB) + m * 256) = s1[m];
    }
    constexpr uint32_t LO = (k & 1) * LUB;  // this step's u buffer
    // LU element (row m+R+dr, column R+dc of this lane's rows)
    auto Lat = [&](int m, int dr, int dc) -> LF & {
      return lds(lbL + LO + (uint32_t)(((m + R + dr) * LW + dc) * 4));
    };
    // This step's buffer of u_{n+j} (E1 row of row m shifted by dr, column
    // lane+dc)
    auto Gat = [&](int j, int m, int dr, int dc) -> LF & {
      return lds(lbW + (uint32_t)((2 * (j - 1) + (k & 1)) * LRB) +
                 (uint32_t)(((m + R + dr) * 64 + R + dc) * 4));
    };

    // u(p) into LDS: the lanes' rows and the j-halo rows; u_{n+j} at the
    // plane of step j+1 (p-j(R+1)), j = 1 .. S-1, of the lanes' rows.
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      Lat(m, 0, 0) = q[m][R + 1];
#pragma unroll
      for (int j = 1; j < S; j++) Gat(j, m, 0, 0) = G[j][m][gcen<S, R>(j)];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0) lds(hadr[h] + LO) = hq[h][0];
    __syncthreads();

    // u_{n+M-2} at step M's plane: u_n from the u queue, else the oldest
    // entry of queue M-2.
    auto older = [&](auto jc, int m) __attribute__((always_inline)) -> float {
      constexpr int J = decltype(jc)::value;
      if constexpr (J == 0)
        return q[m][0];
      else
        return G[J][m][0];
    };
    auto c2of = [&](int M) -> float {
      return M == 2 ? a.c2b : M == 3 ? a.c2c : M == 4 ? a.c2d : a.c2e;
    };

    // The last step writes plane io = p-(S-1)(R+1) of the output tile: the
    // two newest fields there.  Rows without output points (a wave-uniform
    // condition) are skipped.
    const int io = p - (S - 1) * (R + 1);
    const bool store = io >= ib && io < ie;
    const BufRes Wr = buf_res(Wo), Zr = buf_res(Zo);
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      if (!rowout[m]) continue;
      constexpr int J = S - 1, C = gcen<S, R>(J);
      auto pair = [&](int ax, int r) -> float {
        if (ax == 0) return G[J][m][C - r] + G[J][m][C + r];
        if (ax == 2) return Gat(J, m, 0, -r) + Gat(J, m, 0, r);
        const float lo = (m - r >= 0) ? G[J][(m - r) >= 0 ? m - r : 0][C] : Gat(J, m, -r, 0);
        const float hi = (m + r < MJ) ? G[J][(m + r) < MJ ? m + r : 0][C] : Gat(J, m, r, 0);
        return lo + hi;
      };
      float Lp = a.w[0] * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
      for (int t = 1; t < T; t++) Lp = __builtin_fmaf(a.w[t], pair(Tm::axis[t], Tm::r[t]), Lp);
      const float cen = G[J][m][C];
      const float z = __builtin_fmaf(
          c2of(S), sc[S][m],
          __builtin_fmaf(a.c1, Lp,
                         __builtin_fmaf(a.c0, cen, -older(std::integral_constant<int, J - 1>{}, m))));
      if (store && kout) {
        buf_store<float, 2>(cen, Wr, off[m]);
        buf_store<float, 2>(z, Zr, off[m]);
      }
    }
    // Steps S-1 .. 2, each into the newest entry of its queue.  Outside the
    // interior u_{n+M} takes u_{n+M-2}'s value there (the one-step scheme
    // writes u_{n+M} into u_{n+M-2}'s buffer, whose halo no sweep writes).
    // (Only instantiated for S >= 3: an unused closure over the queues
    // changed the register allocation of the two-step pass.)
    if constexpr (S >= 3) {
      // Step M (M >= 2) at plane p-(M-1)(R+1) of row m: the stencil of u_{n+M-1}
      // (queue M-1: i neighbours and in-wave j neighbours; LDS: the others).
      auto lap = [&](auto jc, int m) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value, C = gcen<S, R>(J);
        auto pair = [&](int ax, int r) -> float {
          if (ax == 0) return G[J][m][C - r] + G[J][m][C + r];
          if (ax == 2) return Gat(J, m, 0, -r) + Gat(J, m, 0, r);
          const float lo = (m - r >= 0) ? G[J][(m - r) >= 0 ? m - r : 0][C] : Gat(J, m, -r, 0);
          const float hi = (m + r < MJ) ? G[J][(m + r) < MJ ? m + r : 0][C] : Gat(J, m, r, 0);
          return lo + hi;
        };
        float Lp = a.w[0] * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
        for (int t = 1; t < T; t++) Lp = __builtin_fmaf(a.w[t], pair(Tm::axis[t], Tm::r[t]), Lp);
        return Lp;
      };
      static_for<S - 2>([&](auto ic) __attribute__((always_inline)) {
        constexpr int M = S - 1 - decltype(ic)::value;  // S-1 .. 2
        const int pm = p - (M - 1) * (R + 1);
        const bool pinm = pm >= 0 && pm < a.xs;
#pragma unroll
        for (int m = 0; m < MJ; m++) {
          const float old = older(std::integral_constant<int, M - 2>{}, m);
          if (!rowst[M][m]) {
            G[M][m][glen<S, R>(M) - 1] = old;
            continue;
          }
          const float Lp = lap(std::integral_constant<int, M - 1>{}, m);
          const float cen = G[M - 1][m][gcen<S, R>(M - 1)];
          const float z = __builtin_fmaf(c2of(M), sc[M][m],
                                         __builtin_fmaf(a.c1, Lp, __builtin_fmaf(a.c0, cen, -old)));
          G[M][m][glen<S, R>(M) - 1] = (inner[m] && pinm) ? z : old;
        }
      });
    }
    if (store) {
      Wo += PB;
      Zo += PB;
    }

    // Step one at plane p: u_{n+1} on E1, into the newest queue entry.
    const bool pin = p >= 0 && p < a.xs;
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      auto pair = [&](int ax, int r) -> float {
        if (ax == 0) return q[m][R + 1 - r] + q[m][R + 1 + r];
        if (ax == 2) return Lat(m, 0, -r) + Lat(m, 0, r);
        const float lo = (m - r >= 0) ? q[(m - r) >= 0 ? m - r : 0][R + 1] : Lat(m, -r, 0);
        const float hi = (m + r < MJ) ? q[(m + r) < MJ ? m + r : 0][R + 1] : Lat(m, r, 0);
        return lo + hi;
      };
      float Lp = a.w[0] * pair(Tm::axis[0], Tm::r[0]);
#pragma unroll
      for (int t = 1; t < T; t++) Lp = __builtin_fmaf(a.w[t], pair(Tm::axis[t], Tm::r[t]), Lp);
      const float u0 = q[m][R + 1];
      const float w1 =
          __builtin_fmaf(a.c2a, s1[m], __builtin_fmaf(a.c1, Lp, __builtin_fmaf(a.c0, u0, -vv[m])));
      G[1][m][glen<S, R>(1) - 1] = (inner[m] && pin) ? w1 : vv[m];
    }

#pragma unroll
    for (int m = 0; m < MJ; m++) {
#pragma unroll
      for (int t = 0; t < Q - 1; t++) q[m][t] = q[m][t + 1];
#pragma unroll
      for (int j = 1; j < S; j++)
#pragma unroll
        for (int t = 0; t < glen<S, R>(j) - 1; t++) G[j][m][t] = G[j][m][t + 1];
    }
#pragma unroll
    for (int t = 0; t < R; t++) {
#pragma unroll
      for (int h = 0; h < HJ; h++) hq[h][t] = hq[h][t + 1];
    }
  };
  for (int p = p0; p <= plast; p += U) {
    static_for<U>([&](auto kc) __attribute__((always_inline)) {
      constexpr int k = decltype(kc)::value;
      step(p + k, (k & 1) ? SB : SA, kc);
    });
  }
}

template <int NPTS, int MJ, int NW, int S>
__global__ __launch_bounds__(64 * NW) void sweep2_kernel(const Sweep2Args a) {
  // Work item -> (chunk, tile): k fastest, then j, then the chunk.  (Tile
  // orders that group the workgroups of an XCD into compact 2D patches, so
  // their overlapping reads would meet in L2, measured 2-8 % slower.)
  int bid = blockIdx.x;
  if (a.remap) bid = xcd_remap(bid, (int)gridDim.x);
  const int tk = bid % a.ntk;
  const int tt = bid / a.ntk;
  sweep2_tile<NPTS, MJ, NW, S>(a, tk, tt % a.ntj, tt / a.ntj);
}

// ---- many grids ("shots") in one launch per pass ----------------------------

// A shot's u/v buffers: base a, base b and their twins (element (0,0,0) of
// the interior), and s.  A pass reads the pair of phase ph (0: bases,
// 1: twins) and writes the other pair.
struct Batch2Shot {
  const void *buf[4];  // a, b, twin of a, twin of b
  const void *s;
  int64_t pitch, plane;
  int xs, ys, zs, ntk, ntj, chunk_len;
};

struct Batch2Item {
  int shot, tk, tj, ch;
};

struct Batch2Args {
  const Batch2Shot *shots;
  const Batch2Item *items;  // sorted work list
  int total_items;
  unsigned *ticket;      // queue counter
  unsigned ticket_base;  // tickets taken by earlier launches since the last reset
  int ub;                // buffer of u: 0 = a, 1 = b
  int phase;             // 0: read the bases, write the twins; 1: the reverse
  float c0, c1, c2a, c2b, c2c, c2d, c2e;
  float w[12];
};

// Every block takes tickets until it draws one past the end, so a launch
// consumes exactly total_items + gridDim.x tickets.  S steps per pass.
template <int NPTS, int MJ, int NW, int S>
__global__ __launch_bounds__(64 * NW) void batch2_kernel(const Batch2Args b) {
  __shared__ int sh_item;
  for (;;) {
    if (threadIdx.x == 0) sh_item = (int)(atomicAdd(b.ticket, 1u) - b.ticket_base);
    __syncthreads();
    const int item = sh_item;
    __syncthreads();
    if (item >= b.total_items) break;
    const Batch2Item it = b.items[item];
    const Batch2Shot &sh = b.shots[it.shot];
    const int vb = 1 - b.ub, in = 2 * b.phase, out = 2 - in;
    Sweep2Args a;
    a.u = sh.buf[b.ub + in];
    a.v = sh.buf[vb + in];
    a.s = sh.s;
    // the two newest fields go to the twins of the buffers the one-step
    // scheme writes them to: S even, u_{n+S-1} over v and u_{n+S} over u;
    // S odd, u_{n+S-1} over u and u_{n+S} over v
    a.o1 = const_cast<void *>(sh.buf[(S % 2 == 0 ? vb : b.ub) + out]);
    a.o2 = const_cast<void *>(sh.buf[(S % 2 == 0 ? b.ub : vb) + out]);
    a.pitch = sh.pitch;
    a.plane = sh.plane;
    a.xs = sh.xs;
    a.ys = sh.ys;
    a.zs = sh.zs;
    a.ntk = sh.ntk;
    a.ntj = sh.ntj;
    a.chunks = 0;
    a.chunk_len = sh.chunk_len;
    a.remap = 0;
    a.c0 = b.c0;
    a.c1 = b.c1;
    a.c2a = b.c2a;
    a.c2b = b.c2b;
    a.c2c = b.c2c;
    a.c2d = b.c2d;
    a.c2e = b.c2e;
#pragma unroll
    for (int t = 0; t < 12; t++) a.w[t] = b.w[t];
    sweep2_tile<NPTS, MJ, NW, S>(a, it.tk, it.tj, it.ch);
    __syncthreads();  // LDS is reused by the next item
  }
}

}  // namespace sth

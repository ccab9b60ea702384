// sweep2h_kernel.h -- two LEAPFROG steps per pass over FP16 storage.
//
// The FP32 two-step pass (sweep2_kernel.h) with the data kept in its stored
// binary16 form, as in the packed one-step sweep (sweep16_kernel.h):
//
//  * a lane holds two consecutive k points as one 32-bit word in every
//    queue and LDS plane, so the 64 lanes of a row cover 128 columns (E1:
//    EJ rows x 128 columns); the output tile is 96 columns (192 B, three
//    whole 64-byte segments no other tile writes) centred with KO = 16
//    columns on each side, enough for step one's columns widened by R and
//    their own R-wide k-halo, so the pass loads no separate k-halo;
//  * neighbour pairs are one v_fma_mix_f32 each, the centre, v and s terms
//    take their halves inside the fused multiply-adds, and results round to
//    binary16 in pairs (v_cvt_pk_f16_f32);
//  * step two reads u_{n+1} as stored: step one's results are rounded to
//    binary16 (with their output exponent) before they enter the queue and
//    the LDS plane, so the pass is the one-step scheme's two steps;
//  * the output exponents: u_{n+1}'s from the rigorous bound
//    B1 = (|c0| + |c1| gain) M_u + M_v + |c2a| M_s of the maxima M of the
//    stored inputs (as the one-step sweep); u_{n+2}'s from the same bound
//    with B1 in place of u_{n+1}'s maximum, which the pass does not know
//    before it ends: B2 = (|c0| + |c1| gain) B1 + M_u + |c2b| M_s.  B1 bounds
//    every stored value of u_{n+1}, so no store overflows;
//  * the maxima of both stored outputs are reduced in the pass (as unsigned
//    maxima of the stored bits without the sign, which order like the
//    magnitudes) and published for the next pass.
#pragma once

#include <utility>

#include "sweep16_kernel.h"
#include "sweep2_kernel.h"

namespace sth {

template <int NPTS, int MJ, int NW>
struct Fused2hShape {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int EJ = NW * MJ;       // E1 rows
  static constexpr int OJ = EJ - 2 * R;    // output rows
  static constexpr int OK = 96, KO = 16;   // output columns; lanes' columns on each side
  static constexpr int KH = (R + 1) / 2;   // padding words on each side of an LDS row
  static constexpr int LW = 64 + 2 * KH;   // words per LDS row
  static constexpr int LU = (EJ + 2 * R) * LW;  // u plane with the j-halo rows
  static constexpr int LP = EJ * LW;            // u_{n+1} plane
  static constexpr int LS = EJ * 64;            // one slot of the s ring
  static constexpr int SR = R + 1;              // s ring slots
  static constexpr int LDS_WORDS = 2 * LU + 2 * LP + SR * LS;
};

// The coefficients of both steps on the stored numbers (e1, e2: the output
// exponents of the two steps).
struct Coef2h {
  float k0a, k1a, k2a, kva;  // step one: c0 2^(eu-e1), c1 2^(eu-e1), c2a 2^(es-e1), 2^(ev-e1)
  float k0b, k1b, k2b, kvb;  // step two: c0 2^(e1-e2), c1 2^(e1-e2), c2b 2^(es-e2), 2^(eu-e2)
};

// The output exponents of the pass's two steps (see the file comment).
__host__ __device__ inline void q16_pass_exps(const Q16Slot &cur, int ubuf, int vbuf, float c0,
                                              float c1, float c2a, float c2b, double gain,
                                              int *e1, int *e2) {
  auto M = [&](int b) {
    float m;
    memcpy(&m, &cur.maxbits[b], 4);
    return ldexp((double)m, cur.exp[b]);
  };
  const double gmul = fabs((double)c0) + fabs((double)c1) * gain;
  const double B1 = gmul * M(ubuf) + M(vbuf) + fabs((double)c2a) * M(2);
  const double B2 = gmul * B1 + M(ubuf) + fabs((double)c2b) * M(2);
  *e1 = scale_exp(B1);
  *e2 = scale_exp(B2);
}

// Packed unsigned maximum of two binary16 magnitudes per word.
__device__ __forceinline__ uint32_t pk_umax16(uint32_t a, uint32_t b) {
  using U2 = unsigned short __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(
      uint32_t, __builtin_elementwise_max(__builtin_bit_cast(U2, a), __builtin_bit_cast(U2, b)));
}

// The Laplacian sums of a lane's two columns (halves 0 and 1), term by term
// in the evaluation order, so the words of a term are dead once both
// columns have used them.  iw(r): the words of the i-neighbours at -r, +r;
// jw(r): the j-neighbours; kw(d): the word d to the right of the lane's
// (d = -KH .. KH).
template <int NPTS, int KH, class IW, class JW, class KWf>
__device__ __forceinline__ void lap_pair(const float *w, IW &&iw, JW &&jw, KWf &&kw, float &L0,
                                         float &L1) {
  using Tm = Terms<NPTS>;
  static_for<Tm::T>([&](auto tc) __attribute__((always_inline)) {
    constexpr int t = decltype(tc)::value;
    constexpr int ax = Tm::axis[t], r = Tm::r[t];
    float p0, p1;
    if constexpr (ax == 2) {
      // halves 2KH + c -/+ r of the words kw(-KH) .. kw(KH)
      constexpr int l0 = 2 * KH - r, h0 = 2 * KH + r, l1 = l0 + 1, h1 = h0 + 1;
      p0 = hadd<l0 & 1>(kw(std::integral_constant<int, (l0 >> 1) - KH>{}),
                        kw(std::integral_constant<int, (h0 >> 1) - KH>{}));
      p1 = hadd<l1 & 1>(kw(std::integral_constant<int, (l1 >> 1) - KH>{}),
                        kw(std::integral_constant<int, (h1 >> 1) - KH>{}));
    } else {
      const auto xy = ax == 0 ? iw(std::integral_constant<int, r>{})
                              : jw(std::integral_constant<int, r>{});
      p0 = hadd<0>(xy.x, xy.y);
      p1 = hadd<1>(xy.x, xy.y);
    }
    if constexpr (t == 0) {
      L0 = w[0] * p0;
      L1 = w[0] * p1;
    } else {
      L0 = __builtin_fmaf(w[t], p0, L0);
      L1 = __builtin_fmaf(w[t], p1, L1);
    }
  });
}

// One work item: output tile (tk, tj), i-chunk ch.  red: 2 NW floats.
template <int NPTS, int MJ, int NW>
__device__ __forceinline__ void sweep2h_tile(const Sweep2Args &a, const Coef2h &K, const int tk,
                                             const int tj, const int ch, uint32_t *SM,
                                             float *red) {
  using Tm = Terms<NPTS>;
  using F = Fused2hShape<NPTS, MJ, NW>;
  constexpr int R = Tm::R, T = Tm::T;
  constexpr int EJ = F::EJ, OJ = F::OJ, OK = F::OK, KO = F::KO, KH = F::KH, LW = F::LW,
                SR = F::SR;
  constexpr int HJ = (2 * R + NW - 1) / NW;  // u halo rows per wave
  constexpr int Q = 2 * R + 2;               // u queue: planes p-R-1 .. p+R
  constexpr int GL = 2 * R + 2;              // u_{n+1} queue: planes p-2R-1 .. p
  constexpr int GC = R;                      // its entry at step two's plane p-R-1
  constexpr int U = 2 * (R + 1);             // steps per loop iteration
  static_assert(KO >= 2 * R, "step one and its k-halo inside the lanes' columns");
  static_assert(KO % 2 == 0 && OK % 2 == 0, "a lane's pair is all output or none");
  static_assert(F::LDS_WORDS * 4 <= 160 * 1024, "LDS");
  static_assert(U % SR == 0 && U % 2 == 0, "ring slots and buffers repeat with the loop");

  const int ib = ch * a.chunk_len;
  const int ie = min(a.xs, ib + a.chunk_len);
  if (ib >= ie) return;

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int je = tj * OJ - R, ke = tk * OK - KO;  // E1's first row; the lanes' first column
  const int64_t P = a.plane, Pj = a.pitch;

  // The lane's pair starts at column kg (even).  Reads clamp to the pairs
  // that cover the allocated box [-R, zs+R): from the even column at or
  // below -R to the even column at or below zs+R-1 (its second element may
  // lie one past the box, inside the row padding or the next row's halo).
  // E1 points outside the interior take v's value (not a stencil value), so
  // what their clamped neighbours hold never reaches an output.
  const int kg = ke + 2 * lane;
  const int klo = -2 * KH, khi = (a.zs + R - 1) & ~1;
  const int kc = min(max(kg, klo), khi);
  // Byte offsets from the plane base, element (i, -R, -2KH): every read pair
  // starts at or after it.
  auto boff = [&](int j) {
    const int jc = min(max(j, -R), a.ys + R - 1) + R;
    return (uint32_t)(((int64_t)jc * Pj + kc + 2 * KH) * 2);
  };
  using GB = __attribute__((address_space(1))) char;
  auto uni = [](const void *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (GB *)(((uint64_t)hi << 32) | lo);
  };
  // element (i, -R, -2KH) of plane i, i clamped to the allocated planes
  // (2KH >= R: at most one element before the box's corner, inside the
  // allocation's leading 256 B)
  auto pbase = [&](const void *f, int i) {
    const int ic = min(max(i, -R), a.xs + R - 1);
    return uni((const char *)f + ((int64_t)ic * P - R * Pj - 2 * KH) * 2);
  };
  // Load range: the plane and 256 bytes more.  A pair clamped to the box's
  // last column can end up to 4 bytes past the plane (pitch >= zs + 2R may
  // hold exactly), in the next plane or the field's trailing 256 B; the
  // range must not cut it off.
  const uint32_t plane_bytes = (uint32_t)(P * 2) + 256u;
  constexpr uint32_t kNone = 0x80000000u;
  auto ld = [&](GB *p, uint32_t ob) {
    return buf_load<uint32_t, 0>(buf_res_n(p, plane_bytes), ob);
  };

  // Per half: the point is inside the interior in k / an output column.
  const bool kin0 = kg >= 0 && kg < a.zs, kin1 = kg + 1 >= 0 && kg + 1 < a.zs;
  const bool klane = 2 * lane >= KO && 2 * lane < KO + OK;
  const uint32_t kinmask = (kin0 ? 0x0000ffffu : 0u) | (kin1 ? 0xffff0000u : 0u);
  const uint32_t outabs = klane ? (kinmask & 0x7fff7fffu) : 0u;  // magnitudes of output halves
  const bool kout0 = klane && kin0, kout1 = klane && kin1;
  uint32_t off[MJ];
  int rowin[MJ], rowout[MJ];  // wave-uniform
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int er = wv * MJ + m, jg = je + er;
    off[m] = boff(jg);
    rowin[m] = __builtin_amdgcn_readfirstlane((int)(jg >= 0 && jg < a.ys));
    rowout[m] = __builtin_amdgcn_readfirstlane((int)(er >= R && er < R + OJ));
  }
  uint32_t hoff[HJ];
  int hrow[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) {
    const int hh = wv + NW * h;  // halo row id 0 .. 2R-1
    const int trow = hh < R ? hh - R : EJ + hh - R;
    hrow[h] = hh < 2 * R ? (hh < R ? hh : EJ + hh) : -1;
    hoff[h] = hrow[h] >= 0 ? boff(je + trow) : kNone;
  }

  // LDS: per-lane byte bases plus compile-time offsets (each below 64 KiB),
  // passed through an empty asm statement at every step.
  //   lbL: u plane, row wv*MJ, word lane (offsets: buffer, row m+R+dr, word KH+dp);
  //   lbW: u_{n+1} plane, row wv*MJ-R, word lane (offsets: buffer, row m+R+dr, word KH+dp);
  //   lbS: s ring slot 0, row wv*MJ, word lane (offsets: slot, row m).
  using LW32 = __attribute__((address_space(3))) uint32_t;
  constexpr uint32_t LUB = F::LU * 4, LPB = F::LP * 4, LSB = F::LS * 4;
  static_assert(2 * LUB <= 65536 && 2 * LPB <= 65536 && (SR - 1) * LSB + MJ * 256 <= 65536,
                "offsets");
  const uint32_t lu0 = (uint32_t)(uintptr_t)(LW32 *)&SM[0];
  const uint32_t wbase = lu0 + 2 * LUB, sbase = wbase + 2 * LPB;
  uint32_t lbL = lu0 + (uint32_t)((wv * MJ * LW + lane) * 4);
  uint32_t lbW = wbase + (uint32_t)(((wv * MJ - R) * LW + lane) * 4);
  uint32_t lbS = sbase + (uint32_t)((wv * MJ * 64 + lane) * 4);
  uint32_t hadr[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) hadr[h] = lu0 + (uint32_t)((max(hrow[h], 0) * LW + KH + lane) * 4);
  auto lds = [](uint32_t adr) -> LW32 & { return *(LW32 *)(uintptr_t)adr; };

  // Register queues of packed words:
  //   q[m][t] = u_n(p - R - 1 + t), t = 0 .. 2R+1;
  //   G[m][t] = u_{n+1}(p - 2R - 1 + t), t = 0 .. 2R (t = 2R+1 takes step
  //             one's result at p before the shift); step two reads G[GC -+ r].
  uint32_t q[MJ][Q], G[MJ][GL];
  uint32_t hq[HJ][R + 1];
  // Step one from p0 (its planes past ie-1+R are never used), step two
  // (output plane p - R - 1) up to plast.
  const int p0 = ib - R, plast = ie + R;
#pragma unroll
  for (int t = 1; t <= 2 * R; t++) {
    GB *Up = pbase(a.u, p0 - R - 1 + t);
#pragma unroll
    for (int m = 0; m < MJ; m++) q[m][t] = ld(Up, off[m]);
  }
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    q[m][0] = 0u;
#pragma unroll
    for (int t = 0; t < GL; t++) G[m][t] = 0u;
  }
#pragma unroll
  for (int t = 0; t < R; t++) {
    GB *Up = pbase(a.u, p0 + t);
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][t] = ld(Up, hoff[h]);
  }

  // Staged loads of step p: u(p+R) with its halo rows, v(p), s(p); two sets
  // used in turn, filled two steps ahead, every load unconditional.
  struct Staged {
    uint32_t u[MJ], v[MJ], s[MJ], hj[HJ];
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

  // Plane pointers advanced by one plane per step, each held at the last
  // allocated plane: the prefetch of step p reads u(p+2+R), v(p+2), s(p+2);
  // step two stores plane p-R-1, from plane ib on.
  const int64_t PB = P * 2;
  const int last = a.xs + R - 1;
  GB *Un = pbase(a.u, p0 + 2 + R);
  GB *Vn = pbase(a.v, p0 + 2);
  GB *Sn = pbase(a.s, p0 + 2);
  GB *Wo = pbase(a.o1, ib);
  GB *Zo = pbase(a.o2, ib);
  uint32_t mx1 = 0u, mx2 = 0u;  // packed maxima of the stored output magnitudes

  auto step = [&](int p, Staged &n, auto kcst) __attribute__((always_inline)) {
    constexpr int k = decltype(kcst)::value;
    uint32_t vv[MJ], s1[MJ], sc[MJ];
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      q[m][Q - 1] = n.u[m];
      s1[m] = n.s[m];
      vv[m] = n.v[m];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R] = n.hj[h];
    prefetch(Un, Vn, Sn, n);
    Un = p + 3 + R <= last ? Un + PB : Un;
    Vn = p + 3 <= last ? Vn + PB : Vn;
    Sn = p + 3 <= last ? Sn + PB : Sn;

    asm volatile("" : "+v"(lbL), "+v"(lbW), "+v"(lbS));
    // s ring of R+1 slots (the lane's own words): step two's s (plane
    // p-R-1, stored R+1 steps ago) is read from slot k, then s(p) replaces it.
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      sc[m] = lds(lbS + (uint32_t)((k % SR) * LSB) + m * 256);
      lds(lbS + (uint32_t)((k % SR) * LSB) + m * 256) = s1[m];
    }
    constexpr uint32_t LO = (k & 1) * LUB, PO = (k & 1) * LPB;
    auto Lat = [&](int m, int dr, int dp) -> LW32 & {
      return lds(lbL + LO + (uint32_t)(((m + R + dr) * LW + KH + dp) * 4));
    };
    auto Wat = [&](int m, int dr, int dp) -> LW32 & {
      return lds(lbW + PO + (uint32_t)(((m + R + dr) * LW + KH + dp) * 4));
    };
    // u(p) into LDS (the lanes' rows and the j-halo rows); u_{n+1} at step
    // two's plane p-R-1.
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      Lat(m, 0, 0) = q[m][R + 1];
      Wat(m, 0, 0) = G[m][GC];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0) lds(hadr[h] + LO) = hq[h][0];
    __syncthreads();

    // Step two writes plane io = p-R-1 of the output tile: u_{n+1} as stored
    // and u_{n+2}.  Rows without output points are skipped (uniform).
    const int io = p - R - 1;
    const bool store = io >= ib && io < ie;
    const BufRes Wr = buf_res(Wo), Zr = buf_res(Zo);
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      if (!rowout[m]) continue;
      const uint32_t cw = G[m][GC];
      using W2 = uint32_t __attribute__((ext_vector_type(2)));
      float L0, L1;
      lap_pair<NPTS, KH>(
          a.w,
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{G[m][GC - r], G[m][GC + r]};
          },
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{m - r >= 0 ? G[m - r >= 0 ? m - r : 0][GC] : (uint32_t)Wat(m, -r, 0),
                      m + r < MJ ? G[m + r < MJ ? m + r : 0][GC] : (uint32_t)Wat(m, r, 0)};
          },
          [&](auto dc) -> uint32_t {
            constexpr int d = decltype(dc)::value;
            return d == 0 ? cw : (uint32_t)Wat(m, 0, d);
          },
          L0, L1);
      // -u_n at io: the oldest entry of the u queue
      float z[2];
      static_for<2>([&](auto cc) __attribute__((always_inline)) {
        constexpr int c = decltype(cc)::value;
        const float t0 = __builtin_fmaf(K.k0b, hwide<c>(cw), hnegmul<c>(q[m][0], K.kvb));
        float o = __builtin_fmaf(K.k2b, hwide<c>(sc[m]), __builtin_fmaf(K.k1b, c ? L1 : L0, t0));
        __asm__ volatile("" : "+v"(o));
        z[c] = o;
      });
      using H2 = _Float16 __attribute__((ext_vector_type(2)));
      H2 zh;
      zh.x = (_Float16)z[0];
      zh.y = (_Float16)z[1];
      const uint32_t zw = __builtin_bit_cast(uint32_t, zh);
      if (store && rowin[m]) {
        if (kout0 && kout1) {
          buf_store<uint32_t, 2>(cw, Wr, off[m]);
          buf_store<uint32_t, 2>(zw, Zr, off[m]);
        } else if (kout0) {
          buf_store<uint16_t, 2>((uint16_t)cw, Wr, off[m]);
          buf_store<uint16_t, 2>((uint16_t)zw, Zr, off[m]);
        }
        // maxima of the stored magnitudes of the output points
        mx1 = pk_umax16(mx1, cw & outabs);
        mx2 = pk_umax16(mx2, zw & outabs);
      }
    }
    if (store) {
      Wo += PB;
      Zo += PB;
    }

    // Step one at plane p: u_{n+1} on E1, rounded to binary16, into the
    // newest queue entry; points outside the interior take v's value.
    const bool pin = p >= 0 && p < a.xs;
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      const uint32_t cw = q[m][R + 1];
      using W2 = uint32_t __attribute__((ext_vector_type(2)));
      float L0, L1;
      lap_pair<NPTS, KH>(
          a.w,
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{q[m][R + 1 - r], q[m][R + 1 + r]};
          },
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{m - r >= 0 ? q[m - r >= 0 ? m - r : 0][R + 1] : (uint32_t)Lat(m, -r, 0),
                      m + r < MJ ? q[m + r < MJ ? m + r : 0][R + 1] : (uint32_t)Lat(m, r, 0)};
          },
          [&](auto dc) -> uint32_t {
            constexpr int d = decltype(dc)::value;
            return d == 0 ? cw : (uint32_t)Lat(m, 0, d);
          },
          L0, L1);
      float w1[2];
      static_for<2>([&](auto cc) __attribute__((always_inline)) {
        constexpr int c = decltype(cc)::value;
        const float t0 = __builtin_fmaf(K.k0a, hwide<c>(cw), hnegmul<c>(vv[m], K.kva));
        float o = __builtin_fmaf(K.k2a, hwide<c>(s1[m]), __builtin_fmaf(K.k1a, c ? L1 : L0, t0));
        __asm__ volatile("" : "+v"(o));
        w1[c] = o;
      });
      using H2 = _Float16 __attribute__((ext_vector_type(2)));
      H2 wh;
      wh.x = (_Float16)w1[0];
      wh.y = (_Float16)w1[1];
      const uint32_t ww = __builtin_bit_cast(uint32_t, wh);
      // halves inside the interior take the new value, the others v's
      G[m][GL - 1] = (rowin[m] && pin) ? ((ww & kinmask) | (vv[m] & ~kinmask)) : vv[m];
    }

#pragma unroll
    for (int m = 0; m < MJ; m++) {
#pragma unroll
      for (int t = 0; t < Q - 1; t++) q[m][t] = q[m][t + 1];
#pragma unroll
      for (int t = 0; t < GL - 1; t++) G[m][t] = G[m][t + 1];
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

  // Maxima of the item: the larger half, widened, reduced over the
  // workgroup, one atomicMax per output field.
  using H2 = _Float16 __attribute__((ext_vector_type(2)));
  auto hmax = [](uint32_t x) {
    const uint32_t b = max(x & 0xffffu, x >> 16);
    return (float)__builtin_bit_cast(H2, b).x;
  };
  float m1 = hmax(mx1), m2 = hmax(mx2);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    m1 = fmaxf(m1, __shfl_xor(m1, o, 64));
    m2 = fmaxf(m2, __shfl_xor(m2, o, 64));
  }
  if (lane == 0) {
    red[wv] = m1;
    red[NW + wv] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float b1 = red[0], b2 = red[NW];
#pragma unroll
    for (int t = 1; t < NW; t++) {
      b1 = fmaxf(b1, red[t]);
      b2 = fmaxf(b2, red[NW + t]);
    }
    Q16Slot &nx = a.qs[(a.slot + 1) % 3];
    atomicMax(&nx.maxbits[a.vbuf], __float_as_uint(b1));
    atomicMax(&nx.maxbits[a.ubuf], __float_as_uint(b2));
  }
}

template <int NPTS, int MJ, int NW>
__global__ __launch_bounds__(64 * NW) void sweep2h_kernel(const Sweep2Args a) {
  using F = Fused2hShape<NPTS, MJ, NW>;
  __shared__ __attribute__((aligned(16))) uint32_t SM[F::LDS_WORDS];
  __shared__ float red[2 * NW];
  const Q16Slot &cur = a.qs[a.slot];
  int e1, e2;
  q16_pass_exps(cur, a.ubuf, a.vbuf, a.c0, a.c1, a.c2a, a.c2b, a.gain, &e1, &e2);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // The next slot: u_{n+1} lands in v's buffer, u_{n+2} in u's; their
    // maxima are reduced by the blocks into it (cleared by the previous
    // launch); the slot after it is cleared for the next launch.
    Q16Slot &nxt = a.qs[(a.slot + 1) % 3];
    Q16Slot &clr = a.qs[(a.slot + 2) % 3];
    nxt.exp[a.vbuf] = e1;
    nxt.exp[a.ubuf] = e2;
    nxt.exp[2] = cur.exp[2];
    nxt.maxbits[2] = cur.maxbits[2];
    clr.maxbits[0] = clr.maxbits[1] = clr.maxbits[2] = 0u;
  }
  const int eu = cur.exp[a.ubuf], ev = cur.exp[a.vbuf], es = cur.exp[2];
  Coef2h K;
  K.k0a = ldexpf(a.c0, eu - e1);
  K.k1a = ldexpf(a.c1, eu - e1);
  K.k2a = ldexpf(a.c2a, es - e1);
  K.kva = ldexpf(1.0f, ev - e1);
  K.k0b = ldexpf(a.c0, e1 - e2);
  K.k1b = ldexpf(a.c1, e1 - e2);
  K.k2b = ldexpf(a.c2b, es - e2);
  K.kvb = ldexpf(1.0f, eu - e2);
  // the coefficients are the same in every lane: keep them in scalar registers
  auto uf = [](float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
  };
  K.k0a = uf(K.k0a);
  K.k1a = uf(K.k1a);
  K.k2a = uf(K.k2a);
  K.kva = uf(K.kva);
  K.k0b = uf(K.k0b);
  K.k1b = uf(K.k1b);
  K.k2b = uf(K.k2b);
  K.kvb = uf(K.kvb);
  int bid = blockIdx.x;
  if (a.remap) bid = xcd_remap(bid, (int)gridDim.x);
  const int tk = bid % a.ntk;
  const int tt = bid / a.ntk;
  sweep2h_tile<NPTS, MJ, NW>(a, K, tk, tt % a.ntj, tt / a.ntj, SM, red);
}

}  // namespace sth

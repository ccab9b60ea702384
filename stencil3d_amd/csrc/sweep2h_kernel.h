// sweep2h_kernel.h -- two to five LEAPFROG steps per pass over 16-bit storage.
//
// The FP32 multi-step pass (sweep2_kernel.h) with the data kept in its
// stored 16-bit form (FP16, as in the packed one-step sweep of
// sweep16_kernel.h, or BF16):
//
//  * a lane holds two consecutive k points as one 32-bit word in every
//    queue and LDS plane, so the 64 lanes of a row cover 128 columns (E1:
//    EJ rows x 128 columns); the output tile is 96 columns (192 B, three
//    whole 64-byte segments no other tile writes) centred with KO = 16
//    columns on each side, enough for step one's columns widened by
//    (S-1)R and their own R-wide k-halo (S R <= 16), so the pass loads no
//    separate k-halo;
//  * FP16: neighbour pairs are one v_fma_mix_f32 each, the centre, v and s
//    terms take their halves inside the fused multiply-adds, and results
//    round to binary16 in pairs (v_cvt_pk_f16_f32); BF16: a half widens
//    exactly by a shift or a mask, and results round to nearest even in
//    integer arithmetic (the one-step sweep's encoding);
//  * step M+1 reads u_{n+M} as stored: each step's results are rounded to
//    the 16-bit format (with their output exponent) before they enter the
//    queue and the LDS plane, so a pass is S of the one-step scheme's steps;
//  * the output exponents: step M's from the bound
//      B_M = (|c0| + |c1| gain) B_{M-1} + B_{M-2} + |c2_M| M_s,
//      B_0 = M_u, B_{-1} = M_v
//    of the maxima M of the stored inputs: B_1 is the one-step sweep's
//    bound, and each later bound stands in for the maximum of an output
//    the pass does not know before it ends.  B_M bounds every stored value
//    of u_{n+M}, so no store overflows;
//  * the maxima of the two stored outputs are reduced in the pass (as
//    unsigned maxima of the stored bits without the sign, which order like
//    the magnitudes) and published for the next pass.
#pragma once

#include <utility>

#include "sweep16_kernel.h"
#include "sweep2_kernel.h"

namespace sth {

// S steps per pass: step one runs on the output tile widened by WD = (S-1)R
// on every side (E1: EJ rows x 128 columns).
template <int NPTS, int MJ, int NW, int S = 2>
struct Fused2hShape {
  static constexpr int R = Terms<NPTS>::R;
  static constexpr int WD = (S - 1) * R;
  static constexpr int EJ = NW * MJ;       // E1 rows
  static constexpr int OJ = EJ - 2 * WD;   // output rows
  static constexpr int OK = 96, KO = 16;   // output columns; lanes' columns on each side
  static constexpr int KH = (R + 1) / 2;   // padding words on each side of an LDS row
  static constexpr int LW = 64 + 2 * KH;   // words per LDS row
  static constexpr int LU = (EJ + 2 * R) * LW;  // u plane with the j-halo rows
  static constexpr int LP = EJ * LW;            // one plane of u_{n+j}
  static constexpr int LS = EJ * 64;            // one slot of the s ring
  static constexpr int SR = (S - 1) * (R + 1);  // s ring slots
  static constexpr int LDS_WORDS = 2 * LU + 2 * (S - 1) * LP + SR * LS;
};

// The coefficients of the pass's steps on the stored numbers: step M
// (index M-1) computes u_{n+M} with exponent e_M from u_{n+M-1} (exponent
// e_{M-1}) and u_{n+M-2} (e_{M-2}); e_0 = e_u, e_{-1} = e_v:
//   k0 = c0 2^(e_{M-1}-e_M), k1 = c1 2^(e_{M-1}-e_M), k2 = c2_M 2^(e_s-e_M),
//   kv = 2^(e_{M-2}-e_M).
template <int S>
struct Coef2h {
  float k0[S], k1[S], k2[S], kv[S];
};

// The output exponents e_1 .. e_S of a pass's steps (see the file comment).
template <int S>
__host__ __device__ inline void q16_pass_exps(const Q16Slot &cur, int ubuf, int vbuf, float c0,
                                              float c1, const float *c2, double gain, int *e) {
  auto M = [&](int b) {
    float m;
    memcpy(&m, &cur.maxbits[b], 4);
    return ldexp((double)m, cur.exp[b]);
  };
  const double gmul = fabs((double)c0) + fabs((double)c1) * gain;
  const double Ms = M(2);
  double b2 = M(vbuf), b1 = M(ubuf);  // B_{M-2}, B_{M-1}
  for (int s = 0; s < S; s++) {
    const double B = gmul * b1 + b2 + fabs((double)c2[s]) * Ms;
    e[s] = scale_exp(B);
    b2 = b1;
    b1 = B;
  }
}

// A 16-bit pair per 32-bit word: half H widened to FP32, the FP32 sum of
// half H of two words (one rounding: the sum of the widened values), -(half
// H of x * k) (one rounding, as the oracle's -(v * kv)), and two FP32 results
// rounded into one word.
template <int FMT>
struct Pk16;
template <>
struct Pk16<FMT_F16> {
  template <int H>
  __device__ static float wide(uint32_t x) { return hwide<H>(x); }
  template <int H>
  __device__ static float add(uint32_t x, uint32_t y) { return hadd<H>(x, y); }
  template <int H>
  __device__ static float negmul(uint32_t x, float k) { return hnegmul<H>(x, k); }
  __device__ static uint32_t pack(float a, float b) {
    using H2 = _Float16 __attribute__((ext_vector_type(2)));
    H2 r;
    r.x = (_Float16)a;
    r.y = (_Float16)b;
    return __builtin_bit_cast(uint32_t, r);
  }
};
template <>
struct Pk16<FMT_BF16> {
  template <int H>
  __device__ static float wide(uint32_t x) {
    return __uint_as_float(H ? (x & 0xffff0000u) : (x << 16));
  }
  template <int H>
  __device__ static float add(uint32_t x, uint32_t y) { return wide<H>(x) + wide<H>(y); }
  template <int H>
  __device__ static float negmul(uint32_t x, float k) { return -(wide<H>(x) * k); }
  __device__ static uint32_t pack(float a, float b) {
    return (uint32_t)Codec<FMT_BF16>::enc(a) | ((uint32_t)Codec<FMT_BF16>::enc(b) << 16);
  }
};

// Packed unsigned maximum of two 16-bit magnitudes per word.
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
template <int NPTS, int KH, int FMT, class IW, class JW, class KWf>
__device__ __forceinline__ void lap_pair(const float *w, IW &&iw, JW &&jw, KWf &&kw, float &L0,
                                         float &L1) {
  using PK = Pk16<FMT>;
  using Tm = Terms<NPTS>;
  static_for<Tm::T>([&](auto tc) __attribute__((always_inline)) {
    constexpr int t = decltype(tc)::value;
    constexpr int ax = Tm::axis[t], r = Tm::r[t];
    float p0, p1;
    if constexpr (ax == 2) {
      // halves 2KH + c -/+ r of the words kw(-KH) .. kw(KH)
      constexpr int l0 = 2 * KH - r, h0 = 2 * KH + r, l1 = l0 + 1, h1 = h0 + 1;
      p0 = PK::template add<l0 & 1>(kw(std::integral_constant<int, (l0 >> 1) - KH>{}),
                                    kw(std::integral_constant<int, (h0 >> 1) - KH>{}));
      p1 = PK::template add<l1 & 1>(kw(std::integral_constant<int, (l1 >> 1) - KH>{}),
                                    kw(std::integral_constant<int, (h1 >> 1) - KH>{}));
    } else {
      const auto xy = ax == 0 ? iw(std::integral_constant<int, r>{})
                              : jw(std::integral_constant<int, r>{});
      p0 = PK::template add<0>(xy.x, xy.y);
      p1 = PK::template add<1>(xy.x, xy.y);
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

// One step's update of a lane's word: both halves in the one-step sweep's
// order, each result rounded to FP32 and then (in pairs) to the 16-bit
// format.
//   out = k2 s + (k1 L + (k0 c - kv v))
template <int FMT>
__device__ __forceinline__ uint32_t update_pair(float k0, float k1, float k2, float kv,
                                                uint32_t c, uint32_t v, uint32_t s, float L0,
                                                float L1) {
  using PK = Pk16<FMT>;
  float o[2];
  static_for<2>([&](auto cc) __attribute__((always_inline)) {
    constexpr int h = decltype(cc)::value;
    const float t0 = __builtin_fmaf(k0, PK::template wide<h>(c), PK::template negmul<h>(v, kv));
    float x = __builtin_fmaf(k2, PK::template wide<h>(s), __builtin_fmaf(k1, h ? L1 : L0, t0));
    // keep the last fma apart from the conversion below (two roundings,
    // FP32 then 16-bit, as the oracle rounds)
    __asm__ volatile("" : "+v"(x));
    o[h] = x;
  });
  return PK::pack(o[0], o[1]);
}

// One work item: output tile (tk, tj), i-chunk ch.  red: 2 NW floats.
//
// Step one runs at plane p on E1 from the u queue; step M (2 <= M <= S) at
// plane p-(M-1)(R+1) on E1 less (M-1)R rows on each side, from the register
// queue of step M-1's values (its i neighbours and in-wave j neighbours) and
// an LDS plane of them (the other j and k neighbours), into a queue of its
// own; the last step writes u_{n+S-1} and u_{n+S} of the output tile.
template <int NPTS, int MJ, int NW, int S, int FMT>
__device__ __forceinline__ void sweep2h_tile(const Sweep2Args &a, const Coef2h<S> &K,
                                             const int tk, const int tj, const int ch,
                                             uint32_t *SM, float *red) {
  using Tm = Terms<NPTS>;
  using F = Fused2hShape<NPTS, MJ, NW, S>;
  constexpr int R = Tm::R;
  constexpr int EJ = F::EJ, OJ = F::OJ, OK = F::OK, KO = F::KO, KH = F::KH, LW = F::LW,
                SR = F::SR, WD = F::WD;
  constexpr int HJ = (2 * R + NW - 1) / NW;  // u halo rows per wave
  constexpr int Q = 2 * R + 2;               // u queue: planes p-R-1 .. p+R
  // Queue j (1 <= j < S) holds the lane's u_{n+j} words: glen(j) entries
  // (2R+2 for the last one, one more for the others: step j+2 takes its
  // -u_{n+j} term from the oldest entry).  GL: the array length.
  constexpr int GL = 2 * R + 3;
  // Steps per loop iteration: a multiple of the s ring and of 2 (the LDS
  // buffers alternate).
  constexpr int U = (S - 1) % 2 == 0 ? (S - 1) * (R + 1) : 2 * (S - 1) * (R + 1);
  static_assert(S >= 2 && S <= 5, "two to five steps per pass");
  static_assert(KO >= WD + R, "step one and its k-halo inside the lanes' columns");
  static_assert(KO % 2 == 0 && OK % 2 == 0, "a lane's pair is all output or none");
  static_assert(F::LDS_WORDS * 4 + 2 * NW * 4 <= 160 * 1024, "LDS");
  static_assert(U % SR == 0 && U % 2 == 0, "ring slots and buffers repeat with the loop");

  const int ib = ch * a.chunk_len;
  const int ie = min(a.xs, ib + a.chunk_len);
  if (ib >= ie) return;

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int je = tj * OJ - WD, ke = tk * OK - KO;  // E1's first row; the lanes' first column
  const int64_t P = a.plane, Pj = a.pitch;

  // The lane's pair starts at column kg (even).  Reads clamp to the pairs
  // that cover the allocated box [-R, zs+R): from the even column at or
  // below -R to the even column at or below zs+R-1 (its second element may
  // lie one past the box, inside the row padding or the next row's halo).
  // E1 points outside the interior take the older field's value (not a
  // stencil value), so what their clamped neighbours hold never reaches an
  // output.
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
  // wave-uniform: the row is inside the interior / an output row / one of
  // step M's rows (2 <= M < S: E1 less (M-1)R on each side)
  int rowin[MJ], rowout[MJ], rowst[S][MJ];
#pragma unroll
  for (int m = 0; m < MJ; m++) {
    const int er = wv * MJ + m, jg = je + er;
    off[m] = boff(jg);
    rowin[m] = __builtin_amdgcn_readfirstlane((int)(jg >= 0 && jg < a.ys));
    rowout[m] = __builtin_amdgcn_readfirstlane((int)(er >= WD && er < WD + OJ));
#pragma unroll
    for (int M = 0; M < S; M++)
      rowst[M][m] =
          __builtin_amdgcn_readfirstlane((int)(er >= (M - 1) * R && er < EJ - (M - 1) * R));
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
  //   lbW: u_{n+1} plane, row wv*MJ-R, word lane (offsets: plane 2(j-1) +
  //        buffer, row m+R+dr, word KH+dp);
  //   lbS: s ring slot 0, row wv*MJ, word lane (offsets: slot, row m).
  using LW32 = __attribute__((address_space(3))) uint32_t;
  constexpr uint32_t LUB = F::LU * 4, LPB = F::LP * 4, LSB = F::LS * 4;
  static_assert(2 * LUB <= 65536 && (2 * S - 3) * LPB + (MJ + 2 * R) * LW * 4 <= 65536 &&
                    (SR - 1) * LSB + MJ * 256 <= 65536,
                "offsets");
  const uint32_t lu0 = (uint32_t)(uintptr_t)(LW32 *)&SM[0];
  const uint32_t wbase = lu0 + 2 * LUB, sbase = wbase + 2 * (S - 1) * LPB;
  uint32_t lbL = lu0 + (uint32_t)((wv * MJ * LW + lane) * 4);
  uint32_t lbW = wbase + (uint32_t)(((wv * MJ - R) * LW + lane) * 4);
  uint32_t lbS = sbase + (uint32_t)((wv * MJ * 64 + lane) * 4);
  uint32_t hadr[HJ];
#pragma unroll
  for (int h = 0; h < HJ; h++) hadr[h] = lu0 + (uint32_t)((max(hrow[h], 0) * LW + KH + lane) * 4);
  auto lds = [](uint32_t adr) -> LW32 & { return *(LW32 *)(uintptr_t)adr; };

  // Register queues of packed words:
  //   q[m][t]    = u_n(p - R - 1 + t), t = 0 .. 2R+1;
  //   G[j][m][t] = u_{n+j}(p_j - L_j + 1 + t), t = 0 .. L_j-2, with p_j =
  //                p-(j-1)(R+1) the plane of step j and L_j = glen(j) (t =
  //                L_j-1 takes step j's result at p_j before the shift).  Step
  //                j+1 reads entries gcen(j) -+ r (its plane p_{j+1} and i
  //                neighbours), step j+2 entry 0 (u_{n+j} at p_{j+2}).
  uint32_t q[MJ][Q], G[S][MJ][GL];
  uint32_t hq[HJ][R + 1];
  // Step one from p0 (its planes past ie-1+WD are never used), the last
  // step (output plane p-(S-1)(R+1)) up to plast.
  const int p0 = ib - WD, plast = ie + (S - 1) * (R + 1) - 1;
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
    for (int j = 1; j < S; j++)
#pragma unroll
      for (int t = 0; t < glen<S, R>(j); t++) G[j][m][t] = 0u;
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

  // Plane pointers advanced by one plane per step, each held inside
  // [-R, xs+R): the prefetch of step p reads u(p+2+R), v(p+2), s(p+2); the
  // last step stores plane p-(S-1)(R+1), from plane ib on.  The first step
  // of a chunk is WD planes below it, so the lower bound only binds for
  // WD > R+2 (v, s) and WD > 2R+2 (u); elsewhere the test folds away.
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
    uint32_t vv[MJ], s1[MJ], sc[S + 1][MJ];
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      q[m][Q - 1] = n.u[m];
      s1[m] = n.s[m];
      vv[m] = n.v[m];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++) hq[h][R] = n.hj[h];
    prefetch(Un, Vn, Sn, n);
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
        sc[M][m] = lds(lbS + (uint32_t)(((k + (S - M) * (R + 1)) % SR) * LSB) + m * 256);
      lds(lbS + (uint32_t)((k % SR) * LSB) + m * 256) = s1[m];
    }
    constexpr uint32_t LO = (k & 1) * LUB;
    auto Lat = [&](int m, int dr, int dp) -> LW32 & {
      return lds(lbL + LO + (uint32_t)(((m + R + dr) * LW + KH + dp) * 4));
    };
    // this step's buffer of u_{n+j}
    auto Wat = [&](int j, int m, int dr, int dp) -> LW32 & {
      return lds(lbW + (uint32_t)((2 * (j - 1) + (k & 1)) * LPB) +
                 (uint32_t)(((m + R + dr) * LW + KH + dp) * 4));
    };
    // u(p) into LDS (the lanes' rows and the j-halo rows); u_{n+j} at the
    // plane of step j+1 (p-j(R+1)), j = 1 .. S-1.
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      Lat(m, 0, 0) = q[m][R + 1];
#pragma unroll
      for (int j = 1; j < S; j++) Wat(j, m, 0, 0) = G[j][m][gcen<S, R>(j)];
    }
#pragma unroll
    for (int h = 0; h < HJ; h++)
      if (hrow[h] >= 0) lds(hadr[h] + LO) = hq[h][0];
    __syncthreads();

    // u_{n+M-2} at step M's plane: u_n from the u queue, else the oldest
    // entry of queue M-2.
    auto older = [&](auto jc, int m) __attribute__((always_inline)) -> uint32_t {
      constexpr int J = decltype(jc)::value;
      if constexpr (J == 0)
        return q[m][0];
      else
        return G[J][m][0];
    };
    // The Laplacian of u_{n+J} at row m, both columns: queue J (i neighbours
    // and in-wave j neighbours) and its LDS plane (the others).
    auto lapq = [&](auto jc, int m, float &L0, float &L1) __attribute__((always_inline)) {
      constexpr int J = decltype(jc)::value, C = gcen<S, R>(J);
      using W2 = uint32_t __attribute__((ext_vector_type(2)));
      const uint32_t cw = G[J][m][C];
      lap_pair<NPTS, KH, FMT>(
          a.w,
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{G[J][m][C - r], G[J][m][C + r]};
          },
          [&](auto rc) -> W2 {
            constexpr int r = decltype(rc)::value;
            return W2{m - r >= 0 ? G[J][m - r >= 0 ? m - r : 0][C] : (uint32_t)Wat(J, m, -r, 0),
                      m + r < MJ ? G[J][m + r < MJ ? m + r : 0][C] : (uint32_t)Wat(J, m, r, 0)};
          },
          [&](auto dc) -> uint32_t {
            constexpr int d = decltype(dc)::value;
            return d == 0 ? cw : (uint32_t)Wat(J, m, 0, d);
          },
          L0, L1);
    };

    // The last step writes plane io = p-(S-1)(R+1) of the output tile: the
    // two newest fields there.  Rows without output points are skipped
    // (uniform).
    const int io = p - (S - 1) * (R + 1);
    const bool store = io >= ib && io < ie;
    const BufRes Wr = buf_res(Wo), Zr = buf_res(Zo);
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      if (!rowout[m]) continue;
      constexpr int J = S - 1;
      const uint32_t cw = G[J][m][gcen<S, R>(J)];
      float L0, L1;
      lapq(std::integral_constant<int, J>{}, m, L0, L1);
      const uint32_t zw =
          update_pair<FMT>(K.k0[S - 1], K.k1[S - 1], K.k2[S - 1], K.kv[S - 1], cw,
                           older(std::integral_constant<int, J - 1>{}, m), sc[S][m], L0, L1);
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

    // Steps S-1 .. 2, each into the newest entry of its queue; halves
    // outside the interior take u_{n+M-2}'s value there.  (Only
    // instantiated for S >= 3.)
    if constexpr (S >= 3) {
      static_for<S - 2>([&](auto ic) __attribute__((always_inline)) {
        constexpr int M = S - 1 - decltype(ic)::value;  // S-1 .. 2
        const int pm = p - (M - 1) * (R + 1);
        const bool pinm = pm >= 0 && pm < a.xs;
#pragma unroll
        for (int m = 0; m < MJ; m++) {
          const uint32_t old = older(std::integral_constant<int, M - 2>{}, m);
          if (!rowst[M][m]) {
            G[M][m][glen<S, R>(M) - 1] = old;
            continue;
          }
          float L0, L1;
          lapq(std::integral_constant<int, M - 1>{}, m, L0, L1);
          const uint32_t zw =
              update_pair<FMT>(K.k0[M - 1], K.k1[M - 1], K.k2[M - 1], K.kv[M - 1],
                               G[M - 1][m][gcen<S, R>(M - 1)], old, sc[M][m], L0, L1);
          G[M][m][glen<S, R>(M) - 1] =
              (rowin[m] && pinm) ? ((zw & kinmask) | (old & ~kinmask)) : old;
        }
      });
    }

    // Step one at plane p: u_{n+1} on E1, rounded to binary16, into the
    // newest queue entry; points outside the interior take v's value.
    const bool pin = p >= 0 && p < a.xs;
#pragma unroll
    for (int m = 0; m < MJ; m++) {
      const uint32_t cw = q[m][R + 1];
      using W2 = uint32_t __attribute__((ext_vector_type(2)));
      float L0, L1;
      lap_pair<NPTS, KH, FMT>(
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
      const uint32_t ww =
          update_pair<FMT>(K.k0[0], K.k1[0], K.k2[0], K.kv[0], cw, vv[m], s1[m], L0, L1);
      // halves inside the interior take the new value, the others v's
      G[1][m][glen<S, R>(1) - 1] =
          (rowin[m] && pin) ? ((ww & kinmask) | (vv[m] & ~kinmask)) : vv[m];
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

  // Maxima of the item: the larger half, widened, reduced over the
  // workgroup, one atomicMax per output field.
  auto hmax = [](uint32_t x) { return Pk16<FMT>::template wide<0>(max(x & 0xffffu, x >> 16)); };
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
    // u_{n+S-1} (o1) and u_{n+S} (o2): the one-step scheme writes u_{n+j}
    // into v's buffer for odd j, into u's for even j
    Q16Slot &nx = a.qs[(a.slot + 1) % 3];
    atomicMax(&nx.maxbits[(S - 1) % 2 ? a.vbuf : a.ubuf], __float_as_uint(b1));
    atomicMax(&nx.maxbits[S % 2 ? a.vbuf : a.ubuf], __float_as_uint(b2));
  }
}

template <int NPTS, int MJ, int NW, int S, int FMT>
__global__ __launch_bounds__(64 * NW) void sweep2h_kernel(const Sweep2Args a) {
  using F = Fused2hShape<NPTS, MJ, NW, S>;
  __shared__ __attribute__((aligned(16))) uint32_t SM[F::LDS_WORDS];
  __shared__ float red[2 * NW];
  const Q16Slot &cur = a.qs[a.slot];
  const float c2[5] = {a.c2a, a.c2b, a.c2c, a.c2d, a.c2e};
  int e[S];
  q16_pass_exps<S>(cur, a.ubuf, a.vbuf, a.c0, a.c1, c2, a.gain, e);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // The next slot: u_{n+S-1} and u_{n+S} with their exponents; their
    // maxima are reduced by the blocks into it (cleared by the previous
    // launch); the slot after it is cleared for the next launch.
    Q16Slot &nxt = a.qs[(a.slot + 1) % 3];
    Q16Slot &clr = a.qs[(a.slot + 2) % 3];
    nxt.exp[(S - 1) % 2 ? a.vbuf : a.ubuf] = e[S - 2];
    nxt.exp[S % 2 ? a.vbuf : a.ubuf] = e[S - 1];
    nxt.exp[2] = cur.exp[2];
    nxt.maxbits[2] = cur.maxbits[2];
    clr.maxbits[0] = clr.maxbits[1] = clr.maxbits[2] = 0u;
  }
  // step M's inputs: u_{n+M-1} (exponent ein[M]) and u_{n+M-2} (ein[M-1])
  const int es = cur.exp[2];
  int ein[S + 1];
  ein[0] = cur.exp[a.vbuf];
  ein[1] = cur.exp[a.ubuf];
#pragma unroll
  for (int M = 2; M <= S; M++) ein[M] = e[M - 2];
  // the coefficients are the same in every lane: keep them in scalar registers
  auto uf = [](float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
  };
  Coef2h<S> K;
#pragma unroll
  for (int M = 1; M <= S; M++) {
    const int eo = e[M - 1];
    K.k0[M - 1] = uf(ldexpf(a.c0, ein[M] - eo));
    K.k1[M - 1] = uf(ldexpf(a.c1, ein[M] - eo));
    K.k2[M - 1] = uf(ldexpf(c2[M - 1], es - eo));
    K.kv[M - 1] = uf(ldexpf(1.0f, ein[M - 1] - eo));
  }
  int bid = blockIdx.x;
  if (a.remap) bid = xcd_remap(bid, (int)gridDim.x);
  const int tk = bid % a.ntk;
  const int tt = bid / a.ntk;
  sweep2h_tile<NPTS, MJ, NW, S, FMT>(a, K, tk, tt % a.ntj, tt / a.ntj, SM, red);
}

}  // namespace sth

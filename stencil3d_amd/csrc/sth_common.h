// sth_common.h -- geometry, device layout and stencil constants shared by the
// host side and the gfx950 kernels of libstencil_hip.
#pragma once

#include <stdint.h>

#include "stencil_hip.h"

namespace sth {

constexpr int kXcds = 8;          // MI355X: 8 XCDs, 32 CUs each
constexpr int kRowAlignBytes = 256;

inline int radius_of(uint32_t npts) {
  return npts == 7 ? 1 : npts == 13 ? 2 : npts == 25 ? 4 : -1;
}
inline int elem_bytes_of(uint32_t dtype) {
  return dtype == STH_FP32 ? 4 : dtype == STH_FP64 ? 8
       : (dtype == STH_FP16 || dtype == STH_BF16) ? 2 : -1;
}

// Device layout: halo = R on every side; rows padded to a multiple of 256 B
// and shifted so that interior k = 0 of every row is 256-B aligned (the
// reference aligns whole arrays to 64 KiB and staggers them, main.cpp:244-275;
// on the GPU what matters is that each wave's 64-lane row segment maps onto
// whole 128-B cache lines).
inline int compute_layout(const sth_geom &g, sth_layout_info &L) {
  const int R = radius_of(g.npts);
  const int eb = elem_bytes_of(g.dtype);
  if (R < 0 || eb < 0) return STH_ERR_ARG;
  if (g.xs < 1 || g.ys < 1 || g.zs < 1 || g.xs > 65535 || g.ys > 65535 || g.zs > 65535)
    return STH_ERR_ARG;
  const uint64_t align = kRowAlignBytes / eb;  // elements per 256 B
  L.radius = (uint32_t)R;
  L.elem_bytes = (uint32_t)eb;
  L.pitch = ((uint64_t)g.zs + 2 * R + align - 1) / align * align;
  L.plane = L.pitch * ((uint64_t)g.ys + 2 * R);
  // The kernels address a plane with 32-bit lane byte offsets (and use
  // offsets from 2^31 up as "no load"): a plane must stay below 2 GiB.
  if (L.plane * (uint64_t)eb >= (1ull << 31)) return STH_ERR_ARG;
  L.origin = align + (uint64_t)R * L.pitch + (uint64_t)R * L.plane;
  L.elems = L.plane * ((uint64_t)g.xs + 2 * R) + align;
  L.bytes = (L.elems * eb + 255) / 256 * 256;
  return STH_OK;
}

// Stencil constants of stencil_7.h:4-8, stencil_13.h:13-22 and
// stencil_25.h:19-34 (non-ISOTROPIC), rounded to FP32 exactly as the
// reference forms them (cy = cx*1.1f, cz = cx*0.9f in float).
struct Coeffs {
  float coef0, A;
  float w[3][4];  // [axis i,j,k][r-1]
};

inline int stencil_coeffs(uint32_t npts, Coeffs &c) {
  for (int a = 0; a < 3; a++)
    for (int r = 0; r < 4; r++) c.w[a][r] = 0.0f;
  float cx[4] = {0, 0, 0, 0};
  int R;
  if (npts == 7) {
    c.coef0 = 1.5f;
    c.A = 3.5f;  // stencil_7.h:38
    c.w[0][0] = c.w[1][0] = c.w[2][0] = 0.5f;
    return 1;
  } else if (npts == 13) {
    c.coef0 = 1.5f;
    c.A = 3.5f;  // stencil_13.h:60
    cx[0] = 0.3f;
    cx[1] = 0.2f;
    R = 2;
  } else if (npts == 25) {
    c.coef0 = 1.25f;
    c.A = 6.25f;  // stencil_25.h:82
    cx[0] = 0.4f;
    cx[1] = 0.3f;
    cx[2] = 0.2f;
    cx[3] = 0.1f;
    R = 4;
  } else {
    return -1;
  }
  volatile float f11 = 1.1f, f09 = 0.9f;  // products rounded to FP32 on the host
  for (int r = 0; r < R; r++) {
    c.w[0][r] = cx[r];
    c.w[1][r] = cx[r] * f11;
    c.w[2][r] = cx[r] * f09;
  }
  return R;
}

// The same weights as the reference's -DUSE_DOUBLE build forms them
// (config.h:96-100 makes `real` double): the float literals are promoted
// (cx1 = 0.4f, stencil_25.h:22), and cy = cx*1.1f, cz = cx*0.9f are products
// rounded in double (stencil_13.h:18-21, stencil_25.h:26-33).  The 7-point
// weights are the literal 0.5 (stencil_7.h:4-8).
inline int stencil_coeffs_double(uint32_t npts, double w[3][4]) {
  Coeffs c;
  const int R = stencil_coeffs(npts, c);
  if (R < 0) return -1;
  for (int a = 0; a < 3; a++)
    for (int r = 0; r < 4; r++) w[a][r] = (double)c.w[a][r];
  if (npts == 7) return R;
  volatile double f11 = (double)1.1f, f09 = (double)0.9f;
  for (int r = 0; r < R; r++) {
    const double cx = (double)c.w[0][r];  // the literal itself
    w[1][r] = cx * f11;
    w[2][r] = cx * f09;
  }
  return R;
}

// Evaluation order of the neighbour pairs, as the reference's expressions
// are written (stencil_7.h:20; stencil_13.h:40-43; stencil_25.h:64-67):
// 7-point (i,1),(j,1),(k,1); 13/25-point k(R..1), j(R..1), i(R..1).
template <int NPTS>
struct Terms;
template <>
struct Terms<7> {
  static constexpr int R = 1, T = 3;
  static constexpr int axis[T] = {0, 1, 2};
  static constexpr int r[T] = {1, 1, 1};
};
template <>
struct Terms<13> {
  static constexpr int R = 2, T = 6;
  static constexpr int axis[T] = {2, 2, 1, 1, 0, 0};
  static constexpr int r[T] = {2, 1, 2, 1, 2, 1};
};
template <>
struct Terms<25> {
  static constexpr int R = 4, T = 12;
  static constexpr int axis[T] = {2, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0, 0};
  static constexpr int r[T] = {4, 3, 2, 1, 4, 3, 2, 1, 4, 3, 2, 1};
};

inline int term_order(uint32_t npts, int axis[12], int r[12]) {
  if (npts == 7) {
    for (int t = 0; t < 3; t++) { axis[t] = Terms<7>::axis[t]; r[t] = Terms<7>::r[t]; }
    return 3;
  }
  if (npts == 13) {
    for (int t = 0; t < 6; t++) { axis[t] = Terms<13>::axis[t]; r[t] = Terms<13>::r[t]; }
    return 6;
  }
  if (npts == 25) {
    for (int t = 0; t < 12; t++) { axis[t] = Terms<25>::axis[t]; r[t] = Terms<25>::r[t]; }
    return 12;
  }
  return -1;
}

// Constant-mode gain of the leapfrog neighbour sum: sum over terms of
// 2*|wh|, wh = 2*w (so the 7-point weights are 1 and the gain is 6).
inline double lap_gain(uint32_t npts) {
  Coeffs c;
  int ax[12], rr[12];
  if (stencil_coeffs(npts, c) < 0) return -1.0;
  const int T = term_order(npts, ax, rr);
  double S = 0.0;
  for (int t = 0; t < T; t++) {
    const double wh = 2.0 * (double)c.w[ax[t]][rr[t] - 1];
    S += 2.0 * (wh < 0 ? -wh : wh);
  }
  return S;
}

}  // namespace sth

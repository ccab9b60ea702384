"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the compiled
reference (oracle/_ref/), for tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only.

TEST INFRASTRUCTURE ONLY: the product (stencil3d_amd, libstencil_hip) never
imports this module.  Arrays are numpy arrays in the reference's dense padded
layout (grid.cpp:20-25): shape (xs+2xh, ys+2yh, zs+2zh), C order, k fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REFERENCE_SRC = "/root/reference"


class Geom(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("xs", "ys", "zs", "xh", "yh", "zh")]


@dataclass(frozen=True)
class Grid:
    """Interior extents (i slowest, k unit stride) and halo widths."""

    xs: int
    ys: int
    zs: int
    xh: int
    yh: int
    zh: int

    @property
    def shape(self):
        return (self.xs + 2 * self.xh, self.ys + 2 * self.yh, self.zs + 2 * self.zh)

    @property
    def interior(self):
        return (slice(self.xh, self.xh + self.xs), slice(self.yh, self.yh + self.ys),
                slice(self.zh, self.zh + self.zs))

    def c(self):
        return Geom(self.xs, self.ys, self.zs, self.xh, self.yh, self.zh)

    def zeros(self, dtype=np.float32):
        return np.zeros(self.shape, dtype=dtype)


def build(quiet=True):
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, stdout=out)
    if os.path.isdir(REFERENCE_SRC):
        subprocess.run(["make", "-C", HERE, "-j8", "ref"], check=True, stdout=out)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        G = P(Geom)
        f32p, f64p, u16p = P(C.c_float), P(C.c_double), P(C.c_uint16)
        sig = {
            "orc_coeffs": (C.c_int, [C.c_int, P(C.c_float), P(C.c_float), f32p]),
            "orc_terms": (C.c_int, [C.c_int, P(C.c_int), P(C.c_int)]),
            "orc_ref_sweep_f32": (C.c_int, [C.c_int, G, f32p, f32p, f32p, C.c_int, C.c_int]),
            "orc_ref_sweep_f64": (C.c_int, [C.c_int, G, f64p, f64p, f64p, C.c_int, C.c_int,
                                            C.c_int]),
            "orc_leapfrog_f32": (C.c_int, [C.c_int, G, f32p, f32p, f32p, C.c_float, C.c_float,
                                           f32p, C.c_int, C.c_int]),
            "orc_leapfrog_f64": (C.c_int, [C.c_int, G, f64p, f64p, f64p, C.c_double, C.c_double,
                                           f64p, C.c_int, C.c_int]),
            "orc_lap_gain": (C.c_double, [C.c_int]),
            "orc_f32_to_16": (C.c_uint16, [C.c_int, C.c_float]),
            "orc_16_to_f32": (C.c_float, [C.c_int, C.c_uint16]),
            "orc_scale_exp": (C.c_int, [C.c_double]),
            "orc_quantize": (C.c_float, [C.c_int, G, f32p, u16p, C.c_int]),
            "orc_dequantize": (None, [C.c_int, G, u16p, C.c_int, f32p]),
            "orc_leapfrog_q16": (C.c_int, [C.c_int, C.c_int, G, u16p, u16p, u16p, P(C.c_int),
                                           f32p, C.c_float, C.c_float, f32p, C.c_int, C.c_int]),
            "orc_leapfrog_q16_pass2": (C.c_int, [C.c_int, C.c_int, G, u16p, u16p, u16p,
                                                 P(C.c_int), f32p, C.c_float, C.c_float, f32p,
                                                 C.c_int, C.c_int]),
            "orc_leapfrog_q16_passes": (C.c_int, [C.c_int, C.c_int, G, u16p, u16p, u16p,
                                                  P(C.c_int), f32p, C.c_float, C.c_float, f32p,
                                                  C.c_int, C.c_int, C.c_int]),
            "orc_splitmix64_mix": (C.c_uint64, [C.c_uint64]),
            "orc_uniform01": (C.c_double, [C.c_uint64, C.c_int, C.c_int64, C.c_int64, C.c_int64]),
            "orc_fill_uniform": (None, [G, f32p, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                        C.c_float]),
            "orc_fill_uniform_f64": (None, [G, f64p, C.c_uint64, C.c_int, C.c_double,
                                            C.c_double, C.c_double]),
            "orc_fill_gaussian": (None, [G, f32p, C.c_double, C.c_double, C.c_double,
                                         C.c_uint64, C.c_int]),
            "orc_fill_gaussian_f64": (None, [G, f64p, C.c_double, C.c_double, C.c_double,
                                             C.c_uint64, C.c_int]),
            "orc_ricker": (C.c_double, [C.c_double, C.c_double]),
            "orc_init_arrays": (None, [f32p, f32p, f32p, C.c_int64]),
            "orc_check": (C.c_int64, [G, f32p, C.c_float]),
            "orc_norms_f32": (None, [G, f32p, P(C.c_double), P(C.c_double)]),
            "orc_norms_f64": (None, [G, f64p, P(C.c_double), P(C.c_double)]),
            "orc_error_f32_vs_f64": (None, [G, f32p, f64p, P(C.c_double), P(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ct))


def f32p(a):
    assert a.dtype == np.float32
    return _p(a, C.c_float)


def f64p(a):
    assert a.dtype == np.float64
    return _p(a, C.c_double)


def u16p(a):
    assert a.dtype == np.uint16
    return _p(a, C.c_uint16)


def coeffs(npts):
    c0, A, w = C.c_float(), C.c_float(), (C.c_float * 12)()
    R = lib().orc_coeffs(npts, C.byref(c0), C.byref(A), w)
    if R < 0:
        raise ValueError(f"unsupported stencil size {npts}")
    return c0.value, A.value, np.array(list(w), dtype=np.float32).reshape(3, 4), R


def terms(npts):
    ax, r = (C.c_int * 12)(), (C.c_int * 12)()
    T = lib().orc_terms(npts, ax, r)
    return [(ax[t], r[t]) for t in range(T)]


def radius(npts):
    return {7: 1, 13: 2, 25: 4}[npts]


def ref_sweep_f32(npts, g: Grid, v, s, u, iterations, nthreads=0):
    rc = lib().orc_ref_sweep_f32(npts, C.byref(g.c()), f32p(v), f32p(s), f32p(u), iterations,
                                 nthreads)
    assert rc == 0


def ref_sweep_f64(npts, g: Grid, v, s, u, iterations, coeff_mode=0, nthreads=0):
    rc = lib().orc_ref_sweep_f64(npts, C.byref(g.c()), f64p(v), f64p(s), f64p(u), iterations,
                                 coeff_mode, nthreads)
    assert rc == 0


def leapfrog_f32(npts, g: Grid, a, b, s, c0, c1, c2, steps, nthreads=0):
    """Returns the index (0: a, 1: b) of the buffer holding the newest field."""
    c2a = None if c2 is None else np.ascontiguousarray(c2, dtype=np.float32)
    rc = lib().orc_leapfrog_f32(npts, C.byref(g.c()), f32p(a), f32p(b), f32p(s), c0, c1,
                                None if c2a is None else f32p(c2a), steps, nthreads)
    assert rc >= 0
    return rc


def leapfrog_f64(npts, g: Grid, a, b, s, c0, c1, c2, steps, nthreads=0):
    c2a = None if c2 is None else np.ascontiguousarray(c2, dtype=np.float64)
    rc = lib().orc_leapfrog_f64(npts, C.byref(g.c()), f64p(a), f64p(b), f64p(s), c0, c1,
                                None if c2a is None else f64p(c2a), steps, nthreads)
    assert rc >= 0
    return rc


def lap_gain(npts):
    return lib().orc_lap_gain(npts)


def to16(fmt, x: float) -> int:
    return lib().orc_f32_to_16(fmt, x)


def from16(fmt, h: int) -> float:
    return lib().orc_16_to_f32(fmt, h)


def scale_exp(B: float) -> int:
    return lib().orc_scale_exp(B)


def quantize(fmt, g: Grid, x, e):
    q = np.zeros(g.shape, dtype=np.uint16)
    mx = lib().orc_quantize(fmt, C.byref(g.c()), f32p(x), u16p(q), e)
    return q, mx


def dequantize(fmt, g: Grid, q, e):
    x = np.zeros(g.shape, dtype=np.float32)
    lib().orc_dequantize(fmt, C.byref(g.c()), u16p(q), e, f32p(x))
    return x


def leapfrog_q16(npts, fmt, g: Grid, a, b, s, exps, maxs, c0, c1, c2, steps, nthreads=0,
                 pass2=False, plen=None):
    """exps/maxs: lists of 3 (a, b, s), updated in place.  Returns newest index.
    pass2: the steps in passes of two (orc_leapfrog_q16_pass2); plen: in
    passes of plen steps (orc_leapfrog_q16_passes)."""
    ce = (C.c_int * 3)(*exps)
    cm = (C.c_float * 3)(*maxs)
    c2a = None if c2 is None else np.ascontiguousarray(c2, dtype=np.float32)
    args = (npts, fmt, C.byref(g.c()), u16p(a), u16p(b), u16p(s), ce, cm,
            c0, c1, None if c2a is None else f32p(c2a), steps)
    if plen is not None:
        rc = lib().orc_leapfrog_q16_passes(*args, plen, nthreads)
    else:
        fn = lib().orc_leapfrog_q16_pass2 if pass2 else lib().orc_leapfrog_q16
        rc = fn(*args, nthreads)
    assert rc >= 0
    exps[:] = list(ce)
    maxs[:] = list(cm)
    return rc


def splitmix64_mix(z: int) -> int:
    return lib().orc_splitmix64_mix(z)


def uniform01(seed, field, i, j, k):
    return lib().orc_uniform01(seed, field, i, j, k)


def fill_uniform(g: Grid, seed, field, lo=-1.0, hi=1.0, halo=0.0, dtype=np.float32):
    a = g.zeros(dtype)
    if dtype == np.float32:
        lib().orc_fill_uniform(C.byref(g.c()), f32p(a), seed, field, lo, hi, halo)
    else:
        lib().orc_fill_uniform_f64(C.byref(g.c()), f64p(a), seed, field, lo, hi, halo)
    return a


def fill_gaussian(g: Grid, amp, sigma, noise, seed, field, dtype=np.float32):
    a = g.zeros(dtype)
    if dtype == np.float32:
        lib().orc_fill_gaussian(C.byref(g.c()), f32p(a), amp, sigma, noise, seed, field)
    else:
        lib().orc_fill_gaussian_f64(C.byref(g.c()), f64p(a), amp, sigma, noise, seed, field)
    return a


def ricker(f, tau):
    return lib().orc_ricker(f, tau)


def init_arrays(g: Grid):
    v, s, u = g.zeros(), g.zeros(), g.zeros()
    lib().orc_init_arrays(f32p(v), f32p(s), f32p(u), v.size)
    return v, s, u


def check(g: Grid, v, expected):
    return lib().orc_check(C.byref(g.c()), f32p(v), expected)


def norms(g: Grid, x):
    mx, l2 = C.c_double(), C.c_double()
    if x.dtype == np.float32:
        lib().orc_norms_f32(C.byref(g.c()), f32p(x), C.byref(mx), C.byref(l2))
    else:
        lib().orc_norms_f64(C.byref(g.c()), f64p(x), C.byref(mx), C.byref(l2))
    return mx.value, l2.value


def error_vs_f64(g: Grid, x, r):
    mr, lr = C.c_double(), C.c_double()
    lib().orc_error_f32_vs_f64(C.byref(g.c()), f32p(x), f64p(r), C.byref(mr), C.byref(lr))
    return mr.value, lr.value


# ---------------------------------------------------------------- reference
def ref_available():
    return os.path.exists(os.path.join(REF_DIR, "libref7_3d.so"))


_ref_libs = {}


def ref_lib(npts, kind="3d"):
    key = (npts, kind)
    if key not in _ref_libs:
        path = os.path.join(REF_DIR, f"libref{npts}_{kind}.so")
        L = C.CDLL(path)
        if kind.endswith("f64"):
            fn = L.ref_shim_iterate_f64
            fn.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_uint] * 8
        else:
            fn = L.ref_shim_iterate
            fn.argtypes = [C.POINTER(C.c_float)] * 3 + [C.c_uint] * 8
        fn.restype = C.c_int
        _ref_libs[key] = fn
    return _ref_libs[key]


def ref_iterate(npts, g: Grid, v, s, u, iterations, kind="3d", zb=8):
    """Runs the compiled reference's ITERATE on the arrays (in place on v).
    kind '3d' = iterate_3d (valid only for xs == ys, iterators.cpp:146),
    '2d' = iterate_2d_jk_blocked (-DPAR2D), '3d_f64' = the USE_DOUBLE build."""
    fn = ref_lib(npts, kind)
    conv = f64p if kind.endswith("f64") else f32p
    rc = fn(conv(v), conv(s), conv(u), g.xs, g.ys, g.zs, g.xh, g.yh, g.zh, min(zb, g.zs),
            iterations)
    assert rc == 0


def ref_binary(npts):
    return os.path.join(REF_DIR, f"hstencil3d_{npts}")

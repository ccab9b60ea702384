"""stencil3d_amd -- Python binding of libstencil_hip, the MI355X (gfx950) sweep
library of stencil3d.

The library is the product; this module only loads it with ctypes and moves
numpy arrays in and out.  It mirrors the reference's operator interface
(iterators.h:6-35): ``init_arrays``, ``iterate`` (ITERATE), ``check``, plus the
leapfrog and 16-bit storage forms.  Arrays use the reference's dense padded
layout (grid.cpp:20-25): shape (xs+2xh, ys+2yh, zs+2zh), C order, k fastest.

There is no CPU fallback: if libstencil_hip.so is missing, or no GPU is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libstencil_hip.so")
FRONTEND_PATH = os.path.join(PKG, "hstencil3d_hip")

STH_OK, STH_ERR_ARG, STH_ERR_HIP, STH_ERR_NOMEM, STH_ERR_STATE, STH_ERR_NODEV, \
    STH_ERR_UNSUPPORTED = range(7)
DTYPES = {"fp32": 0, "fp16": 1, "bf16": 2, "fp64": 3}
WEIGHTS = {"fp32": 0, "reference": 1}  # sth_weights (FP64 storage)
FIELDS = {"v": 0, "s": 1, "u": 2}


class SthError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {_lib_or_none_strerror(status)} (status {status})")


class Geom(C.Structure):
    _fields_ = [("xs", C.c_uint32), ("ys", C.c_uint32), ("zs", C.c_uint32),
                ("npts", C.c_uint32), ("dtype", C.c_uint32)]


class LayoutInfo(C.Structure):
    _fields_ = [("radius", C.c_uint32), ("elem_bytes", C.c_uint32), ("pitch", C.c_uint64),
                ("plane", C.c_uint64), ("origin", C.c_uint64), ("elems", C.c_uint64),
                ("bytes", C.c_uint64)]


class HostLayout(C.Structure):
    _fields_ = [("xh", C.c_uint32), ("yh", C.c_uint32), ("zh", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("rows_per_wave", C.c_uint32), ("waves", C.c_uint32), ("chunks", C.c_uint32),
                ("xcd_remap", C.c_uint32), ("vec", C.c_uint32), ("steps_per_pass", C.c_uint32),
                ("launch", C.c_uint32), ("ref_sweeps", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol include/stencil_hip.h declares, with its ctypes signature.
_P = C.POINTER
_VP = C.c_void_p
SIGNATURES = {
    "sth_strerror": (C.c_char_p, [C.c_int]),
    "sth_version": (C.c_int, []),
    "sth_layout": (C.c_int, [_P(Geom), _P(LayoutInfo)]),
    "sth_host_elems": (C.c_int, [_P(Geom), _P(HostLayout), _P(C.c_uint64)]),
    "sth_device_count": (C.c_int, [_P(C.c_int)]),
    "sth_create": (C.c_int, [_P(Geom), _P(_VP)]),
    "sth_destroy": (C.c_int, [_VP]),
    "sth_get_layout": (C.c_int, [_VP, _P(LayoutInfo)]),
    "sth_init_const": (C.c_int, [_VP, C.c_float, C.c_float, C.c_float]),
    "sth_init_uniform": (C.c_int, [_VP, C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                   C.c_float]),
    "sth_init_gaussian": (C.c_int, [_VP, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_uint64, C.c_int]),
    "sth_init_point": (C.c_int, [_VP, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float]),
    "sth_upload": (C.c_int, [_VP, C.c_int, _P(C.c_float), _P(HostLayout)]),
    "sth_download": (C.c_int, [_VP, C.c_int, _P(C.c_float), _P(HostLayout)]),
    "sth_upload_f64": (C.c_int, [_VP, C.c_int, _P(C.c_double), _P(HostLayout)]),
    "sth_download_f64": (C.c_int, [_VP, C.c_int, _P(C.c_double), _P(HostLayout)]),
    "sth_upload_raw16": (C.c_int, [_VP, C.c_int, _P(C.c_uint16), C.c_int, _P(HostLayout)]),
    "sth_download_raw16": (C.c_int, [_VP, C.c_int, _P(C.c_uint16), _P(C.c_int),
                                     _P(C.c_float), _P(HostLayout)]),
    "sth_save": (C.c_int, [_VP, C.c_char_p]),
    "sth_load": (C.c_int, [_VP, C.c_char_p]),
    "sth_iterate": (C.c_int, [_VP, C.c_uint32]),
    "sth_leapfrog": (C.c_int, [_VP, C.c_uint32, C.c_float, C.c_float, _P(C.c_float)]),
    "sth_leapfrog_noalloc": (C.c_int, [_P(Geom), _VP, _VP, _VP, C.c_uint32, C.c_float,
                                       C.c_float, _P(C.c_float), _VP]),
    "sth_leapfrog_c0": (C.c_int, [C.c_uint32, C.c_float, _P(C.c_float)]),
    "sth_ricker": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
                             _P(C.c_float)]),
    "sth_sync": (C.c_int, [_VP]),
    "sth_check": (C.c_int, [_VP, C.c_float, _P(C.c_uint64)]),
    "sth_norms": (C.c_int, [_VP, C.c_int, _P(C.c_double), _P(C.c_double)]),
    "sth_compare": (C.c_int, [_VP, C.c_int, _VP, C.c_int, _P(C.c_double), _P(C.c_double)]),
    "sth_energy": (C.c_int, [_VP, C.c_float, C.c_float, _P(C.c_double)]),
    "sth_set_weights": (C.c_int, [_VP, C.c_int]),
    "sth_get_weights": (C.c_int, [_VP, _P(C.c_int)]),
    "sth_set_tuning": (C.c_int, [_VP, _P(Tuning)]),
    "sth_get_tuning": (C.c_int, [_VP, _P(Tuning)]),
    "sth_autotune": (C.c_int, [_VP, C.c_int, C.c_uint32, _P(Tuning), _P(C.c_float)]),
    "sth_timer_start": (C.c_int, [_VP]),
    "sth_timer_read": (C.c_int, [_VP, _P(C.c_double), _P(C.c_uint64)]),
    "sth_stream_copy": (C.c_int, [_VP, C.c_uint64, C.c_uint32, _P(C.c_double)]),
    "sth_bw_probe": (C.c_int, [_VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                               C.c_uint32, _P(C.c_double)]),
    "sth_place": (C.c_int, [_VP, C.c_uint32, _P(C.c_float), _P(C.c_uint32)]),
    "sth_batch_create": (C.c_int, [_P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32,
                                   _P(_VP)]),
    "sth_batch_destroy": (C.c_int, [_VP]),
    "sth_batch_shot": (_VP, [_VP, C.c_uint32]),
    "sth_batch_size": (C.c_uint32, [_VP]),
    "sth_batch_set_tuning": (C.c_int, [_VP, _P(Tuning), C.c_uint32, C.c_uint32]),
    "sth_batch_get_tuning": (C.c_int, [_VP, _P(Tuning), _P(C.c_uint32), _P(C.c_uint32)]),
    "sth_batch_leapfrog": (C.c_int, [_VP, C.c_uint32, C.c_float, C.c_float, _P(C.c_float),
                                     C.c_uint32]),
    "sth_batch_sync": (C.c_int, [_VP]),
    "sth_batch_autotune": (C.c_int, [_VP, C.c_uint32, _P(Tuning), _P(C.c_uint32),
                                     _P(C.c_float)]),
    "sth_batch_best_method": (C.c_int, [_VP, _P(C.c_uint32)]),
    "sth_batch_timer_start": (C.c_int, [_VP]),
    "sth_batch_timer_read": (C.c_int, [_VP, _P(C.c_double), _P(C.c_uint64)]),
}

BATCH_PERSISTENT, BATCH_SEQUENTIAL, BATCH_TWO_STEP, BATCH_THREE_STEP, BATCH_FUSED = 0, 1, 2, 3, 4
BATCH_GRAPH = 0x100

_lib = None


def lib():
    """Loads libstencil_hip.so (built in-tree by stencil3d_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m stencil3d_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _lib_or_none_strerror(status):
    try:
        return lib().sth_strerror(status).decode()
    except Exception:  # noqa: BLE001 - only used to format an error message
        return "error"


def _ok(status, what):
    if status != STH_OK:
        raise SthError(status, what)


def radius(npts):
    return {7: 1, 13: 2, 25: 4}[npts]


def layout(xs, ys, zs, npts, dtype="fp32"):
    g = Geom(xs, ys, zs, npts, DTYPES[dtype])
    out = LayoutInfo()
    _ok(lib().sth_layout(C.byref(g), C.byref(out)), "sth_layout")
    return out


def leapfrog_c0(npts, c1):
    """Centre coefficient that keeps constant fields constant (FP32)."""
    out = C.c_float()
    _ok(lib().sth_leapfrog_c0(npts, c1, C.byref(out)), "sth_leapfrog_c0")
    return out.value


def ricker(steps, dt=1e-3, f=15.0, t0=None, scale=None):
    """Per-step Ricker source coefficients (float32): scale * wavelet(n dt - t0);
    defaults t0 = 1/f, scale = dt^2."""
    out = np.zeros(steps, np.float32)
    t0 = 1.0 / f if t0 is None else t0
    scale = dt * dt if scale is None else scale
    _ok(lib().sth_ricker(steps, dt, f, t0, scale, out.ctypes.data_as(_P(C.c_float))),
        "sth_ricker")
    return out


def device_count():
    n = C.c_int(0)
    lib().sth_device_count(C.byref(n))
    return n.value


class Stencil:
    """One run of the sweep on the current GPU: three device fields of an
    xs x ys x zs interior (i slowest, k unit stride) for an npts-point stencil,
    stored as dtype ('fp32', 'fp16', 'bf16', 'fp64')."""

    def __init__(self, xs, ys, zs, npts=7, dtype="fp32", _handle=None):
        self.xs, self.ys, self.zs, self.npts, self.dtype = xs, ys, zs, npts, dtype
        self.R = radius(npts)
        self._g = Geom(xs, ys, zs, npts, DTYPES[dtype])
        self._owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            _ok(lib().sth_create(C.byref(self._g), C.byref(h)), "sth_create")
            self._h = h
        else:
            self._h = C.c_void_p(_handle)

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().sth_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("context closed")
        return self._h

    def layout(self):
        out = LayoutInfo()
        _ok(lib().sth_get_layout(self.handle, C.byref(out)), "sth_get_layout")
        return out

    # -- fields
    def host_shape(self, halo):
        xh, yh, zh = _halo3(halo)
        return (self.xs + 2 * xh, self.ys + 2 * yh, self.zs + 2 * zh)

    def init_const(self, v=0.0, s=1.0, u=1.0):
        """init_arrays (iterators.cpp:292-301): v=0, s=1, u=1 everywhere."""
        _ok(lib().sth_init_const(self.handle, v, s, u), "sth_init_const")

    def init_uniform(self, field, seed, field_id, lo=-1.0, hi=1.0, halo_value=0.0):
        _ok(lib().sth_init_uniform(self.handle, FIELDS[field], seed, field_id, lo, hi,
                                   halo_value), "sth_init_uniform")

    def init_gaussian(self, field, amp=1.0, sigma=8.0, noise=1e-3, seed=1, field_id=0):
        _ok(lib().sth_init_gaussian(self.handle, FIELDS[field], amp, sigma, noise, seed,
                                    field_id), "sth_init_gaussian")

    def init_point(self, field, i, j, k, value=1.0):
        _ok(lib().sth_init_point(self.handle, FIELDS[field], i, j, k, value), "sth_init_point")

    def upload(self, field, arr, halo):
        """float32 arrays go through sth_upload, float64 through sth_upload_f64."""
        hl = HostLayout(*_halo3(halo))
        if np.asarray(arr).dtype == np.float64:
            arr = np.ascontiguousarray(arr)
            assert arr.shape == self.host_shape(halo)
            _ok(lib().sth_upload_f64(self.handle, FIELDS[field],
                                     arr.ctypes.data_as(_P(C.c_double)), C.byref(hl)),
                "sth_upload_f64")
            return
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        assert arr.shape == self.host_shape(halo), (arr.shape, self.host_shape(halo))
        _ok(lib().sth_upload(self.handle, FIELDS[field],
                             arr.ctypes.data_as(_P(C.c_float)), C.byref(hl)), "sth_upload")

    def download(self, field, halo, out=None, dtype=np.float32):
        if out is None:
            out = np.zeros(self.host_shape(halo), dtype=dtype)
        assert out.flags["C_CONTIGUOUS"] and out.shape == self.host_shape(halo)
        hl = HostLayout(*_halo3(halo))
        if out.dtype == np.float64:
            _ok(lib().sth_download_f64(self.handle, FIELDS[field],
                                       out.ctypes.data_as(_P(C.c_double)), C.byref(hl)),
                "sth_download_f64")
            return out
        assert out.dtype == np.float32
        _ok(lib().sth_download(self.handle, FIELDS[field],
                               out.ctypes.data_as(_P(C.c_float)), C.byref(hl)), "sth_download")
        return out

    def upload_raw16(self, field, arr, exp, halo):
        arr = np.ascontiguousarray(arr, dtype=np.uint16)
        assert arr.shape == self.host_shape(halo)
        hl = HostLayout(*_halo3(halo))
        _ok(lib().sth_upload_raw16(self.handle, FIELDS[field],
                                   arr.ctypes.data_as(_P(C.c_uint16)), exp, C.byref(hl)),
            "sth_upload_raw16")

    def download_raw16(self, field, halo):
        out = np.zeros(self.host_shape(halo), dtype=np.uint16)
        e, mx = C.c_int(), C.c_float()
        hl = HostLayout(*_halo3(halo))
        _ok(lib().sth_download_raw16(self.handle, FIELDS[field],
                                     out.ctypes.data_as(_P(C.c_uint16)), C.byref(e),
                                     C.byref(mx), C.byref(hl)), "sth_download_raw16")
        return out, e.value, mx.value

    def save(self, directory):
        """Writes u, v, s and a JSON manifest into an existing directory."""
        _ok(lib().sth_save(self.handle, os.fsencode(directory)), "sth_save")

    def load(self, directory):
        _ok(lib().sth_load(self.handle, os.fsencode(directory)), "sth_load")

    # -- the sweep
    def iterate(self, iterations):
        """ITERATE replacement: REF sweeps, v updated in place."""
        _ok(lib().sth_iterate(self.handle, iterations), "sth_iterate")

    def leapfrog(self, steps, c0, c1, c2=None):
        c2a = None
        if c2 is not None:
            c2a = np.ascontiguousarray(np.broadcast_to(np.asarray(c2, np.float32), (steps,)))
        ptr = None if c2a is None else c2a.ctypes.data_as(_P(C.c_float))
        _ok(lib().sth_leapfrog(self.handle, steps, c0, c1, ptr), "sth_leapfrog")

    def sync(self):
        _ok(lib().sth_sync(self.handle), "sth_sync")

    def check(self, expected):
        """check (iterators.cpp:304-325): count of |v - expected| > 0.1."""
        n = C.c_uint64()
        _ok(lib().sth_check(self.handle, expected, C.byref(n)), "sth_check")
        return n.value

    def norms(self, field):
        mx, l2 = C.c_double(), C.c_double()
        _ok(lib().sth_norms(self.handle, FIELDS[field], C.byref(mx), C.byref(l2)), "sth_norms")
        return mx.value, l2.value

    def energy(self, c0, c1):
        """The LEAPFROG scheme's discrete energy <u,u> + <v,v> - <u,(c0 + c1 L) v>
        over the interior (u newest, v the one before), in FP64 on the GPU;
        conserved by steps with s = 0 and zero halos."""
        q = C.c_double()
        _ok(lib().sth_energy(self.handle, c0, c1, C.byref(q)), "sth_energy")
        return q.value

    def compare(self, field, ref, ref_field=None):
        """(max-rel, L2-rel) of this context's field against `ref`'s (same
        geometry; e.g. an FP64 twin run), computed on the GPU."""
        mr, lr = C.c_double(), C.c_double()
        _ok(lib().sth_compare(self.handle, FIELDS[field], ref.handle,
                              FIELDS[ref_field or field], C.byref(mr), C.byref(lr)),
            "sth_compare")
        return mr.value, lr.value

    def set_weights(self, mode):
        """FP64 storage: 'fp32' (FP32 constants promoted, the default) or
        'reference' (the reference's -DUSE_DOUBLE weights)."""
        _ok(lib().sth_set_weights(self.handle, WEIGHTS[mode]), "sth_set_weights")

    def get_weights(self):
        m = C.c_int()
        _ok(lib().sth_get_weights(self.handle, C.byref(m)), "sth_get_weights")
        return {v: k for k, v in WEIGHTS.items()}[m.value]

    # -- schedule and timing
    def set_tuning(self, rows_per_wave=0, waves=0, chunks=0, xcd_remap=1, vec=0,
                   steps_per_pass=0, launch=0, ref_sweeps=0):
        t = Tuning(rows_per_wave, waves, chunks, xcd_remap, vec, steps_per_pass, launch,
                   ref_sweeps)
        _ok(lib().sth_set_tuning(self.handle, C.byref(t)), "sth_set_tuning")

    def get_tuning(self):
        t = Tuning()
        _ok(lib().sth_get_tuning(self.handle, C.byref(t)), "sth_get_tuning")
        return t.as_dict()

    def autotune(self, leapfrog=True, reps=3):
        t, ms = Tuning(), C.c_float()
        _ok(lib().sth_autotune(self.handle, 1 if leapfrog else 0, reps, C.byref(t),
                               C.byref(ms)), "sth_autotune")
        return t.as_dict(), ms.value

    def timer_start(self):
        _ok(lib().sth_timer_start(self.handle), "sth_timer_start")

    def stream_copy(self, nbytes, reps=5):
        """Plain copy kernel u -> v of nbytes; returns mean ms per copy."""
        ms = C.c_double()
        _ok(lib().sth_stream_copy(self.handle, nbytes, reps, C.byref(ms)), "sth_stream_copy")
        return ms.value

    def bw_probe(self, reads, writes, lane_bytes, nbytes, reps=5, grid=0):
        """Streaming probe (reads of u, s, v; writes of v and u's twin) of
        nbytes per stream; returns mean ms per launch."""
        ms = C.c_double()
        _ok(lib().sth_bw_probe(self.handle, reads, writes, lane_bytes, nbytes, grid, reps,
                               C.byref(ms)), "sth_bw_probe")
        return ms.value

    def place(self, tries=3):
        """Keeps the fastest of the current field buffers and up to `tries`
        further allocations of them (the LEAPFROG default timed on each);
        returns (ms per step of each set, 0 where none was allocated; the
        index kept).  Field contents are undefined afterwards."""
        ms = (C.c_float * (tries + 1))()
        chosen = C.c_uint32()
        _ok(lib().sth_place(self.handle, tries, ms, C.byref(chosen)), "sth_place")
        return list(ms), chosen.value

    def timer_read(self):
        ms, n = C.c_double(), C.c_uint64()
        _ok(lib().sth_timer_read(self.handle, C.byref(ms), C.byref(n)), "sth_timer_read")
        return ms.value, n.value


class Batch:
    """Many independent grids advanced together (sth_batch_*).  dims is a list
    of (xs, ys, zs); shot(s) returns a Stencil view of one shot's fields."""

    def __init__(self, dims, npts=13, dtype="fp32"):
        flat = (C.c_uint32 * (3 * len(dims)))(*[int(d) for xyz in dims for d in xyz])
        h = C.c_void_p()
        _ok(lib().sth_batch_create(flat, len(dims), npts, DTYPES[dtype], C.byref(h)),
            "sth_batch_create")
        self._h = h
        self.dims = [tuple(d) for d in dims]
        self.npts, self.dtype = npts, dtype
        self._shots = [Stencil(*d, npts=npts, dtype=dtype,
                               _handle=lib().sth_batch_shot(h, q))
                       for q, d in enumerate(self.dims)]

    def close(self):
        if getattr(self, "_h", None):
            for s in self._shots:
                s.close()
            lib().sth_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self._shots)

    def shot(self, q) -> Stencil:
        return self._shots[q]

    def set_tuning(self, rows_per_wave=0, waves=0, planes=0, grid_blocks=0, xcd_remap=1, vec=0):
        t = Tuning(rows_per_wave, waves, 0, xcd_remap, vec)
        _ok(lib().sth_batch_set_tuning(self._h, C.byref(t), planes, grid_blocks),
            "sth_batch_set_tuning")

    def get_tuning(self):
        t, p, g = Tuning(), C.c_uint32(), C.c_uint32()
        _ok(lib().sth_batch_get_tuning(self._h, C.byref(t), C.byref(p), C.byref(g)),
            "sth_batch_get_tuning")
        d = t.as_dict()
        d.update(planes=p.value, grid_blocks=g.value)
        return d

    def leapfrog(self, steps, c0, c1, c2=None, method=BATCH_PERSISTENT, graph=False):
        c2a = None
        if c2 is not None:
            c2a = np.ascontiguousarray(np.broadcast_to(np.asarray(c2, np.float32), (steps,)))
        ptr = None if c2a is None else c2a.ctypes.data_as(_P(C.c_float))
        m = method | (BATCH_GRAPH if graph else 0)
        _ok(lib().sth_batch_leapfrog(self._h, steps, c0, c1, ptr, m), "sth_batch_leapfrog")

    def sync(self):
        _ok(lib().sth_batch_sync(self._h), "sth_batch_sync")

    def autotune(self, reps=2):
        t, p, ms = Tuning(), C.c_uint32(), C.c_float()
        _ok(lib().sth_batch_autotune(self._h, reps, C.byref(t), C.byref(p), C.byref(ms)),
            "sth_batch_autotune")
        d = t.as_dict()
        d.update(planes=p.value)
        return d, ms.value

    def best_method(self):
        """The method the last autotune found fastest (BATCH_PERSISTENT or
        BATCH_FUSED)."""
        m = C.c_uint32()
        _ok(lib().sth_batch_best_method(self._h, C.byref(m)), "sth_batch_best_method")
        return m.value

    def timer_start(self):
        _ok(lib().sth_batch_timer_start(self._h), "sth_batch_timer_start")

    def timer_read(self):
        ms, n = C.c_double(), C.c_uint64()
        _ok(lib().sth_batch_timer_read(self._h, C.byref(ms), C.byref(n)),
            "sth_batch_timer_read")
        return ms.value, n.value


def _halo3(halo):
    if isinstance(halo, int):
        return (halo, halo, halo)
    return tuple(int(h) for h in halo)


# Names of the reference's operator interface (iterators.h:6-35).
def init_arrays(ctx: Stencil):
    ctx.init_const(0.0, 1.0, 1.0)


def iterate_3d(ctx: Stencil, iterations: int):
    ctx.iterate(iterations)


def check(ctx: Stencil, maxiter: int) -> int:
    return ctx.check(float(maxiter))

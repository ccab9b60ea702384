"""sth_leapfrog_noalloc: the LEAPFROG sweep on fields the caller allocated
(here with hipMalloc from the HIP runtime directly, in the device layout
sth_layout describes), with the library allocating nothing.  Checked bit for
bit against sth_leapfrog on a context holding the same fields, on two calls
in a row (the second on a smaller grid, with another stencil), on the null
stream and on a stream of the caller's, FP32 and FP64."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


H2D, D2H = 1, 2


def dev_index(L, shape, R):
    """Device-layout element index of every cell of a host array with an
    R-wide halo (shape (xs+2R, ys+2R, zs+2R))."""
    a, b, c = (np.arange(n, dtype=np.int64) - R for n in shape)
    return (L.origin + c[None, None, :] + b[None, :, None] * L.pitch +
            a[:, None, None] * L.plane)


class DevField:
    def __init__(self, L, dt):
        self.L, self.dt = L, dt
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), L.bytes) == 0
        self.host = np.zeros(L.bytes // np.dtype(dt).itemsize, dt)

    def put(self, arr, idx):
        self.host[:] = 0
        self.host[idx] = arr
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.L.bytes, H2D) == 0

    def get(self, idx):
        assert hip().hipMemcpy(self.host.ctypes.data, self.ptr, self.L.bytes, D2H) == 0
        return self.host[idx]

    def free(self):
        hip().hipFree(self.ptr)


def run_case(gpu, orc, npts, dtype, shape, steps, seed, use_stream):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    dt = np.float64 if dtype == "fp64" else np.float32
    u = orc.fill_uniform(g, seed, 0, halo=0.25).astype(dt)
    v = orc.fill_uniform(g, seed, 1, halo=-0.5).astype(dt)
    s = orc.fill_uniform(g, seed, 2, halo=0.125).astype(dt)
    c1 = float(np.float32(0.16))
    c0 = gpu.leapfrog_c0(npts, c1)
    c2 = np.linspace(-2e-3, 3e-3, steps).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        for name, x in (("u", u), ("v", v), ("s", s)):
            ctx.upload(name, x, R)
        ctx.leapfrog(steps, c0, c1, c2)
        want_u = ctx.download("u", R, dtype=dt)
        want_v = ctx.download("v", R, dtype=dt)
    L = gpu.layout(*shape, npts, dtype)
    idx = dev_index(L, g.shape, R)
    fields = [DevField(L, dt) for _ in range(3)]
    stream = C.c_void_p()
    try:
        for f, x in zip(fields, (u, v, s)):
            f.put(x, idx)
        if use_stream:
            assert hip().hipStreamCreate(C.byref(stream)) == 0
        geom = gpu.Geom(*shape, npts, gpu.DTYPES[dtype])
        rc = gpu.lib().sth_leapfrog_noalloc(
            C.byref(geom), fields[0].ptr, fields[1].ptr, fields[2].ptr, steps, c0, c1,
            c2.ctypes.data_as(C.POINTER(C.c_float)), stream)
        assert rc == gpu.STH_OK, gpu.lib().sth_strerror(rc)
        if use_stream:
            assert hip().hipStreamSynchronize(stream) == 0
        else:
            assert hip().hipDeviceSynchronize() == 0
        a, b = fields[0].get(idx), fields[1].get(idx)
    finally:
        for f in fields:
            f.free()
        if use_stream and stream.value:
            hip().hipStreamDestroy(stream)
    got_u, got_v = (a, b) if steps % 2 == 0 else (b, a)
    assert got_u.tobytes() == want_u.tobytes()
    assert got_v.tobytes() == want_v.tobytes()


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_noalloc_matches_leapfrog_twice(gpu, orc, dtype):
    # a larger grid first (several tiles and chunks, the persistent queue),
    # then a smaller one with another stencil: the second call must start
    # from cleared counters
    run_case(gpu, orc, 25, dtype, (48, 70, 150), 5, 3, use_stream=False)
    run_case(gpu, orc, 13, dtype, (20, 26, 66), 4, 4, use_stream=True)


def test_noalloc_rejects_what_it_cannot_run(gpu):
    geom16 = gpu.Geom(8, 8, 8, 7, gpu.DTYPES["fp16"])
    geom = gpu.Geom(8, 8, 8, 7, gpu.DTYPES["fp32"])
    p = C.c_void_p(1 << 20)  # never dereferenced: the calls fail their checks first
    L = gpu.lib()
    assert L.sth_leapfrog_noalloc(C.byref(geom16), p, C.c_void_p(2 << 20), C.c_void_p(3 << 20),
                                  1, 1.0, 0.1, None, None) == gpu.STH_ERR_UNSUPPORTED
    assert L.sth_leapfrog_noalloc(C.byref(geom), p, p, C.c_void_p(3 << 20), 1, 1.0, 0.1, None,
                                  None) == gpu.STH_ERR_ARG
    bad = gpu.Geom(0, 8, 8, 7, gpu.DTYPES["fp32"])
    assert L.sth_leapfrog_noalloc(C.byref(bad), p, C.c_void_p(2 << 20), C.c_void_p(3 << 20), 1,
                                  1.0, 0.1, None, None) == gpu.STH_ERR_ARG

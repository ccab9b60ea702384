"""The LEAPFROG scheme's discrete energy on the GPU (sth_energy):

    Q = <u,u> + <v,v> - <u, M v>,   M = c0 + c1 L   (u newest, v the one before)

against the same sum formed in numpy from the downloaded fields, with M v taken
from one step of the oracle's FP64 LEAPFROG (started from (v, 0) with s = 0,
which leaves c0 v + c1 L v); and its conservation over steps with s = 0 and
zero halos (exact for FP64 up to rounding, within the storage format's
rounding for FP32 and 16 bits).  tests/test_oracle.py pins the conservation of
the same Q on the oracle side."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (14, 12, 70)  # zs > 64: two wave columns, and an odd-free row for 16-bit pairs


def energy_np(orc, npts, g, un, um, c0, c1):
    """(Q, <un,un> + <um,um>) in FP64 from host arrays with halos."""
    z = g.zeros(np.float64)
    y, mv = np.asarray(um, np.float64).copy(), g.zeros(np.float64)
    assert orc.leapfrog_f64(npts, g, y, mv, z, float(c0), float(c1), None, 1) == 1
    I = g.interior
    a = np.asarray(un, np.float64)[I]
    b = np.asarray(um, np.float64)[I]
    sq = float((a * a).sum() + (b * b).sum())
    return sq - float((a * mv[I]).sum()), sq


def setup(gpu, orc, npts, dtype, seed):
    R = orc.radius(npts)
    g = orc.Grid(*SHAPE, R, R, R)
    u = orc.fill_uniform(g, seed, 0)
    v = orc.fill_uniform(g, seed, 2)
    ctx = gpu.Stencil(*SHAPE, npts=npts, dtype=dtype)
    ctx.upload("u", u, R)
    ctx.upload("v", v, R)
    ctx.upload("s", g.zeros(), R)
    return g, R, ctx


def fields(ctx, R):
    return (ctx.download("u", R, dtype=np.float64), ctx.download("v", R, dtype=np.float64))


@pytest.mark.parametrize("dtype", ["fp32", "fp64", "fp16", "bf16"])
@pytest.mark.parametrize("npts", [7, 13, 25])
def test_energy_matches_numpy(gpu, orc, npts, dtype):
    """sth_energy equals Q formed from the downloaded fields, before and after
    steps (16-bit: the decoded stored values, each field with its own scale)."""
    c1 = np.float32(0.1)
    c0 = np.float32(2.0 - orc.lap_gain(npts) * float(c1))
    g, R, ctx = setup(gpu, orc, npts, dtype, seed=4)
    with ctx:
        for steps in (0, 3, 4):
            if steps:
                ctx.leapfrog(steps, float(c0), float(c1))
            q = ctx.energy(float(c0), float(c1))
            un, um = fields(ctx, R)
            want, scale = energy_np(orc, npts, g, un, um, c0, c1)
            assert abs(q - want) <= 1e-12 * scale, (steps, q, want)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_energy_conserved(gpu, orc, npts):
    """s = 0, zero halos: FP64 keeps Q to rounding, FP32 within 1e-5, and a
    wrong operator (c0 off by 1e-3) visibly does not conserve the GPU's Q."""
    c1 = np.float32(0.1)
    c0 = np.float32(2.0 - orc.lap_gain(npts) * float(c1))
    drift = {}
    for dtype, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        g, R, ctx = setup(gpu, orc, npts, dtype, seed=5)
        with ctx:
            q0 = ctx.energy(float(c0), float(c1))
            ctx.leapfrog(20, float(c0), float(c1))
            q1 = ctx.energy(float(c0), float(c1))
        drift[dtype] = abs(q1 - q0) / abs(q0)
        assert drift[dtype] <= tol, (dtype, q0, q1)
    # the same Q measured with a perturbed c0 is no longer invariant
    g, R, ctx = setup(gpu, orc, npts, "fp64", seed=5)
    with ctx:
        bad = float(c0) + 1e-3
        q0 = ctx.energy(bad, float(c1))
        ctx.leapfrog(20, float(c0), float(c1))
        q1 = ctx.energy(bad, float(c1))
    assert abs(q1 - q0) / abs(q0) > 1e3 * drift["fp64"] + 1e-9


def test_energy_rejects_bad_arguments(gpu):
    import ctypes as C
    lib = gpu.lib()
    assert lib.sth_energy(None, 1.0, 0.1, C.byref(C.c_double())) != 0
    with gpu.Stencil(8, 8, 8, npts=7) as ctx:
        assert lib.sth_energy(ctx.handle, 1.0, 0.1, None) != 0

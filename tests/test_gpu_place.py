"""GPU tests of sth_place, the choice of the physical memory behind the
fields: the candidates are timed and the fastest kept, and the context then
behaves exactly like a fresh one (LEAPFROG bit-identical to the oracle, fused
pass and one-step alike)."""
import numpy as np
import pytest

from test_gpu_fused import coeffs, oracle_run, same_bits, seeded

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("npts,shape,spp", [(25, (80, 320, 384), 0), (13, (30, 200, 240), 1)])
def test_place_then_leapfrog_bit_exact(gpu, orc, npts, shape, spp):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 41, 0.25)
    c0, c1 = coeffs(orc, npts)
    steps = 5
    c2 = np.linspace(1e-3, 3e-3, steps).astype(np.float32)
    want_u, want_v = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, steps)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.set_tuning(steps_per_pass=spp)
        ms, chosen = ctx.place(2)
        assert len(ms) == 3 and 0 <= chosen <= 2
        assert ms[0] > 0 and ms[chosen] > 0
        assert all(m == 0 or m >= ms[chosen] for m in ms)
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.leapfrog(steps, c0, c1, c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    assert same_bits(gu, want_u)
    assert same_bits(gv, want_v)


def test_place_fp64_and_rejections(gpu, orc):
    npts, shape = 7, (20, 30, 40)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 43, 0.5)
    u, v, s = (x.astype(np.float64) for x in (u, v, s))
    c1 = float(np.float32(0.25))
    c0 = float(np.float32(2.0 - orc.lap_gain(npts) * c1))
    c2 = np.array([1e-3, -2e-3, 3e-3], np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype="fp64") as ctx:
        ms, chosen = ctx.place(1)
        assert ms[chosen] > 0
        for f, x in (("u", u), ("v", v), ("s", s)):
            ctx.upload(f, x, R)
        ctx.leapfrog(3, c0, c1, c2)
        gu = ctx.download("u", R, dtype=np.float64)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f64(npts, g, a, b, s, c0, c1, c2.astype(np.float64), 3)
    assert np.array_equal((a if newest == 0 else b).view(np.uint64), gu.view(np.uint64))
    with gpu.Stencil(*shape, npts=npts) as ctx:
        with pytest.raises(gpu.SthError) as e:
            ctx.place(9)
        assert e.value.status == gpu.STH_ERR_ARG



@pytest.mark.parametrize("npts,shape,spp", [(25, (20, 30, 200), 1), (7, (40, 130, 300), 3)])
def test_place_fp16_then_leapfrog_bit_exact(gpu, orc, npts, shape, spp):
    """Placement of 16-bit fields: the kept buffers hold no scale state and
    no zero halo until the fields are set (a LEAPFROG call before that is
    refused); after an upload the steps (the packed sweep, or for 7 points the
    three-step pass) are bit-identical to the oracle's emulation."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 47, 0.0)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(1, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.array([1e-2, -2e-2, 3e-2, 0.0], np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype="fp16") as ctx:
        ms, chosen = ctx.place(2)
        assert ms[chosen] > 0
        with pytest.raises(gpu.SthError) as e:
            ctx.leapfrog(1, c0, c1)
        assert e.value.status == gpu.STH_ERR_UNSUPPORTED
        assert ctx.get_tuning()["steps_per_pass"] == spp
        for f, q, ex in zip(("u", "v", "s"), qs, exps):
            ctx.upload_raw16(f, q, ex, R)
        ctx.leapfrog(4, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, 1, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 4, plen=spp)
    assert eu == exps[newest]
    assert np.array_equal(gu[g.interior], (a if newest == 0 else b)[g.interior])

"""Degenerate and tiny grids through every path: one-point rows, one-row
planes, one plane, widths below one wave, for REF and LEAPFROG in FP32, the
16-bit forms, the FP32 and FP16 passes of two to five steps and both launch
forms.  Every
result must equal the oracle bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 3, 5), (2, 1, 70), (3, 2, 1), (1, 1, 200), (5, 1, 1), (2, 9, 3)]


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", SHAPES)
def test_tiny_fp32(gpu, orc, npts, shape):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u = orc.fill_uniform(g, 41, 0, halo=0.5)
    s = orc.fill_uniform(g, 41, 1, halo=0.25)
    v = orc.fill_uniform(g, 41, 2, halo=-0.5)
    c1 = np.float32(0.16)
    c0 = np.float32(2.0 - orc.lap_gain(npts) * float(c1))
    c2 = np.full(5, 1e-3, np.float32)
    want_ref = v.copy()
    orc.ref_sweep_f32(npts, g, want_ref, s, u, 2)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(npts, g, a, b, s, float(c0), float(c1), c2, 5)
    want_u = a if newest == 0 else b
    I = g.interior
    for tuning in ({"launch": 1, "steps_per_pass": 1}, {"launch": 2, "steps_per_pass": 1},
                   {"steps_per_pass": 2}) + (({"steps_per_pass": 3},) if npts != 25 else ()) \
            + (({"steps_per_pass": 4}, {"steps_per_pass": 5}) if npts == 7 else ()):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(**tuning)
            for f, x in (("u", u), ("v", v), ("s", s)):
                ctx.upload(f, x, R)
            ctx.iterate(2)
            got_ref = ctx.download("v", R)
            ctx.upload("v", v, R)
            ctx.leapfrog(5, float(c0), float(c1), c2)
            got_u = ctx.download("u", R)
        assert same_bits(got_ref[I], want_ref[I]), tuning
        assert same_bits(got_u[I], want_u[I]), tuning


@pytest.mark.parametrize("fmt,dtype", [(1, "fp16"), (2, "bf16")])
@pytest.mark.parametrize("shape", SHAPES)
def test_tiny_q16(gpu, orc, fmt, dtype, shape):
    npts = 13
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    xs = [orc.fill_uniform(g, 43, k) for k in range(3)]
    xs[2] = xs[2] * np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in xs:
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.2, 2.0 - orc.lap_gain(npts) * 0.2
    c2 = np.linspace(0.5, -0.5, 3).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(3, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 3)
    wu, weu, wmu = (a, exps[0], maxs[0]) if newest == 0 else (b, exps[1], maxs[1])
    I = g.interior
    assert eu == weu
    assert np.array_equal(gu[I], wu[I])
    assert np.float32(mu) == np.float32(wmu)


@pytest.mark.parametrize("fmt,dtype", [(1, "fp16"), (2, "bf16")])
@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", SHAPES)
def test_tiny_16bit_passes(gpu, orc, fmt, dtype, npts, shape):
    """16-bit passes of every length the stencil has (FP16: two to five steps
    for 7 points, two and three for 13, two for 25; BF16: two and three for
    7 points, two for 13) asked for on degenerate grids,
    where the passes' warm-up planes and edge tiles lie mostly outside the
    grid: five steps (passes and the steps left over), bits and exponents
    equal to the oracle's emulation of the same passes."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    xs = [orc.fill_uniform(g, 47, k) for k in range(3)]
    xs[2] = xs[2] * np.float32(1e-3)
    q0, e0, m0 = [], [], []
    for x in xs:
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        q0.append(q)
        e0.append(e)
        m0.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 5).astype(np.float32)
    I = g.interior
    longest = {1: {7: 5, 13: 3, 25: 2}, 2: {7: 3, 13: 2, 25: 1}}[fmt][npts]
    for spp in range(2, longest + 1):
        with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            for f, q, e in zip(("u", "v", "s"), q0, e0):
                ctx.upload_raw16(f, q, e, R)
            ctx.leapfrog(5, c0, c1, c2)
            gu, eu, mu = ctx.download_raw16("u", R)
        a, b = q0[0].copy(), q0[1].copy()
        exps, maxs = list(e0), list(m0)
        newest = orc.leapfrog_q16(npts, fmt, g, a, b, q0[2], exps, maxs, np.float32(c0),
                                  np.float32(c1), c2, 5, plen=spp)
        wu, weu, wmu = (a, exps[0], maxs[0]) if newest == 0 else (b, exps[1], maxs[1])
        assert eu == weu, spp
        assert np.array_equal(gu[I], wu[I]), spp
        assert np.float32(mu) == np.float32(wmu), spp

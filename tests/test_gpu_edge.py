"""GPU parity of the folded last k-tile: when a row's length leaves a last
k-tile of at most 64/f columns, that tile runs with f rows per wave and one k
point per lane.  The zs values below leave 16, 42, 64 and 72 columns after
the 128-column tiles (folds 4 or 2, 1, 1, none); every case must equal the
oracle bit for bit (REF, LEAPFROG FP32; LEAPFROG FP16/BF16 on the stored bits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ZS = [144, 170, 192, 200]  # q16 tiles run unfolded (FP32 only folds)
TILES = [(4, 8, 2), (2, 8, 2)]


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("zs", ZS)
@pytest.mark.parametrize("tile", TILES, ids=[f"{t[0]}x{t[1]}v{t[2]}" for t in TILES])
def test_folded_edge_fp32(gpu, orc, npts, zs, tile):
    shape = (7, 45, zs)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u = orc.fill_uniform(g, 21, 0, halo=0.25)
    s = orc.fill_uniform(g, 21, 1, halo=0.5)
    v = orc.fill_uniform(g, 21, 2, halo=-0.25)
    c1 = np.float32(0.16)
    c0 = np.float32(2.0 - orc.lap_gain(npts) * float(c1))
    c2 = np.full(3, 1e-3, np.float32)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.set_tuning(tile[0], tile[1], 0, vec=tile[2], steps_per_pass=1)
        for f, a in (("u", u), ("v", v), ("s", s)):
            ctx.upload(f, a, R)
        ctx.iterate(2)
        ref_v = ctx.download("v", R)
        ctx.upload("v", v, R)
        ctx.leapfrog(3, float(c0), float(c1), c2)
        gu = ctx.download("u", R)
    want = v.copy()
    orc.ref_sweep_f32(npts, g, want, s, u, 2)
    I = g.interior
    assert same_bits(ref_v[I], want[I])
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(npts, g, a, b, s, float(c0), float(c1), c2, 3)
    assert same_bits(gu[I], (a if newest == 0 else b)[I])


@pytest.mark.parametrize("fmt,dtype", [(1, "fp16"), (2, "bf16")])
@pytest.mark.parametrize("zs", ZS)
@pytest.mark.parametrize("tile", TILES, ids=[f"{t[0]}x{t[1]}v{t[2]}" for t in TILES])
def test_folded_edge_q16(gpu, orc, fmt, dtype, zs, tile):
    npts = 25
    shape = (6, 37, zs)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u = orc.fill_uniform(g, 31, 0)
    v = orc.fill_uniform(g, 31, 1)
    s = orc.fill_uniform(g, 31, 2) * np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 4).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        ctx.set_tuning(tile[0], tile[1], 0, vec=tile[2])
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(4, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 4)
    wu, weu, wmu = (a, exps[0], maxs[0]) if newest == 0 else (b, exps[1], maxs[1])
    I = g.interior
    assert eu == weu
    assert np.array_equal(gu[I], wu[I])
    assert np.float32(mu) == np.float32(wmu)

"""GPU parity: libstencil_hip (through its C ABI) against the CPU oracle on the
same seeded fields.  FP32 REF and LEAPFROG must be bit-identical; the 16-bit
storage form must be bit-identical to the oracle's emulation of it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (12, 14, 20), (33, 19, 70), (5, 7, 129)]


def seeded(orc, g, seed, halo=0.0):
    u = orc.fill_uniform(g, seed, 0, halo=halo)
    s = orc.fill_uniform(g, seed, 1, halo=halo)
    v = orc.fill_uniform(g, seed, 2, halo=halo)
    return v, s, u


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7fffffff), ai)
    bi = np.where(bi < 0, -(bi & 0x7fffffff), bi)
    return int(np.abs(ai - bi).max()) if a.size else 0


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", SHAPES)
def test_ref_sweep_bit_exact(gpu, orc, npts, shape):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=0)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        for f, a in (("v", v), ("s", s), ("u", u)):
            ctx.upload(f, a, R)
        ctx.iterate(5)
        out = ctx.download("v", R)
    orc.ref_sweep_f32(npts, g, v, s, u, 5)
    I = g.interior
    assert same_bits(out[I], v[I]), f"max ulp {ulp_diff(out[I], v[I])}"


# (rows per wave, waves, k points per lane, i-chunks): every compiled tile
TILES = [(1, 4, 1, 1), (2, 4, 1, 3), (4, 4, 1, 2), (8, 4, 1, 1), (1, 8, 1, 5), (2, 8, 1, 1),
         (4, 8, 1, 4), (8, 8, 1, 2), (2, 2, 1, 1), (4, 2, 1, 7), (8, 2, 1, 3), (1, 4, 2, 2),
         (2, 4, 2, 1), (4, 4, 2, 3), (2, 8, 2, 2), (4, 8, 2, 1), (2, 2, 2, 4), (4, 2, 2, 2)]


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("tile", TILES, ids=[f"{t[0]}x{t[1]}v{t[2]}c{t[3]}" for t in TILES])
def test_ref_every_tile_schedule(gpu, orc, npts, tile):
    """Every compiled tile/chunk schedule gives the same bits (the schedule is
    result-neutral, like the reference's iterators).  zs = 75 is odd, so the
    two-points-per-lane tiles end on a half-valid lane; zs > 128 gives them a
    second, partial k-tile."""
    shape = (21, 37, 75) if tile[2] == 1 else (21, 19, 139)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=3)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.set_tuning(tile[0], tile[1], tile[3], xcd_remap=tile[3] % 2, vec=tile[2])
        for f, a in (("v", v), ("s", s), ("u", u)):
            ctx.upload(f, a, R)
        ctx.iterate(3)
        out = ctx.download("v", R)
    orc.ref_sweep_f32(npts, g, v, s, u, 3)
    I = g.interior
    assert same_bits(out[I], v[I])


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_device_seeded_init_matches_oracle(gpu, orc, npts):
    shape = (9, 10, 67)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.init_uniform("u", 7, 0, -1.0, 1.0, 0.25)
        got = ctx.download("u", R)
    want = orc.fill_uniform(g, 7, 0, -1.0, 1.0, 0.25)
    assert same_bits(got, want)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_reference_kat(gpu, npts):
    """The reference's own check: u=s=1, v=0 everywhere (halos included),
    one warm-up sweep then 20 more -> v == #sweeps within 0.1
    (main.cpp:289-312, iterators.cpp:292-325)."""
    with gpu.Stencil(24, 24, 24, npts=npts) as ctx:
        gpu.init_arrays(ctx)
        gpu.iterate_3d(ctx, 1)
        assert gpu.check(ctx, 1) == 0
        gpu.iterate_3d(ctx, 20)
        assert gpu.check(ctx, 21) == 0
        # and a wrong expectation is caught
        assert gpu.check(ctx, 22) == 24 ** 3


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_kat_exact_recurrence(gpu, orc, npts):
    """With u = s = 1 the stencil value P is a constant, so every interior v
    follows the scalar FP32 recurrence v <- fma(1, P, fma(-A, 1, v)) exactly."""
    from fractions import Fraction as F

    def rnd(x):
        return float(np.float32(float(x)))  # exact rational -> nearest float32

    coef0, A, w, R = orc.coeffs(npts)
    terms = orc.terms(npts)
    a0, r0 = terms[0]
    p = rnd(F(float(coef0)) + F(rnd(F(float(w[a0, r0 - 1])) * 2)))
    for ax, r in terms[1:]:
        p = rnd(F(float(w[ax, r - 1])) * 2 + F(p))
    vN = 0.0
    for _ in range(13):
        vN = rnd(F(p) + F(rnd(F(vN) - F(float(A)))))
    with gpu.Stencil(10, 11, 12, npts=npts) as ctx:
        ctx.init_const(0.0, 1.0, 1.0)
        ctx.iterate(13)
        out = ctx.download("v", R)
    g = orc.Grid(10, 11, 12, R, R, R)
    assert np.all(out[g.interior] == np.float32(vN))


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", [(16, 16, 16), (19, 33, 70)])
@pytest.mark.parametrize("vec", [1, 2])
def test_leapfrog_bit_exact(gpu, orc, npts, shape, vec):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=0)  # zero halos
    c1 = np.float32(0.25 if npts != 25 else 0.16)
    c0 = np.float32(2.0 - orc.lap_gain(npts) * float(c1))
    c2 = (np.arange(1, 8, dtype=np.float32) * np.float32(1e-3)).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.set_tuning(vec=vec)
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.leapfrog(7, float(c0), float(c1), c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(npts, g, a, b, s, float(c0), float(c1), c2, 7)
    want_u, want_v = (a, b) if newest == 0 else (b, a)
    I = g.interior
    assert same_bits(gu[I], want_u[I])
    assert same_bits(gv[I], want_v[I])


@pytest.mark.parametrize("fmt,dtype", [(1, "fp16"), (2, "bf16")])
@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("vec", [1, 2])
def test_q16_leapfrog_bit_exact(gpu, orc, fmt, dtype, npts, vec):
    shape = (14, 17, 71)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=5)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 9).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        ctx.set_tuning(vec=vec)
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(9, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
        gv, ev, mv = ctx.download_raw16("v", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 9)
    (wu, weu, wmu), (wv, wev, wmv) = (
        ((a, exps[0], maxs[0]), (b, exps[1], maxs[1])) if newest == 0 else
        ((b, exps[1], maxs[1]), (a, exps[0], maxs[0])))
    I = g.interior
    assert (eu, ev) == (weu, wev)
    assert np.array_equal(gu[I], wu[I]) and np.array_equal(gv[I], wv[I])
    assert np.float32(mu) == np.float32(wmu) and np.float32(mv) == np.float32(wmv)


PACKED16_TILES = [(1, 4), (2, 4), (4, 4), (2, 8), (4, 8), (2, 2), (4, 2), (1, 8), (3, 8),
                  (1, 16), (2, 16), (3, 16)]


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("tile", PACKED16_TILES)
@pytest.mark.parametrize("shape,chunks", [((23, 37, 201), 2), ((40, 19, 130), 3)])
def test_fp16_packed_tiles_bit_exact(gpu, orc, npts, tile, shape, chunks):
    """The packed FP16 sweep (two k points per lane in one 32-bit word) on
    every tile it compiles: rows and columns past the interior (an odd zs
    leaves a lane with one point inside), several i-chunks whose lengths are
    not whole turns of the kernel's rings; bits, exponents and maxima equal
    to the oracle's binary16 emulation."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=11 + npts)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(1, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 7).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype="fp16") as ctx:
        ctx.set_tuning(rows_per_wave=tile[0], waves=tile[1], vec=2, chunks=chunks)
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(7, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
        gv, ev, mv = ctx.download_raw16("v", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, 1, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 7)
    (wu, weu, wmu), (wv, wev, wmv) = (
        ((a, exps[0], maxs[0]), (b, exps[1], maxs[1])) if newest == 0 else
        ((b, exps[1], maxs[1]), (a, exps[0], maxs[0])))
    I = g.interior
    assert (eu, ev) == (weu, wev)
    assert np.array_equal(gu[I], wu[I]) and np.array_equal(gv[I], wv[I])
    assert np.float32(mu) == np.float32(wmu) and np.float32(mv) == np.float32(wmv)


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("tile", [(2, 8), (2, 16)])
@pytest.mark.parametrize("zs", [16, 144, 145, 160, 161, 192, 1040])
def test_fp16_packed_folded_edge_bit_exact(gpu, orc, npts, tile, zs):
    """The packed FP16 sweep's last k-tile, when it holds at most 64 of its
    128 columns, runs folded on the default 2x8 tile (8, 4 or 2 rows per lane
    group, 8, 4 or 2 tiles' rows per item; 2x16 runs it unfolded): widths 16
    (a grid of one folded tile), 16, 17 and 32, 33 and 64 past the last full
    tile, and 1040; 150 rows, so several folded tiles run down j and the last
    one is partly outside.  Bits, exponents and maxima equal to the oracle's
    binary16 emulation."""
    R = orc.radius(npts)
    shape = (9, 150, zs)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=17 + npts)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(1, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 5).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype="fp16") as ctx:
        ctx.set_tuning(rows_per_wave=tile[0], waves=tile[1], vec=2, chunks=2,
                       steps_per_pass=1)
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(5, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
        gv, ev, mv = ctx.download_raw16("v", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, 1, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 5)
    (wu, weu, wmu), (wv, wev, wmv) = (
        ((a, exps[0], maxs[0]), (b, exps[1], maxs[1])) if newest == 0 else
        ((b, exps[1], maxs[1]), (a, exps[0], maxs[0])))
    I = g.interior
    assert (eu, ev) == (weu, wev)
    assert np.array_equal(gu[I], wu[I]) and np.array_equal(gv[I], wv[I])
    assert np.float32(mu) == np.float32(wmu) and np.float32(mv) == np.float32(wmv)


# (storage format, stencil, rows per wave, waves, steps per pass): every
# compiled 16-bit pass
Q16_PASSES = [(1, 7, 4, 16, 2), (1, 13, 3, 16, 2), (1, 13, 4, 16, 2), (1, 25, 3, 16, 2),
              (1, 7, 3, 16, 3), (1, 13, 3, 16, 3), (1, 7, 2, 16, 4), (1, 7, 2, 16, 5),
              (2, 7, 4, 16, 2), (2, 7, 2, 16, 3), (2, 13, 2, 16, 2)]


@pytest.mark.parametrize("fmt,npts,mj,nw,sp", Q16_PASSES,
                         ids=[f"{'fp16' if f == 1 else 'bf16'}-{n}pt-{m}x{w}-s{s}"
                              for f, n, m, w, s in Q16_PASSES])
@pytest.mark.parametrize("shape,chunks", [((30, 61, 201), 2), ((44, 130, 97), 3),
                                          ((17, 12, 40), 1)])
def test_q16_pass_bit_exact(gpu, orc, fmt, npts, mj, nw, sp, shape, chunks):
    """The 16-bit passes of two to five steps (FP16, BF16): seven steps
    (passes of sp and the steps left over as one shorter pass or a single
    step), several i-chunks, output tiles cut by the grid's edges in j and k
    (an odd zs leaves a lane with one point inside), a source term; bits,
    exponents and maxima equal to the oracle's 16-bit emulation in the same
    passes."""
    dtype = "fp16" if fmt == 1 else "bf16"
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v, s, u = seeded(orc, g, seed=23 + npts)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 7).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        ctx.set_tuning(rows_per_wave=mj, waves=nw, chunks=chunks, steps_per_pass=sp)
        t = ctx.get_tuning()
        assert (t["steps_per_pass"], t["rows_per_wave"], t["waves"]) == (sp, mj, nw)
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(7, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
        gv, ev, mv = ctx.download_raw16("v", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 7, plen=sp)
    (wu, weu, wmu), (wv, wev, wmv) = (
        ((a, exps[0], maxs[0]), (b, exps[1], maxs[1])) if newest == 0 else
        ((b, exps[1], maxs[1]), (a, exps[0], maxs[0])))
    I = g.interior
    assert (eu, ev) == (weu, wev)
    assert np.array_equal(gu[I], wu[I]) and np.array_equal(gv[I], wv[I])
    assert np.float32(mu) == np.float32(wmu) and np.float32(mv) == np.float32(wmv)


def test_norms_match_oracle(gpu, orc):
    shape = (20, 21, 22)
    g = orc.Grid(*shape, 1, 1, 1)
    x = orc.fill_uniform(g, 11, 0)
    with gpu.Stencil(*shape, npts=7) as ctx:
        ctx.upload("u", x, 1)
        mx, l2 = ctx.norms("u")
    omx, ol2 = orc.norms(g, x)
    assert mx == omx
    assert abs(l2 - ol2) <= 1e-12 * ol2


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_fp64_ref_and_leapfrog_bit_exact(gpu, orc, npts):
    """FP64 storage: REF with the FP32 constants promoted, and LEAPFROG, both
    bit-identical to the oracle's FP64 sweeps."""
    shape = (13, 22, 45)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, s, v = (orc.fill_uniform(g, 8, f).astype(np.float64) for f in range(3))
    with gpu.Stencil(*shape, npts=npts, dtype="fp64") as ctx:
        for f, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(f, x, R)
        ctx.iterate(4)
        got = ctx.download("v", R, dtype=np.float64)
    want = v.copy()
    orc.ref_sweep_f64(npts, g, want, s, u, 4, coeff_mode=0)
    assert np.array_equal(got[g.interior], want[g.interior])

    c1 = float(np.float32(0.16))
    c0 = float(np.float32(2.0 - orc.lap_gain(npts) * 0.16))
    c2 = np.array([1e-3, -2e-3, 3e-3], np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype="fp64") as ctx:
        for f, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(f, x, R)
        ctx.leapfrog(3, c0, c1, c2)
        got = ctx.download("u", R, dtype=np.float64)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f64(npts, g, a, b, s, c0, c1, c2.astype(np.float64), 3)
    want = a if newest == 0 else b
    assert np.array_equal(got[g.interior], want[g.interior])


@pytest.mark.parametrize("dtype", ["fp32", "fp16"])
@pytest.mark.parametrize("npts", [7, 25])
def test_launch_modes_agree(gpu, orc, dtype, npts):
    """One workgroup per tile and the persistent work-queue launch (several
    calls in a row, so the queue's ticket base advances between launches)
    give the same bits."""
    shape = (40, 70, 300)
    R = orc.radius(npts)
    c1 = 0.16
    c0 = 2.0 - orc.lap_gain(npts) * c1
    out = []
    for launch in (1, 2):
        with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
            ctx.set_tuning(0, 0, 3, launch=launch, steps_per_pass=1 if dtype == "fp32" else 0)
            assert ctx.get_tuning()["launch"] == launch
            ctx.init_gaussian("u", 1.0, 6.0, 1e-3, 4, 0)
            ctx.init_gaussian("v", 1.0, 6.0, 1e-3, 4, 1)
            ctx.init_point("s", 20, 35, 150, 1.0)
            for rep in range(3):
                ctx.leapfrog(3, c0, c1, 1e-3)
            out.append(ctx.download("u", R, dtype=np.float64))
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("dtype", ["fp32", "fp16"])
def test_compare_matches_numpy(gpu, orc, dtype):
    """sth_compare (the error of a run against its FP64 twin, on the GPU)
    against the same figures computed with numpy from the downloads."""
    shape, npts = (23, 30, 77), 13
    R = orc.radius(npts)
    c1 = 0.16
    c0 = gpu.leapfrog_c0(npts, c1)
    runs = {}
    ctxs = [gpu.Stencil(*shape, npts=npts, dtype=dt) for dt in (dtype, "fp64")]
    try:
        for ctx in ctxs:
            ctx.init_gaussian("u", 1.0, 4.0, 1e-3, 1, 0)
            ctx.init_gaussian("v", 1.0, 4.0, 1e-3, 1, 1)
            ctx.init_point("s", 11, 15, 38, 1.0)
            ctx.leapfrog(9, c0, c1, 1e-2)
        mr, lr = ctxs[0].compare("u", ctxs[1])
        x = ctxs[0].download("u", R, dtype=np.float64)
        r = ctxs[1].download("u", R, dtype=np.float64)
        with pytest.raises(gpu.SthError):
            ctxs[0].compare("u", ctxs[0])
    finally:
        for ctx in ctxs:
            ctx.close()
    err = np.abs(x - r)
    want_mr = err.max() / np.abs(r).max()
    want_lr = np.sqrt((err ** 2).sum() / (r ** 2).sum())
    assert mr == want_mr  # the maxima are exact (FP64 on the GPU)
    assert lr == pytest.approx(want_lr, rel=1e-9)
    assert 0 < mr < (1e-5 if dtype == "fp32" else 1e-2)

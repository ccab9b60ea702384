"""GPU parity of the two-steps-per-pass LEAPFROG sweep (FP32, temporal
blocking).  A pass must be bit-identical to two one-step sweeps of the
oracle, halos included: the fields carry nonzero halo values, which the
one-step scheme never writes and which a pass must leave where the one-step
scheme leaves them.  Covered: every compiled fused shape, i-chunks, odd step
counts (the last step runs one-step), one-step calls and uploads between
calls, and shapes that are not multiples of the output tile; the same for
three steps per pass (7 and 13 points) and four and five (7 points)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FUSED = {7: [(4, 16), (4, 8)], 13: [(4, 16), (3, 16), (4, 8)], 25: [(3, 16), (4, 12), (3, 8), (4, 8)]}
SHAPES = [(20, 24, 36), (37, 45, 70), (9, 70, 130)]


def seeded(orc, g, seed, halo):
    u = orc.fill_uniform(g, seed, 0, halo=halo)
    v = orc.fill_uniform(g, seed, 1, halo=-halo)
    s = orc.fill_uniform(g, seed, 2, halo=2 * halo)
    return u, v, s


def coeffs(orc, npts):
    c1 = np.float32(0.25 if npts != 25 else 0.16)
    return float(np.float32(2.0 - orc.lap_gain(npts) * float(c1))), float(c1)


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def oracle_run(orc, npts, g, u, v, s, c0, c1, c2, steps):
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, steps)
    return (a, b) if newest == 0 else (b, a)


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", SHAPES)
def test_two_step_passes_bit_exact(gpu, orc, npts, shape):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 3, 0.375)
    c0, c1 = coeffs(orc, npts)
    steps = 5
    c2 = (np.arange(1, steps + 1, dtype=np.float32) * np.float32(1e-3)).astype(np.float32)
    want_u, want_v = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, steps)
    for mj, nw in FUSED[npts]:
        for chunks in (0, 2, 3):
            with gpu.Stencil(*shape, npts=npts) as ctx:
                ctx.set_tuning(mj, nw, chunks, steps_per_pass=2)
                t = ctx.get_tuning()
                assert t["steps_per_pass"] == 2 and (t["rows_per_wave"], t["waves"]) == (mj, nw)
                ctx.upload("u", u, R)
                ctx.upload("v", v, R)
                ctx.upload("s", s, R)
                ctx.timer_start()
                ctx.leapfrog(steps, c0, c1, c2)
                _, launches = ctx.timer_read()
                assert launches == 3  # two passes + one one-step sweep
                gu = ctx.download("u", R)
                gv = ctx.download("v", R)
            assert same_bits(gu, want_u), (mj, nw, chunks)
            assert same_bits(gv, want_v), (mj, nw, chunks)


@pytest.mark.parametrize("npts", [13, 25])
def test_passes_mixed_with_single_steps_and_uploads(gpu, orc, npts):
    """Two passes, one one-step sweep, a new v (other halo values), two more
    passes: the twin buffers must pick up the new halo."""
    shape = (26, 31, 67)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 11, 0.25)
    c0, c1 = coeffs(orc, npts)
    c2 = np.full(4, 2e-3, np.float32)
    # oracle: 4 steps, 1 step, replace v, 4 steps
    a, b = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, 4)
    a, b = oracle_run(orc, npts, g, a, b, s, c0, c1, c2[:1], 1)
    v2 = orc.fill_uniform(g, 12, 1, halo=0.625)
    want_u, want_v = oracle_run(orc, npts, g, a, v2, s, c0, c1, c2, 4)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.set_tuning(steps_per_pass=2)
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.leapfrog(4, c0, c1, c2)
        ctx.set_tuning(steps_per_pass=1)
        ctx.leapfrog(1, c0, c1, c2[:1])
        ctx.set_tuning(steps_per_pass=2)
        ctx.upload("v", v2, R)
        ctx.leapfrog(4, c0, c1, c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    assert same_bits(gu, want_u)
    assert same_bits(gv, want_v)


def test_passes_equal_one_step_sweeps_on_gpu(gpu, orc):
    """A larger grid (several tiles and chunks in every direction): 8 steps
    as passes and as one-step sweeps give the same bits."""
    shape = (96, 100, 200)
    npts, R = 25, 4
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-3, 1e-3, 8).astype(np.float32)
    out = []
    for spp in (1, 2):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            ctx.init_uniform("u", 5, 0, -1.0, 1.0, 0.5)
            ctx.init_uniform("v", 5, 1, -1.0, 1.0, -0.5)
            ctx.init_uniform("s", 5, 2, -1.0, 1.0, 0.25)
            ctx.leapfrog(8, c0, c1, c2)
            out.append((ctx.download("u", R), ctx.download("v", R)))
    assert same_bits(out[0][0], out[1][0])
    assert same_bits(out[0][1], out[1][1])


FUSED3 = {7: [(3, 16), (4, 12)], 13: [(3, 16), (4, 12)]}


@pytest.mark.parametrize("npts", [7, 13])
@pytest.mark.parametrize("shape", SHAPES)
def test_three_step_passes_bit_exact(gpu, orc, npts, shape):
    """Three steps per pass: 6 steps are two passes, 7 add a one-step sweep,
    8 add a two-step pass; every compiled shape, with i-chunks."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 5, 0.375)
    c0, c1 = coeffs(orc, npts)
    c2 = (np.arange(1, 9, dtype=np.float32) * np.float32(-7e-4)).astype(np.float32)
    want = {n: oracle_run(orc, npts, g, u, v, s, c0, c1, c2[:n], n) for n in (6, 7, 8)}
    for mj, nw in FUSED3[npts]:
        for chunks, steps in ((0, 6), (2, 7), (3, 8)):
            with gpu.Stencil(*shape, npts=npts) as ctx:
                ctx.set_tuning(mj, nw, chunks, steps_per_pass=3)
                t = ctx.get_tuning()
                assert t["steps_per_pass"] == 3 and (t["rows_per_wave"], t["waves"]) == (mj, nw)
                ctx.upload("u", u, R)
                ctx.upload("v", v, R)
                ctx.upload("s", s, R)
                ctx.timer_start()
                ctx.leapfrog(steps, c0, c1, c2[:steps])
                _, launches = ctx.timer_read()
                assert launches == (2 if steps == 6 else 3)
                gu = ctx.download("u", R)
                gv = ctx.download("v", R)
            assert same_bits(gu, want[steps][0]), (mj, nw, chunks, steps)
            assert same_bits(gv, want[steps][1]), (mj, nw, chunks, steps)


@pytest.mark.parametrize("npts", [7, 13])
def test_three_step_passes_mixed_with_other_schedules(gpu, orc, npts):
    """A three-step pass, a two-step pass, one one-step sweep, a new v, then
    four steps of three-step tuning (a pass and a sweep): the buffer roles
    after an odd pass and the twins' halos must follow."""
    shape = (26, 31, 67)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 13, 0.25)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(1e-3, 4e-3, 4).astype(np.float32)
    a, b = oracle_run(orc, npts, g, u, v, s, c0, c1, c2[:3], 3)
    a, b = oracle_run(orc, npts, g, a, b, s, c0, c1, c2[:2], 2)
    a, b = oracle_run(orc, npts, g, a, b, s, c0, c1, c2[:1], 1)
    v2 = orc.fill_uniform(g, 14, 1, halo=0.625)
    want_u, want_v = oracle_run(orc, npts, g, a, v2, s, c0, c1, c2, 4)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.set_tuning(steps_per_pass=3)
        ctx.leapfrog(3, c0, c1, c2[:3])
        ctx.set_tuning(steps_per_pass=2)
        ctx.leapfrog(2, c0, c1, c2[:2])
        ctx.set_tuning(steps_per_pass=1)
        ctx.leapfrog(1, c0, c1, c2[:1])
        ctx.set_tuning(steps_per_pass=3)
        ctx.upload("v", v2, R)
        ctx.leapfrog(4, c0, c1, c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    assert same_bits(gu, want_u)
    assert same_bits(gv, want_v)


def test_three_step_passes_equal_one_step_sweeps_on_gpu(gpu, orc):
    """Several tiles and chunks in every direction: 9 steps as three-step
    passes and as one-step sweeps give the same bits."""
    shape = (96, 100, 200)
    npts, R = 13, 2
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-3, 1e-3, 9).astype(np.float32)
    out = []
    for spp in (1, 3):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            ctx.init_uniform("u", 6, 0, -1.0, 1.0, 0.5)
            ctx.init_uniform("v", 6, 1, -1.0, 1.0, -0.5)
            ctx.init_uniform("s", 6, 2, -1.0, 1.0, 0.25)
            ctx.leapfrog(9, c0, c1, c2)
            out.append((ctx.download("u", R), ctx.download("v", R)))
    assert same_bits(out[0][0], out[1][0])
    assert same_bits(out[0][1], out[1][1])


FUSED4 = [(4, 10), (5, 8), (2, 16)]


@pytest.mark.parametrize("shape", SHAPES)
def test_four_step_passes_bit_exact(gpu, orc, shape):
    """Four steps per pass (7 points): 8 steps are two passes; 9, 10 and 11
    add a one-step sweep, a two-step pass or a three-step pass; every
    compiled shape, with i-chunks."""
    npts = 7
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 9, 0.375)
    c0, c1 = coeffs(orc, npts)
    c2 = (np.arange(1, 12, dtype=np.float32) * np.float32(6e-4)).astype(np.float32)
    want = {n: oracle_run(orc, npts, g, u, v, s, c0, c1, c2[:n], n) for n in (8, 9, 10, 11)}
    for mj, nw in FUSED4:
        for chunks, steps in ((0, 8), (2, 9), (3, 10), (0, 11)):
            with gpu.Stencil(*shape, npts=npts) as ctx:
                ctx.set_tuning(mj, nw, chunks, steps_per_pass=4)
                t = ctx.get_tuning()
                assert t["steps_per_pass"] == 4 and (t["rows_per_wave"], t["waves"]) == (mj, nw)
                ctx.upload("u", u, R)
                ctx.upload("v", v, R)
                ctx.upload("s", s, R)
                ctx.timer_start()
                ctx.leapfrog(steps, c0, c1, c2[:steps])
                _, launches = ctx.timer_read()
                assert launches == (2 if steps == 8 else 3)
                gu = ctx.download("u", R)
                gv = ctx.download("v", R)
            assert same_bits(gu, want[steps][0]), (mj, nw, chunks, steps)
            assert same_bits(gv, want[steps][1]), (mj, nw, chunks, steps)


@pytest.mark.parametrize("shape", SHAPES)
def test_five_step_passes_bit_exact(gpu, orc, shape):
    """Five steps per pass (7 points): 10 steps are two passes; 11 to 14 add
    a one-step sweep or a two-, three- or four-step pass; with i-chunks."""
    npts = 7
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 19, 0.375)
    c0, c1 = coeffs(orc, npts)
    c2 = (np.arange(1, 15, dtype=np.float32) * np.float32(-4e-4)).astype(np.float32)
    want = {n: oracle_run(orc, npts, g, u, v, s, c0, c1, c2[:n], n) for n in range(10, 15)}
    for chunks, steps in ((0, 10), (2, 11), (3, 12), (0, 13), (2, 14)):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(2, 16, chunks, steps_per_pass=5)
            t = ctx.get_tuning()
            assert t["steps_per_pass"] == 5 and (t["rows_per_wave"], t["waves"]) == (2, 16)
            ctx.upload("u", u, R)
            ctx.upload("v", v, R)
            ctx.upload("s", s, R)
            ctx.timer_start()
            ctx.leapfrog(steps, c0, c1, c2[:steps])
            _, launches = ctx.timer_read()
            assert launches == (2 if steps == 10 else 3)
            gu = ctx.download("u", R)
            gv = ctx.download("v", R)
        assert same_bits(gu, want[steps][0]), (chunks, steps)
        assert same_bits(gv, want[steps][1]), (chunks, steps)


def test_four_step_passes_mixed_with_other_schedules(gpu, orc):
    """A four-step pass, a three-step pass, one one-step sweep, a new v, then
    five steps of four-step tuning (a pass and a sweep)."""
    npts = 7
    shape = (26, 31, 67)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 15, 0.25)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(1e-3, 5e-3, 5).astype(np.float32)
    a, b = oracle_run(orc, npts, g, u, v, s, c0, c1, c2[:4], 4)
    a, b = oracle_run(orc, npts, g, a, b, s, c0, c1, c2[:3], 3)
    a, b = oracle_run(orc, npts, g, a, b, s, c0, c1, c2[:1], 1)
    v2 = orc.fill_uniform(g, 16, 1, halo=0.625)
    want_u, want_v = oracle_run(orc, npts, g, a, v2, s, c0, c1, c2, 5)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.set_tuning(steps_per_pass=4)
        ctx.leapfrog(4, c0, c1, c2[:4])
        ctx.set_tuning(steps_per_pass=3)
        ctx.leapfrog(3, c0, c1, c2[:3])
        ctx.set_tuning(steps_per_pass=1)
        ctx.leapfrog(1, c0, c1, c2[:1])
        ctx.set_tuning(steps_per_pass=4)
        ctx.upload("v", v2, R)
        ctx.leapfrog(5, c0, c1, c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    assert same_bits(gu, want_u)
    assert same_bits(gv, want_v)


def test_four_step_passes_equal_one_step_sweeps_on_gpu(gpu, orc):
    """Several tiles and chunks in every direction: 12 steps as four-step
    passes and as one-step sweeps give the same bits."""
    shape = (96, 100, 200)
    npts, R = 7, 1
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-3, 1e-3, 12).astype(np.float32)
    out = []
    for spp in (1, 4):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            ctx.init_uniform("u", 7, 0, -1.0, 1.0, 0.5)
            ctx.init_uniform("v", 7, 1, -1.0, 1.0, -0.5)
            ctx.init_uniform("s", 7, 2, -1.0, 1.0, 0.25)
            ctx.leapfrog(12, c0, c1, c2)
            out.append((ctx.download("u", R), ctx.download("v", R)))
    assert same_bits(out[0][0], out[1][0])
    assert same_bits(out[0][1], out[1][1])


def test_five_step_passes_equal_one_step_sweeps_non_cubic(gpu, orc):
    """A non-cubic grid with several tiles and chunks in every direction: 13
    steps as five-step passes (two passes and a three-step pass) and as
    one-step sweeps give the same bits for u and v, halos included."""
    shape = (150, 131, 260)
    npts, R = 7, 1
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-2e-3, 1e-3, 13).astype(np.float32)
    out = []
    for spp in (1, 5):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            ctx.init_uniform("u", 8, 0, -1.0, 1.0, 0.5)
            ctx.init_uniform("v", 8, 1, -1.0, 1.0, -0.5)
            ctx.init_uniform("s", 8, 2, -1.0, 1.0, 0.25)
            ctx.timer_start()
            ctx.leapfrog(13, c0, c1, c2)
            _, launches = ctx.timer_read()
            assert launches == (13 if spp == 1 else 3)
            out.append((ctx.download("u", R), ctx.download("v", R)))
    assert same_bits(out[0][0], out[1][0])
    assert same_bits(out[0][1], out[1][1])


@pytest.mark.parametrize("spp", [4, 5])
def test_four_and_five_steps_only_for_7_points(gpu, spp):
    for dtype in ("fp32", "fp16"):
        for npts in (13, 25):
            with gpu.Stencil(16, 16, 64, npts=npts, dtype=dtype) as ctx:
                with pytest.raises(gpu.SthError):
                    ctx.set_tuning(steps_per_pass=spp)
        with gpu.Stencil(16, 16, 64, npts=7, dtype=dtype) as ctx:
            ctx.set_tuning(steps_per_pass=spp)
            assert ctx.get_tuning()["steps_per_pass"] == spp
            with pytest.raises(gpu.SthError):
                ctx.set_tuning(steps_per_pass=6)


def test_three_steps_rejected_for_25_points_and_16bit(gpu):
    with gpu.Stencil(16, 16, 64, npts=25) as ctx:
        with pytest.raises(gpu.SthError):
            ctx.set_tuning(steps_per_pass=3)
    with gpu.Stencil(16, 16, 64, npts=13, dtype="bf16") as ctx:
        with pytest.raises(gpu.SthError):
            ctx.set_tuning(steps_per_pass=3)


def test_16bit_pass_lengths(gpu):
    """FP16 storage has passes of two steps for every stencil, three for 7
    and 13 points, four and five for 7 points; BF16 two and three steps for
    7 points and two for 13, none for 25."""
    with gpu.Stencil(16, 16, 64, npts=25, dtype="bf16") as ctx:
        with pytest.raises(gpu.SthError):
            ctx.set_tuning(steps_per_pass=2)
    for dtype, lengths in (("fp16", ((25, 2), (13, 3), (7, 5))),
                           ("bf16", ((13, 2), (7, 3)))):
        for npts, longest in lengths:
            with gpu.Stencil(16, 16, 64, npts=npts, dtype=dtype) as ctx:
                for spp in range(2, longest + 1):
                    ctx.set_tuning(steps_per_pass=spp)
                    assert ctx.get_tuning()["steps_per_pass"] == spp
                with pytest.raises(gpu.SthError):
                    ctx.set_tuning(steps_per_pass=longest + 1)


@pytest.mark.parametrize("dtype,npts,shape,spp", [
    ("fp16", 7, (40, 130, 300), 3), ("fp16", 7, (40, 70, 300), 1), ("fp16", 25, (40, 130, 300), 1),
    ("bf16", 7, (40, 130, 300), 2), ("bf16", 13, (40, 130, 300), 1)])
def test_16bit_default_schedule(gpu, orc, dtype, npts, shape, spp):
    """16-bit LEAPFROG's default pass length follows the grid's output tiles
    of each pass's shape: FP16 7 points three steps from 10 tiles (four from
    200; 13 points two steps from 70 tiles, three from 200), 25 points one
    step; BF16 7 points two steps from 8 tiles (three from 50), 13 points two
    from 30.  Five steps are then the passes and the steps left over,
    bit-identical to the oracle's emulation of the same schedule."""
    fmt = 1 if dtype == "fp16" else 2
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 31, 0.0)  # the 16-bit form needs zero halos
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.25, -0.25, 5).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        assert ctx.get_tuning()["steps_per_pass"] == spp
        ctx.upload_raw16("u", qs[0], exps[0], R)
        ctx.upload_raw16("v", qs[1], exps[1], R)
        ctx.upload_raw16("s", qs[2], exps[2], R)
        ctx.leapfrog(5, c0, c1, c2)
        gu, eu, mu = ctx.download_raw16("u", R)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 5, plen=spp)
    want = a if newest == 0 else b
    assert eu == exps[newest]
    assert np.array_equal(gu[g.interior], want[g.interior])


@pytest.mark.parametrize("npts,n,spp", [(7, 12, 1), (7, 48, 5), (13, 112, 1), (13, 128, 3),
                                         (25, 256, 1), (25, 320, 2)])
def test_default_schedule_by_grid_size(gpu, orc, npts, n, spp):
    """The default FP32 LEAPFROG schedule takes several steps per pass where
    that measured faster (enough output tiles and planes for the stencil;
    five steps for 7 points, three for 13, two for 25): twelve steps are
    passes (the remainder one shorter pass) or twelve sweeps, and either way
    bit-identical to the oracle."""
    R = orc.radius(npts)
    shape = (n, n, n)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 23, 0.125)
    c0, c1 = coeffs(orc, npts)
    c2 = np.full(12, 1e-3, np.float32)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        assert ctx.get_tuning()["steps_per_pass"] == spp
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.timer_start()
        ctx.leapfrog(12, c0, c1, c2)
        _, launches = ctx.timer_read()
        gu = ctx.download("u", R)
    assert launches == 12 // spp + (1 if 12 % spp else 0)
    if n <= 160:  # the oracle follows these sizes in seconds
        want_u, _ = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, 12)
        assert same_bits(gu, want_u)


@pytest.mark.parametrize("npts,shape", [(7, (256, 240, 272)), (13, (256, 240, 272)),
                                        (25, (160, 330, 340))])
def test_default_schedule_matches_oracle_mid_size(gpu, orc, npts, shape):
    """The default FP32 LEAPFROG schedule (five, three or two steps per pass)
    on grids of tens of tiles and several i-chunks (the 25-point pass needs
    50 tiles), twelve steps with a source term and nonzero halos:
    bit-identical to the oracle, halos included."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 29, 0.25)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-3e-3, 2e-3, 12).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        spp = ctx.get_tuning()["steps_per_pass"]
        assert spp == {7: 5, 13: 3, 25: 2}[npts]
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.leapfrog(12, c0, c1, c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    want_u, want_v = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, 12)
    assert same_bits(gu, want_u)
    assert same_bits(gv, want_v)


@pytest.mark.parametrize("npts,shape,steps", [(25, (128, 640, 816), 5), (13, (96, 640, 816), 7),
                                              (7, (128, 288, 1056), 11)])
def test_more_tiles_than_cus(gpu, orc, npts, shape, steps):
    """More output tiles than CUs (272 or 264 at the default shapes), so the
    work items of a pass take several rounds of the grid, with the default
    and with an explicit i-chunk count: bit-identical to the oracle."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 31, 0.25)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-3, 2e-3, steps).astype(np.float32)
    want_u, want_v = oracle_run(orc, npts, g, u, v, s, c0, c1, c2, steps)
    for chunks in (0, 3):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(chunks=chunks)
            ctx.upload("u", u, R)
            ctx.upload("v", v, R)
            ctx.upload("s", s, R)
            ctx.timer_start()
            ctx.leapfrog(steps, c0, c1, c2)
            _, launches = ctx.timer_read()
            assert launches == 3  # two passes + one one-step sweep
            gu = ctx.download("u", R)
            gv = ctx.download("v", R)
        assert same_bits(gu, want_u), chunks
        assert same_bits(gv, want_v), chunks


@pytest.mark.parametrize("npts,spp", [(7, 2), (7, 4), (13, 2), (13, 3)])
def test_fp16_pass_across_calls_and_autotune(gpu, orc, npts, spp):
    """FP16 LEAPFROG split over calls of 3, 2 and 1 steps with passes of spp
    steps (each call its own passes and the steps left over as one shorter
    pass or a step), then autotuned (the tuner also times the FP16 passes
    and the taller packed tiles) and stepped on: bit-identical to the
    oracle's emulation of the same schedule."""
    shape = (24, 70, 200)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 53, 0.0)
    s *= np.float32(1e-3)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(1, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.3, -0.3, 6).astype(np.float32)
    a, b = qs[0].copy(), qs[1].copy()
    bufs, cur = [a, b], 0
    with gpu.Stencil(*shape, npts=npts, dtype="fp16") as ctx:
        ctx.set_tuning(steps_per_pass=spp)
        for f, q, ex in zip(("u", "v", "s"), qs, exps):
            ctx.upload_raw16(f, q, ex, R)
        done = 0
        for k in (3, 2, 1):
            ctx.leapfrog(k, c0, c1, c2[done:done + k])
            e3 = [exps[cur], exps[1 - cur], exps[2]]
            m3 = [maxs[cur], maxs[1 - cur], maxs[2]]
            n = orc.leapfrog_q16(npts, 1, g, bufs[cur], bufs[1 - cur], qs[2], e3, m3,
                                 np.float32(c0), np.float32(c1), c2[done:done + k], k,
                                 plen=spp)
            exps[cur], exps[1 - cur] = e3[0], e3[1]
            maxs[cur], maxs[1 - cur] = m3[0], m3[1]
            cur = cur if n == 0 else 1 - cur
            done += k
        gu, eu, _ = ctx.download_raw16("u", R)
        assert eu == exps[cur]
        assert np.array_equal(gu[g.interior], bufs[cur][g.interior])
        # autotune (it steps the fields), then start again from the same fields
        t, ms = ctx.autotune(reps=1)
        assert ms > 0
        tuned = ctx.get_tuning()["steps_per_pass"]
        for f, q, ex in zip(("u", "v", "s"), qs, (orc.scale_exp(float(np.abs(x).max()))
                                                 for x in (u, v, s))):
            ctx.upload_raw16(f, q, ex, R)
        ctx.leapfrog(5, c0, c1, c2[:5])
        gu, eu, _ = ctx.download_raw16("u", R)
    e0 = [orc.scale_exp(float(np.abs(x).max())) for x in (u, v, s)]
    m0 = [orc.quantize(1, g, x, e)[1] for x, e in zip((u, v, s), e0)]
    a, b = qs[0].copy(), qs[1].copy()
    n = orc.leapfrog_q16(npts, 1, g, a, b, qs[2], e0, m0, np.float32(c0), np.float32(c1),
                         c2[:5], 5, plen=tuned)
    assert eu == e0[n]
    assert np.array_equal(gu[g.interior], (a if n == 0 else b)[g.interior])

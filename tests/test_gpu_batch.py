"""GPU tests of the batching form: many uneven shots advanced by one
persistent launch per step (or one plain launch per shot per step), with or
without a hipGraph around the time loop.  Every shot must be bit-identical to
the CPU oracle's LEAPFROG on the same fields, and every method to every other."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [(16, 20, 70), (33, 17, 65), (9, 40, 130), (24, 24, 24), (50, 8, 64)]


def shot_fields(orc, dims, npts, seed):
    R = orc.radius(npts)
    out = []
    for q, d in enumerate(dims):
        g = orc.Grid(*d, R, R, R)
        a = orc.fill_uniform(g, seed + q, 0)
        b = orc.fill_uniform(g, seed + q, 1)
        s = orc.fill_uniform(g, seed + q, 2)
        out.append((g, a, b, s))
    return out


def load(batch, fields, R):
    for q, (g, a, b, s) in enumerate(fields):
        sh = batch.shot(q)
        sh.upload("u", a, R)
        sh.upload("v", b, R)
        sh.upload("s", s, R)


def coeffs(orc, npts, c1=0.2):
    return float(np.float32(2.0 - orc.lap_gain(npts) * c1)), float(np.float32(c1))


@pytest.mark.parametrize("method,graph", [(0, False), (1, False), (0, True), (1, True)])
@pytest.mark.parametrize("npts", [13, 25])
def test_batch_matches_oracle(gpu, orc, method, graph, npts):
    R = orc.radius(npts)
    fields = shot_fields(orc, DIMS, npts, 10)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-2, 1e-2, 6).astype(np.float32)
    with gpu.Batch(DIMS, npts=npts) as bt:
        bt.set_tuning(2, 4, planes=7)
        load(bt, fields, R)
        bt.leapfrog(6, c0, c1, c2, method=method, graph=graph)
        got = [(bt.shot(q).download("u", R), bt.shot(q).download("v", R))
               for q in range(len(DIMS))]
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, 6)
        wu, wv = (a, b) if newest == 0 else (b, a)
        I = g.interior
        assert np.array_equal(got[q][0][I].view(np.uint32), wu[I].view(np.uint32)), q
        assert np.array_equal(got[q][1][I].view(np.uint32), wv[I].view(np.uint32)), q


def test_persistent_equals_sequential_config_shapes(gpu, orc):
    """Eight shots with sides from the batching workload's set
    {96, 128, 192, 256, 320, 384}, 12 steps: the
    persistent, sequential, two- and three-steps-per-pass and graphed runs
    agree bit for bit."""
    rng = np.random.default_rng(2)
    dims = [tuple(int(x) for x in rng.choice([96, 128, 192], 3)) for _ in range(8)]
    npts, R = 13, 2
    c0, c1 = coeffs(orc, npts)
    results = []
    for method, graph in ((0, False), (1, False), (0, True), (gpu.BATCH_TWO_STEP, False),
                          (gpu.BATCH_TWO_STEP, True), (gpu.BATCH_THREE_STEP, False),
                          (gpu.BATCH_THREE_STEP, True)):
        with gpu.Batch(dims, npts=npts) as bt:
            for q in range(len(dims)):
                bt.shot(q).init_gaussian("u", 1.0, 6.0, 1e-3, 1, q)
                bt.shot(q).init_gaussian("v", 1.0, 6.0, 1e-3, 1, 100 + q)
                bt.shot(q).init_point("s", dims[q][0] // 2, dims[q][1] // 3, dims[q][2] // 4, 1.0)
            bt.leapfrog(12, c0, c1, np.full(12, 1e-3, np.float32), method=method, graph=graph)
            results.append([bt.shot(q).download("u", R) for q in range(len(dims))])
    for r in results[1:]:
        for q in range(len(dims)):
            assert np.array_equal(r[q].view(np.uint32), results[0][q].view(np.uint32))


def test_graph_replays_match_eager(gpu, orc):
    """Repeated graphed calls, with eager work and synchronisation between
    them, give the same fields as the same calls made eagerly."""
    npts, R = 13, 2
    c0, c1 = coeffs(orc, npts)
    outs = []
    for graph in (False, True):
        fields = shot_fields(orc, DIMS[:3], npts, 40)
        with gpu.Batch(DIMS[:3], npts=npts) as bt:
            load(bt, fields, R)
            for rep in range(3):
                bt.leapfrog(4, c0, c1, np.full(4, 1e-3 * (rep + 1), np.float32), graph=graph)
                bt.sync()
                # eager work between replays: norms on every shot
                for q in range(3):
                    bt.shot(q).norms("u")
            outs.append([bt.shot(q).download("u", R) for q in range(3)])
    for q in range(3):
        assert np.array_equal(outs[0][q].view(np.uint32), outs[1][q].view(np.uint32))


@pytest.mark.parametrize("tile,planes", [((1, 4, 1), 0), ((4, 4, 1), 16), ((8, 2, 1), 5),
                                         ((2, 8, 1), 33), ((2, 4, 2), 9), ((4, 8, 2), 0)])
def test_batch_tunings_agree(gpu, orc, tile, planes):
    npts, R = 13, 2
    fields = shot_fields(orc, DIMS, npts, 70)
    c0, c1 = coeffs(orc, npts)
    with gpu.Batch(DIMS, npts=npts) as bt:
        bt.set_tuning(tile[0], tile[1], planes=planes, vec=tile[2])
        load(bt, fields, R)
        bt.leapfrog(3, c0, c1)
        got = [bt.shot(q).download("u", R) for q in range(len(DIMS))]
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, None, 3)
        want = a if newest == 0 else b
        assert np.array_equal(got[q][g.interior].view(np.uint32), want[g.interior].view(np.uint32))


def test_default_tuning_large_shot_matches_single_grid(gpu, orc):
    """A shot large enough for the single-grid size-dependent default tile
    (here 8 x 1040 x 1040) under the batch's default tuning: persistent and
    sequential batch steps equal a plain single-grid run, bit for bit."""
    dims = [(8, 1040, 1040), (9, 20, 70)]
    npts, R = 13, 2
    c0, c1 = coeffs(orc, npts)
    want = []
    for d in dims:
        with gpu.Stencil(*d, npts=npts) as ctx:
            ctx.set_tuning(steps_per_pass=1)
            ctx.init_uniform("u", 9, 0, -1.0, 1.0, 0.0)
            ctx.init_uniform("v", 9, 1, -1.0, 1.0, 0.0)
            ctx.init_uniform("s", 9, 2, -1.0, 1.0, 0.0)
            ctx.leapfrog(3, c0, c1, 1e-3)
            want.append(ctx.download("u", R))
    for method in (0, 1):
        with gpu.Batch(dims, npts=npts) as bt:
            for q in range(len(dims)):
                bt.shot(q).init_uniform("u", 9, 0, -1.0, 1.0, 0.0)
                bt.shot(q).init_uniform("v", 9, 1, -1.0, 1.0, 0.0)
                bt.shot(q).init_uniform("s", 9, 2, -1.0, 1.0, 0.0)
            bt.leapfrog(3, c0, c1, np.full(3, 1e-3, np.float32), method=method)
            for q in range(len(dims)):
                got = bt.shot(q).download("u", R)
                assert np.array_equal(got.view(np.uint32), want[q].view(np.uint32)), (method, q)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("steps", [5, 6])
@pytest.mark.parametrize("npts", [7, 13, 25])
def test_two_step_batch_matches_oracle(gpu, orc, graph, steps, npts):
    """Two LEAPFROG steps per pass over all shots (an odd last step runs
    one-step): every shot bit-identical to the oracle, halos included."""
    R = orc.radius(npts)
    fields = shot_fields(orc, DIMS, npts, 40)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-2, 1e-2, steps).astype(np.float32)
    with gpu.Batch(DIMS, npts=npts) as bt:
        load(bt, fields, R)
        bt.timer_start()
        bt.leapfrog(steps, c0, c1, c2, method=gpu.BATCH_TWO_STEP, graph=graph)
        _, launches = bt.timer_read()
        got = [(bt.shot(q).download("u", R), bt.shot(q).download("v", R))
               for q in range(len(DIMS))]
    assert launches == steps // 2 + steps % 2
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, steps)
        wu, wv = (a, b) if newest == 0 else (b, a)
        assert np.array_equal(got[q][0].view(np.uint32), wu.view(np.uint32)), q
        assert np.array_equal(got[q][1].view(np.uint32), wv.view(np.uint32)), q


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("steps", [6, 7, 8])
@pytest.mark.parametrize("npts", [7, 13])
def test_three_step_batch_matches_oracle(gpu, orc, graph, steps, npts):
    """Three LEAPFROG steps per pass over all shots (the one or two steps
    left over run one-step): every shot bit-identical to the oracle, halos
    included."""
    R = orc.radius(npts)
    fields = shot_fields(orc, DIMS, npts, 60)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-2, 1e-2, steps).astype(np.float32)
    with gpu.Batch(DIMS, npts=npts) as bt:
        load(bt, fields, R)
        bt.timer_start()
        bt.leapfrog(steps, c0, c1, c2, method=gpu.BATCH_THREE_STEP, graph=graph)
        _, launches = bt.timer_read()
        got = [(bt.shot(q).download("u", R), bt.shot(q).download("v", R))
               for q in range(len(DIMS))]
    assert launches == steps // 3 + steps % 3
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, steps)
        wu, wv = (a, b) if newest == 0 else (b, a)
        assert np.array_equal(got[q][0].view(np.uint32), wu.view(np.uint32)), q
        assert np.array_equal(got[q][1].view(np.uint32), wv.view(np.uint32)), q


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("npts,spp", [(7, 5), (13, 3), (25, 2)])
def test_fused_batch_matches_oracle(gpu, orc, graph, npts, spp):
    """STH_BATCH_FUSED: passes of the stencil's single-grid default length
    over all shots, the steps left over one-step; bit-identical to the
    oracle, halos included."""
    steps = 12
    R = orc.radius(npts)
    fields = shot_fields(orc, DIMS, npts, 70)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-2, 1e-2, steps).astype(np.float32)
    with gpu.Batch(DIMS, npts=npts) as bt:
        load(bt, fields, R)
        bt.timer_start()
        bt.leapfrog(steps, c0, c1, c2, method=gpu.BATCH_FUSED, graph=graph)
        _, launches = bt.timer_read()
        got = [(bt.shot(q).download("u", R), bt.shot(q).download("v", R))
               for q in range(len(DIMS))]
    assert launches == steps // spp + steps % spp
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, steps)
        wu, wv = (a, b) if newest == 0 else (b, a)
        assert np.array_equal(got[q][0].view(np.uint32), wu.view(np.uint32)), q
        assert np.array_equal(got[q][1].view(np.uint32), wv.view(np.uint32)), q


def test_three_step_batch_rejected_for_25_points(gpu):
    with gpu.Batch(DIMS[:2], npts=25) as bt:
        with pytest.raises(gpu.SthError):
            bt.leapfrog(3, 0.08, 0.16, method=gpu.BATCH_THREE_STEP)


def test_two_step_batch_mixed_with_other_methods(gpu, orc):
    """Calls of every method in a row (the two-step calls swap each shot's
    buffers with its twins), an upload in between: still the oracle's bits."""
    npts = 13
    R = orc.radius(npts)
    fields = shot_fields(orc, DIMS, npts, 50)
    c0, c1 = coeffs(orc, npts)
    plan = [(3, gpu.BATCH_TWO_STEP, False), (2, gpu.BATCH_PERSISTENT, False),
            (5, gpu.BATCH_TWO_STEP, True), (1, gpu.BATCH_SEQUENTIAL, False),
            (4, gpu.BATCH_TWO_STEP, False), (7, gpu.BATCH_THREE_STEP, False),
            (3, gpu.BATCH_THREE_STEP, True), (5, gpu.BATCH_THREE_STEP, False),
            (4, gpu.BATCH_FUSED, False)]
    c2 = np.linspace(-5e-3, 5e-3, sum(p[0] for p in plan)).astype(np.float32)
    new_v = orc.fill_uniform(fields[1][0], 77, 1, halo=0.5)
    with gpu.Batch(DIMS, npts=npts) as bt:
        load(bt, fields, R)
        n = 0
        for k, (steps, method, graph) in enumerate(plan):
            bt.leapfrog(steps, c0, c1, c2[n:n + steps], method=method, graph=graph)
            n += steps
            if k == 1:
                bt.sync()
                bt.shot(1).upload("v", new_v, R)
        got = [bt.shot(q).download("u", R) for q in range(len(DIMS))]
    for q, (g, a, b, s) in enumerate(fields):
        a, b = a.copy(), b.copy()
        n = 0
        for k, (steps, _, _) in enumerate(plan):
            newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2[n:n + steps], steps)
            if newest == 1:
                a, b = b, a  # keep the newest field in a
            n += steps
            if k == 1 and q == 1:
                b = new_v.copy()
        assert np.array_equal(got[q].view(np.uint32), a.view(np.uint32)), q


def test_batch_refuses_mixed_parities(gpu, orc):
    """A shot stepped on its own between batch calls holds u in its other
    buffer: the next batch call refuses (STH_ERR_STATE) instead of stepping
    that shot on swapped buffers; once every shot has the same parity again
    the batch runs and matches the oracle."""
    npts = 13
    R = orc.radius(npts)
    dims = DIMS[:3]
    fields = shot_fields(orc, dims, npts, 40)
    c0, c1 = coeffs(orc, npts)
    with gpu.Batch(dims, npts=npts) as bt:
        load(bt, fields, R)
        bt.leapfrog(2, c0, c1, None)
        bt.sync()
        bt.shot(1).leapfrog(1, c0, c1)
        bt.shot(1).sync()
        with pytest.raises(gpu.SthError) as e:
            bt.leapfrog(2, c0, c1, None)
        assert e.value.status == gpu.STH_ERR_STATE
        for q in (0, 2):
            bt.shot(q).leapfrog(1, c0, c1)
            bt.shot(q).sync()
        bt.leapfrog(3, c0, c1, None, method=gpu.BATCH_TWO_STEP)
        got = [bt.shot(q).download("u", R) for q in range(len(dims))]
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, None, 6)
        want = a if newest == 0 else b
        assert np.array_equal(got[q].view(np.uint32), want.view(np.uint32)), q


def test_batch_autotune_picks_fused_chunks_and_stays_exact(gpu, orc):
    """The batch autotuner times the one-step tiles and the fused pass's
    i-chunk counts, names the faster method, and a run with its choice
    matches the oracle per shot."""
    npts = 13
    R = orc.radius(npts)
    dims = [(40, 60, 100), (64, 48, 96), (30, 96, 128)]
    fields = shot_fields(orc, dims, npts, 50)
    c0, c1 = coeffs(orc, npts)
    c2 = np.linspace(-1e-2, 1e-2, 9).astype(np.float32)
    with gpu.Batch(dims, npts=npts) as bt:
        best, ms = bt.autotune(reps=3)
        assert ms > 0
        method = bt.best_method()
        assert method in (gpu.BATCH_PERSISTENT, gpu.BATCH_FUSED)
        assert best["chunks"] in (0, 1, 2, 3, 4, 6, 8, 12)
        assert bt.get_tuning()["chunks"] == best["chunks"]
        load(bt, fields, R)
        bt.leapfrog(9, c0, c1, c2, method=method, graph=True)
        got = [bt.shot(q).download("u", R) for q in range(len(dims))]
    for q, (g, a, b, s) in enumerate(fields):
        newest = orc.leapfrog_f32(npts, g, a, b, s, c0, c1, c2, 9)
        want = a if newest == 0 else b
        assert np.array_equal(got[q].view(np.uint32), want.view(np.uint32)), q


def test_graph_segments_match_eager(gpu, orc):
    """Graphed calls of more than 1024 launches are captured in several
    graphs: 48 shots x 23 sequential steps (1104 launches: 21 steps, then 2),
    and 1027 five-step passes plus 3 persistent steps on two small 7-point
    shots (1030 launches: 1024 passes, then 3 passes and the steps).  Each
    equals the same call made eagerly, bit for bit."""
    npts, R = 7, 1
    c0, c1 = coeffs(orc, npts)
    dims = [(8 + q % 5, 10 + q % 3, 16 + q % 7) for q in range(48)]
    cases = [(dims, 23, gpu.BATCH_SEQUENTIAL), (dims[:2], 5 * 1027 + 3, gpu.BATCH_FUSED)]
    for dd, steps, method in cases:
        c2 = np.linspace(-1e-3, 1e-3, steps).astype(np.float32)
        outs = []
        for graph in (False, True):
            fields = shot_fields(orc, dd, npts, 70)
            with gpu.Batch(dd, npts=npts) as bt:
                load(bt, fields, R)
                bt.timer_start()
                bt.leapfrog(steps, c0, c1, c2, method=method, graph=graph)
                _, launches = bt.timer_read()
                outs.append([(bt.shot(q).download("u", R), bt.shot(q).download("v", R))
                             for q in range(len(dd))])
            assert launches > 1024
        for q in range(len(dd)):
            for f in range(2):
                assert np.array_equal(outs[0][q][f].view(np.uint32), outs[1][q][f].view(np.uint32)), \
                    (method, q, f)


def test_config5_shape_autotuned_matches_single_grids(gpu, orc):
    """Config 5's exact batch: 64 13-point shots, sides drawn from seed 2
    among {96 ... 384} per axis (non-cubic), a Ricker point source each,
    tiling and method chosen by the autotuner, the time loop in hipGraphs.
    Every shot's u after ten steps equals a single-grid run of the same
    shot (each path is bit-identical to the oracle's LEAPFROG on smaller
    grids), compared on the GPU (max |x - y| = 0)."""
    rng = np.random.default_rng(2)
    dims = [tuple(int(x) for x in rng.choice([96, 128, 192, 256, 320, 384], 3))
            for _ in range(64)]
    srcs = [tuple(int(rng.integers(0, d)) for d in dd) for dd in dims]
    c0, c1 = coeffs(orc, 13, 0.16)
    c2 = gpu.ricker(10, 1e-3, 15.0)

    def init(ctx, q):
        ctx.init_uniform("u", 40 + q, 0)
        ctx.init_uniform("v", 40 + q, 1)
        ctx.init_point("s", *srcs[q], 1.0)

    with gpu.Batch(dims, npts=13) as bt:
        bt.autotune(reps=1)
        method = bt.best_method()
        for q in range(len(dims)):
            init(bt.shot(q), q)
        bt.leapfrog(10, c0, c1, c2, method=method, graph=True)
        bt.sync()
        for q, d in enumerate(dims):
            with gpu.Stencil(*d, npts=13) as one:
                init(one, q)
                one.leapfrog(10, c0, c1, c2)
                mr, _ = bt.shot(q).compare("u", one)
                assert mr == 0.0, (q, d)

"""REF sweeps per launch (sth_tuning.ref_sweeps): iterate_3d updates v in
place from read-only u and s (iterators.cpp:130-157), so a launch can apply
k sweeps to each point in registers.  The result must be bit-identical to k
separate sweeps (the oracle's REF), in FP32 and FP64, for every stencil, for
sweep counts that are and are not multiples of k, and at the headline size
(the reference's known-answer check)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(20, 24, 36), (9, 70, 130)]


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("shape", SHAPES)
def test_ref_sweeps_per_launch_bit_exact(gpu, orc, npts, shape):
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v = orc.fill_uniform(g, 3, 0, halo=0.5)
    s = orc.fill_uniform(g, 3, 1, halo=0.25)
    u = orc.fill_uniform(g, 3, 2, halo=-0.5)
    want = v.copy()
    orc.ref_sweep_f32(npts, g, want, s, u, 7)
    I = g.interior
    for k in (2, 3, 7, 9):
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.set_tuning(ref_sweeps=k)
            assert ctx.get_tuning()["ref_sweeps"] == k
            for f, a in (("v", v), ("s", s), ("u", u)):
                ctx.upload(f, a, R)
            ctx.timer_start()
            ctx.iterate(7)
            _, launches = ctx.timer_read()
            got = ctx.download("v", R)
        assert launches == -(-7 // k), k
        assert same_bits(got[I], want[I]), k


@pytest.mark.parametrize("npts", [7, 25])
def test_fp64_ref_sweeps_per_launch_bit_exact(gpu, orc, npts):
    shape = (13, 22, 45)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, s, v = (orc.fill_uniform(g, 8, f).astype(np.float64) for f in range(3))
    want = v.copy()
    orc.ref_sweep_f64(npts, g, want, s, u, 5, coeff_mode=0)
    with gpu.Stencil(*shape, npts=npts, dtype="fp64") as ctx:
        ctx.set_tuning(ref_sweeps=4)
        for f, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(f, x, R)
        ctx.iterate(5)
        got = ctx.download("v", R, dtype=np.float64)
    assert np.array_equal(got[g.interior], want[g.interior])


def test_ref_sweeps_per_launch_kat_at_headline_size(gpu):
    """The reference's known-answer run at 1040^3 (u = s = 1, v = 0, 7
    points: v = #sweeps exactly), twenty sweeps in two launches."""
    n = 1040
    with gpu.Stencil(n, n, n, npts=7) as ctx:
        ctx.set_tuning(ref_sweeps=10)
        ctx.init_const(0.0, 1.0, 1.0)
        ctx.timer_start()
        ctx.iterate(20)
        _, launches = ctx.timer_read()
        assert launches == 2
        assert ctx.check(20.0) == 0
        mx, _ = ctx.norms("v")
    assert mx == 20.0


def test_ref_sweeps_out_of_range_is_an_error(gpu):
    with gpu.Stencil(16, 16, 64, npts=7) as ctx:
        with pytest.raises(gpu.SthError):
            ctx.set_tuning(ref_sweeps=4097)

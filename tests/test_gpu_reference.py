"""GPU parity against the reference's own results and the configs' exact
setups, through the C ABI:

* the committed golden vectors (tests/golden: outputs of the compiled
  reference) fed through sth_upload / sth_download with the reference's own
  host halos (the wide-halo cases hold --xh 4 --yh 4 --zh 16 with nonzero
  halo values), FP32 and the -DUSE_DOUBLE build (STH_WEIGHTS_REFERENCE);
* the reference's KAT at 1040^3 in FP64 with the USE_DOUBLE weights;
* the device initialisers the benchmark runs on (Gaussian pulse + noise,
  point source) against the oracle's;
* config 1's exact run (7 points, 128^3, halo 1 = 0, seed 0, 20 steps).
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def cases():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["cases"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden_inputs(orc, c):
    """The case's inputs, regenerated from the oracle's seeded generator and
    checked against the hashes the manifest recorded."""
    g = orc.Grid(*c["shape"], *c["halo"])
    hu, hs, hv = c["halos"]
    u = orc.fill_uniform(g, c["seed"], 0, halo=hu)
    s = orc.fill_uniform(g, c["seed"], 1, halo=hs)
    v = orc.fill_uniform(g, c["seed"], 2, halo=hv)
    if c["dtype"] == "float64":
        u, s, v = u.astype(np.float64), s.astype(np.float64), v.astype(np.float64)
    assert {"u": sha(u), "s": sha(s), "v": sha(v)} == c["inputs_sha256"], c["name"]
    return g, u, s, v


@pytest.mark.parametrize("case", cases(), ids=lambda c: c["name"])
def test_golden_vectors_through_the_boundary(gpu, orc, case):
    """REF sweeps on the GPU from the golden inputs in the reference's host
    layout (their own halo widths), bit-identical to the compiled reference's
    output.  FP64 cases run with the reference's -DUSE_DOUBLE weights."""
    g, u, s, v = golden_inputs(orc, case)
    want = np.fromfile(os.path.join(GOLDEN, case["file"]), dtype=case["dtype"])
    assert sha(want) == case["output_sha256"]
    f64 = case["dtype"] == "float64"
    with gpu.Stencil(*case["shape"], npts=case["npts"], dtype="fp64" if f64 else "fp32") as ctx:
        if f64:
            ctx.set_weights("reference")
            assert ctx.get_weights() == "reference"
        for name, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(name, x, case["halo"])
        ctx.iterate(case["sweeps"])
        out = np.full(g.shape, -7.0, dtype=want.dtype)  # cells the download must not touch
        ctx.download("v", case["halo"], out=out)
    got = np.ascontiguousarray(out[g.interior])
    assert got.tobytes() == want.tobytes(), case["name"]
    # cells beyond the device's radius-wide halo are left as they were
    R = orc.radius(case["npts"])
    xh, yh, zh = case["halo"]
    if zh > R:
        assert np.all(out[:, :, : zh - R] == -7.0)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_fp64_default_weights_are_the_fp32_constants(gpu, orc, npts):
    """The default FP64 weights differ from the USE_DOUBLE ones (13 and 25
    points): the same golden inputs give the oracle's coeff_mode 0 result."""
    c = next(x for x in cases() if x["npts"] == npts and x["dtype"] == "float64")
    g, u, s, v = golden_inputs(orc, c)
    with gpu.Stencil(*c["shape"], npts=npts, dtype="fp64") as ctx:
        assert ctx.get_weights() == "fp32"
        for name, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(name, x, c["halo"])
        ctx.iterate(c["sweeps"])
        got = ctx.download("v", c["halo"], dtype=np.float64)
    want = v.copy()
    orc.ref_sweep_f64(npts, g, want, s, u, c["sweeps"], coeff_mode=0)
    assert np.array_equal(got[g.interior], want[g.interior])
    ref = np.fromfile(os.path.join(GOLDEN, c["file"]), dtype=np.float64)
    assert (np.array_equal(got[g.interior].ravel(), ref)) == (npts == 7)


def test_reference_weights_need_fp64(gpu):
    with gpu.Stencil(8, 8, 8, npts=13) as ctx:
        with pytest.raises(gpu.SthError) as e:
            ctx.set_weights("reference")
        assert e.value.status == gpu.STH_ERR_UNSUPPORTED


@pytest.mark.parametrize("npts", [7, 25])
def test_fp64_reference_kat_at_1040(gpu, orc, npts):
    """The reference's -DUSE_DOUBLE run at the headline size: init_arrays
    (v=0, u=s=1 everywhere), one warm-up and 3 more sweeps; its check
    passes, and the largest |v| is the value the oracle's USE_DOUBLE sweep
    gives every point of a small grid (with u = s = 1 the stencil value is a
    constant, so every interior point follows the same recurrence)."""
    n = 1040
    R = orc.radius(npts)
    g = orc.Grid(6, 6, 6, R, R, R)
    v, s, u = (np.full(g.shape, x, np.float64) for x in (0.0, 1.0, 1.0))
    orc.ref_sweep_f64(npts, g, v, s, u, 4, coeff_mode=1)
    kat = v[g.interior]
    assert np.all(kat == kat.flat[0])
    with gpu.Stencil(n, n, n, npts=npts, dtype="fp64") as ctx:
        ctx.set_weights("reference")
        ctx.init_const(0.0, 1.0, 1.0)
        ctx.iterate(1)
        assert ctx.check(1.0) == 0
        ctx.iterate(3)
        assert ctx.check(4.0) == 0
        mx, l2 = ctx.norms("v")
    assert mx == abs(kat.flat[0])  # the FP64 maximum, exact
    assert abs(l2 - abs(kat.flat[0]) * n ** 1.5) <= 1e-9 * l2


# ------------------------------------------------------------- initialisers
def ulp_diff32(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def ulp_diff64(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFFFFFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFFFFFFFFFF), ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("npts", [13, 25])
@pytest.mark.parametrize("shape", [(40, 33, 70), (25, 48, 130)])
def test_gaussian_init_matches_oracle(gpu, orc, npts, shape):
    """sth_init_gaussian (configs 3-4: pulse sigma 8 at the centre plus
    U[-1e-3, 1e-3) noise from seed 1) against orc_fill_gaussian: FP32 within
    one unit in the last place and at most 0.1 % of the points not
    bit-identical; FP64 within 4 ulps of the pulse term (the device's and the
    host's exp may differ in the last bits of a double; where pulse and noise
    nearly cancel, that is many ulps of the sum); the halo exactly zero."""
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    for field, fid in (("u", 0), ("v", 1)):
        want32 = orc.fill_gaussian(g, 1.0, 8.0, 1e-3, 1, fid)
        want64 = orc.fill_gaussian(g, 1.0, 8.0, 1e-3, 1, fid, dtype=np.float64)
        with gpu.Stencil(*shape, npts=npts) as ctx:
            ctx.init_gaussian(field, 1.0, 8.0, 1e-3, 1, fid)
            got32 = ctx.download(field, R)
        with gpu.Stencil(*shape, npts=npts, dtype="fp64") as ctx:
            ctx.init_gaussian(field, 1.0, 8.0, 1e-3, 1, fid)
            got64 = ctx.download(field, R, dtype=np.float64)
        d32 = ulp_diff32(got32, want32)
        assert d32.max() <= 1 and np.count_nonzero(d32) <= 1e-3 * d32.size
        assert np.all(np.abs(got64 - want64) <= 4 * np.finfo(np.float64).eps *
                      np.maximum(np.abs(want64), 1e-3))
        assert np.count_nonzero(ulp_diff64(got64, want64)) <= 0.05 * got64.size
        halo = np.ones(g.shape, bool)
        halo[g.interior] = False
        assert np.all(got32[halo] == 0.0) and np.all(got64[halo] == 0.0)
        assert got32.max() > 0.5  # the pulse is there


def test_point_source_init(gpu, orc):
    shape = (17, 20, 66)
    g = orc.Grid(*shape, 4, 4, 4)
    with gpu.Stencil(*shape, npts=25) as ctx:
        ctx.init_uniform("s", 3, 2, -1.0, 1.0, 0.5)  # overwritten entirely
        ctx.init_point("s", 8, 10, 33, 1.0)
        got = ctx.download("s", 4)
        with pytest.raises(gpu.SthError):
            ctx.init_point("s", 17, 0, 0, 1.0)
    want = g.zeros()
    want[g.interior][8, 10, 33] = 1.0
    assert np.array_equal(got, want)


# ------------------------------------------------------------ config 1 exactly
def test_config1_exact_run(gpu, orc):
    """Config 1: 7 points, 128^3 interior, 1-cell zero halo, u, v, s
    i.i.d. U[-1, 1) from seed 0, 20 steps; REF (what hstencil3d runs) and
    LEAPFROG with c1 = 0.25, c2 = 1e-6, c0 = 2 - 6 c1, FP32 and FP64:
    bit-identical to the oracle."""
    n, steps = 128, 20
    g = orc.Grid(n, n, n, 1, 1, 1)
    v, s, u = (orc.fill_uniform(g, 0, f) for f in range(3))
    # REF
    with gpu.Stencil(n, n, n, npts=7) as ctx:
        for name, x in (("v", v), ("s", s), ("u", u)):
            ctx.upload(name, x, 1)
        ctx.iterate(steps)
        got = ctx.download("v", 1)
    want = v.copy()
    orc.ref_sweep_f32(7, g, want, s, u, steps)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # LEAPFROG, FP32 (the default five-step passes) and FP64
    c1 = np.float32(0.25)
    c0 = np.float32(2.0 - 6.0 * 0.25)
    c2 = np.full(steps, 1e-6, np.float32)
    with gpu.Stencil(n, n, n, npts=7) as ctx:
        # the device's seeded init is the oracle's generator (same seed and
        # field ids: v 0, s 1, u 2)
        for name, fid in (("v", 0), ("s", 1), ("u", 2)):
            ctx.init_uniform(name, 0, fid, -1.0, 1.0, 0.0)
        for name, fid in (("v", 0), ("s", 1), ("u", 2)):
            assert np.array_equal(ctx.download(name, 1),
                                  orc.fill_uniform(g, 0, fid)), name
        ctx.leapfrog(steps, float(c0), float(c1), c2)
        gu = ctx.download("u", 1)
        gv = ctx.download("v", 1)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(7, g, a, b, s, c0, c1, c2, steps)
    want_u, want_v = (a, b) if newest == 0 else (b, a)
    assert np.array_equal(gu.view(np.uint32), want_u.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), want_v.view(np.uint32))
    with gpu.Stencil(n, n, n, npts=7, dtype="fp64") as ctx:
        for name, x in (("v", v), ("s", s), ("u", u)):
            ctx.upload(name, x.astype(np.float64), 1)
        ctx.leapfrog(steps, float(c0), float(c1), c2)
        gu64 = ctx.download("u", 1, dtype=np.float64)
    a, b = u.astype(np.float64), v.astype(np.float64)
    newest = orc.leapfrog_f64(7, g, a, b, s.astype(np.float64), float(c0), float(c1),
                              c2.astype(np.float64), steps)
    assert np.array_equal(gu64, a if newest == 0 else b)
    # the FP32 result against FP64: the physics check's error figure
    err = np.abs(gu.astype(np.float64) - gu64)
    assert err.max() / np.abs(gu64).max() < 1e-6

"""GPU tests of the hstencil3d_hip front end (the reference driver's CLI and
report, main.cpp:94-332) and of the library's diagnostics."""
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_fe(sth, *args):
    r = subprocess.run([sth.FRONTEND_PATH, *args], capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_frontend_reference_run_passes_its_check(gpu, npts):
    """The reference's default run: init_arrays, warm-up, timed sweeps, two
    checks with 0 errors, the same report line."""
    r = run_fe(gpu, "--xs", "64", "--ys", "64", "--zs", "96", "--iter", "10",
               "--stencil", str(npts))
    assert r.returncode == 0, r.stderr
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"]
    m = re.search(r"finish compute in ([0-9.]+) secs \(([0-9.]+) GFlop/s, mem: ([0-9.]+) GB/s, "
                  r"cache: ([0-9.]+) GB/s\)", r.stdout)
    assert m and float(m.group(1)) > 0
    assert "Gpts/s" in r.stdout


def test_frontend_leapfrog_fp16(gpu):
    r = run_fe(gpu, "--xs", "48", "--ys", "40", "--zs", "70", "--iter", "20", "--stencil", "25",
               "--mode", "leapfrog", "--init", "gauss", "--c1", "0.16", "--c0", "0.08",
               "--dtype", "fp16")
    assert r.returncode == 0, r.stderr
    m = re.search(r"norms: max \|u\| = ([0-9.e+-]+)", r.stdout)
    assert m and 0.0 < float(m.group(1)) < 10.0


def test_frontend_leapfrog_steps_per_pass_and_launch(gpu):
    """FP32 LEAPFROG from the CLI with one, two or three steps per pass and
    either launch form: the printed norms agree to the last digit."""
    outs = []
    for extra in (["--steps-per-pass", "1", "--launch", "1"], ["--steps-per-pass", "2"],
                  ["--steps-per-pass", "1", "--launch", "2"], ["--steps-per-pass", "3"]):
        r = run_fe(gpu, "--xs", "40", "--ys", "56", "--zs", "150", "--iter", "6", "--stencil",
                   "13", "--mode", "leapfrog", "--init", "gauss", "--c1", "0.2", *extra)
        assert r.returncode == 0, r.stderr
        m = re.search(r"norms: max \|u\| = ([0-9.e+-]+).*", r.stdout)
        assert m
        outs.append(m.group(0))
    assert outs[0] == outs[1] == outs[2] == outs[3]


def test_frontend_sweeps_per_launch(gpu):
    """The reference's run (REF, init_arrays, check) with ten sweeps per
    launch: both checks pass, the timed 20 sweeps take 2 launches and the
    reported time is per sweep."""
    r = run_fe(gpu, "--xs", "48", "--ys", "40", "--zs", "96", "--iter", "20", "--stencil", "13",
               "--sweeps-per-launch", "10")
    assert r.returncode == 0, r.stderr
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"]
    assert re.search(r"-- GPU: [0-9.]+ ms/sweep \(HIP events, 2 launches\)", r.stdout)


@pytest.mark.parametrize("stencil,spp", [("25", "3"), ("13", "4"), ("25", "5")])
def test_frontend_rejects_a_pass_the_stencil_lacks(gpu, stencil, spp):
    """A steps-per-pass value the stencil has no pass for is a status from
    sth_set_tuning: the front end reports it and exits 3, nothing runs."""
    r = run_fe(gpu, "--xs", "40", "--ys", "40", "--zs", "64", "--iter", "3", "--stencil", stencil,
               "--mode", "leapfrog", "--steps-per-pass", spp)
    assert r.returncode == 3
    assert "ERROR: tuning: unsupported combination" in r.stderr


def test_frontend_autotune(gpu):
    r = run_fe(gpu, "--xs", "40", "--ys", "40", "--zs", "64", "--iter", "3", "--stencil", "13",
               "--tune")
    assert r.returncode == 0, r.stderr
    assert "autotune:" in r.stdout
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"]


def test_stream_copy_copies(gpu, orc):
    g = orc.Grid(20, 22, 64, 1, 1, 1)
    x = orc.fill_uniform(g, 3, 0)
    with gpu.Stencil(20, 22, 64, npts=7) as ctx:
        ctx.upload("u", x, 1)
        ms = ctx.stream_copy(ctx.layout().bytes // 16 * 16, 2)
        y = ctx.download("v", 1)
    assert ms > 0
    assert np.array_equal(y[g.interior], x[g.interior])


def test_autotune_keeps_results(gpu, orc):
    """The tuner picks a schedule; a run with it still matches the oracle."""
    g = orc.Grid(40, 36, 100, 2, 2, 2)
    u, s, v = (orc.fill_uniform(g, 4, f) for f in range(3))
    with gpu.Stencil(40, 36, 100, npts=13) as ctx:
        best, ms = ctx.autotune(leapfrog=False, reps=1)
        assert ms > 0 and best["rows_per_wave"] in (1, 2, 4, 8)
        for f, a in (("u", u), ("s", s), ("v", v)):
            ctx.upload(f, a, 2)
        ctx.iterate(2)
        got = ctx.download("v", 2)
    orc.ref_sweep_f32(13, g, v, s, u, 2)
    assert np.array_equal(got[g.interior].view(np.uint32), v[g.interior].view(np.uint32))


@pytest.mark.parametrize("dtype", ["fp32", "fp16", "bf16", "fp64"])
def test_save_load_restart_is_exact(gpu, tmp_path, dtype):
    """Checkpoint after 5 steps, restore into a fresh context, run 4 more
    steps on both: identical fields (and manifest geometry checks)."""
    shape, npts = (18, 21, 70), 13
    c1 = 0.16
    c0 = 2.0 - 6.0 * c1
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as a:
        a.init_gaussian("u", 1.0, 3.0, 1e-3, 1, 0)
        a.init_gaussian("v", 1.0, 3.0, 1e-3, 1, 1)
        a.init_point("s", 9, 10, 35, 1.0)
        a.leapfrog(5, c0, c1, [1e-2] * 5)
        a.save(str(tmp_path))
        a.leapfrog(4, c0, c1, [1e-2] * 4)
        ua = a.download("u", 2, dtype=np.float64)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as b:
        b.load(str(tmp_path))
        b.leapfrog(4, c0, c1, [1e-2] * 4)
        ub = b.download("u", 2, dtype=np.float64)
    assert np.array_equal(ua, ub)
    with gpu.Stencil(shape[0] + 1, shape[1], shape[2], npts=npts, dtype=dtype) as c:
        with pytest.raises(gpu.SthError):
            c.load(str(tmp_path))


def test_autotune_cache_round_trip(gpu, tmp_path, monkeypatch):
    """With STH_TUNE_CACHE set, the first autotune stores its result and the
    second one on the same geometry applies it without timing."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("STH_TUNE_CACHE", str(cache))
    with gpu.Stencil(30, 34, 90, npts=25) as ctx:
        ctx.init_const(0.0, 1.0, 1.0)
        best, ms = ctx.autotune(leapfrog=False, reps=1)
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("sth-tune 2 25 0 0 30 34 90")
    with gpu.Stencil(30, 34, 90, npts=25) as ctx:
        best2, ms2 = ctx.autotune(leapfrog=False, reps=1)
        assert ctx.get_tuning()["rows_per_wave"] == best["rows_per_wave"]
    assert best2 == best and abs(ms2 - ms) < 1e-4
    assert len(cache.read_text().splitlines()) == 1


def test_tune_cache_three_step_lines(gpu, tmp_path, monkeypatch):
    """A stored three-steps-per-pass result is applied without timing where
    the stencil has that pass (13 points) and ignored where it has not (25
    points: the tuner times its candidates and appends its own line)."""
    cache = tmp_path / "tune.txt"
    cache.write_text("sth-tune 2 13 0 1 40 44 100 4 12 1 2 3 0.01000\n"
                     "sth-tune 2 25 0 1 30 34 90 3 16 1 2 3 0.01000\n")
    monkeypatch.setenv("STH_TUNE_CACHE", str(cache))
    with gpu.Stencil(40, 44, 100, npts=13) as ctx:
        best, ms = ctx.autotune(leapfrog=True, reps=1)
        t = ctx.get_tuning()
    assert (best["rows_per_wave"], best["waves"], best["steps_per_pass"]) == (4, 12, 3)
    assert ms == pytest.approx(0.01)
    assert (t["rows_per_wave"], t["waves"], t["chunks"], t["steps_per_pass"]) == (4, 12, 2, 3)
    with gpu.Stencil(30, 34, 90, npts=25) as ctx:
        ctx.init_const(0.0, 1.0, 1.0)
        best, ms = ctx.autotune(leapfrog=True, reps=1)
    lines = cache.read_text().splitlines()
    assert len(lines) == 3 and lines[2].startswith("sth-tune 2 25 0 1 30 34 90")
    assert best["steps_per_pass"] in (1, 2) and ms != pytest.approx(0.01)


def test_autotune_7pt_result_is_bit_exact(gpu, orc):
    """The autotuner times every candidate, the fused passes of two to five
    steps among them, keeps the fastest; twelve steps under its choice
    equal the oracle's bits."""
    npts, shape = 7, (40, 44, 100)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u = orc.fill_uniform(g, 31, 0, halo=0.25)
    v = orc.fill_uniform(g, 31, 1, halo=-0.25)
    s = orc.fill_uniform(g, 31, 2, halo=0.5)
    c1 = np.float32(0.25)
    c0 = float(np.float32(2.0 - orc.lap_gain(npts) * float(c1)))
    c2 = np.linspace(-2e-3, 2e-3, 12).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts) as ctx:
        best, ms = ctx.autotune(leapfrog=True, reps=1)
        assert ms > 0.0 and 1 <= best["steps_per_pass"] <= 5
        t = ctx.get_tuning()
        assert t["steps_per_pass"] == best["steps_per_pass"]
        ctx.upload("u", u, R)
        ctx.upload("v", v, R)
        ctx.upload("s", s, R)
        ctx.leapfrog(12, c0, float(c1), c2)
        gu = ctx.download("u", R)
        gv = ctx.download("v", R)
    a, b = u.copy(), v.copy()
    newest = orc.leapfrog_f32(npts, g, a, b, s, c0, float(c1), c2, 12)
    wu, wv = (a, b) if newest == 0 else (b, a)
    assert np.array_equal(gu.view(np.uint32), wu.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), wv.view(np.uint32))


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_frontend_host_fields_with_reference_halos(gpu, npts, dtype):
    """README.md:92's halos (--xh 7 --yh 7 --zh 16) on the reference's host
    arrays: init_arrays on the host over the whole allocation, upload with
    that layout, sweeps, download and the reference's check on the host
    array; FP64 is the -DUSE_DOUBLE build's arithmetic."""
    r = run_fe(gpu, "--xs", "40", "--ys", "26", "--zs", "70", "--xh", "7", "--yh", "7",
               "--zh", "16", "--iter", "10", "--stencil", str(npts), "--host-fields",
               "--dtype", dtype)
    assert r.returncode == 0, r.stderr
    assert "host arrays: halo 7x7x16" in r.stdout
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"]


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_frontend_host_fields_leapfrog_norms_agree(gpu, dtype):
    """A LEAPFROG run through the host arrays prints the same device norms as
    the same run on device-only fields, and the host's max |u| of the
    downloaded field equals the device's."""
    args = ["--xs", "30", "--ys", "33", "--zs", "80", "--iter", "7", "--stencil", "13",
            "--mode", "leapfrog", "--init", "gauss", "--c1", "0.2", "--dtype", dtype]
    a = run_fe(gpu, *args)
    b = run_fe(gpu, *args, "--host-fields", "--xh", "5", "--yh", "6", "--zh", "9")
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    na = re.search(r"norms: max \|u\| = ([0-9.e+-]+), \|\|u\|\|_2 = ([0-9.e+-]+)", a.stdout)
    nb = re.search(r"norms: max \|u\| = ([0-9.e+-]+), \|\|u\|\|_2 = ([0-9.e+-]+)", b.stdout)
    assert na and nb and na.groups() == nb.groups()
    h = re.search(r"host: max \|u\| = ([0-9.e+-]+)", b.stdout)
    assert h and h.group(1) == nb.group(1)


@pytest.mark.parametrize("dtype", ["fp32", "fp16"])
def test_frontend_tune_keeps_the_physics(gpu, dtype):
    """--tune times schedules on the live fields; the timed run must still
    start from the initial state: LEAPFROG on the Gaussian pulse prints the
    same norms with and without --tune."""
    args = ["--xs", "40", "--ys", "40", "--zs", "64", "--iter", "8", "--stencil", "25",
            "--mode", "leapfrog", "--init", "gauss", "--c1", "0.16", "--dtype", dtype]
    a = run_fe(gpu, *args)
    b = run_fe(gpu, *args, "--tune")
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    assert "autotune:" in b.stdout
    na = re.search(r"norms: .*", a.stdout)
    nb = re.search(r"norms: .*", b.stdout)
    assert na and nb and na.group(0) == nb.group(0)


@pytest.mark.parametrize("dtype,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("fp16", 1e-3)])
def test_frontend_leapfrog_reports_conserved_energy(gpu, dtype, tol):
    """A LEAPFROG run on the Gaussian pulse (s = 0, zero halos) prints the
    scheme's discrete energy before and after the timed steps; it is
    conserved up to the storage format's rounding."""
    r = run_fe(gpu, "--xs", "40", "--ys", "44", "--zs", "72", "--iter", "20", "--stencil", "25",
               "--mode", "leapfrog", "--init", "gauss", "--c1", "0.16", "--dtype", dtype)
    assert r.returncode == 0, r.stderr
    m = re.search(r"energy: before ([0-9.e+-]+), after ([0-9.e+-]+), relative change "
                  r"([0-9.e+-]+)", r.stdout)
    assert m, r.stdout
    q0, q1, d = map(float, m.groups())
    assert q0 > 0 and q1 > 0 and d <= tol
    assert abs(abs(q1 - q0) / q0 - d) <= 1e-3 * d + 1e-8  # printed to 10 digits


def test_frontend_default_c0_is_constant_preserving(gpu):
    """Without --c0 the 25-point LEAPFROG takes c0 = 2 - 12 c1 (the value
    sth_leapfrog_c0 gives), not 2 - 6 c1."""
    c0 = gpu.leapfrog_c0(25, 0.16)
    args = ["--xs", "36", "--ys", "30", "--zs", "70", "--iter", "6", "--stencil", "25",
            "--mode", "leapfrog", "--init", "gauss", "--c1", "0.16"]
    outs = []
    for extra in ([], ["--c0", repr(float(c0))], ["--c0", repr(2.0 - 6.0 * 0.16)]):
        r = run_fe(gpu, *args, *extra)
        assert r.returncode == 0, r.stderr
        outs.append(re.search(r"norms: .*", r.stdout).group(0))
    assert outs[0] == outs[1] and outs[0] != outs[2]


@pytest.mark.parametrize("lane", [4, 16])
def test_bw_probe_streams(gpu, orc, lane):
    """The bandwidth probe moves what it says: one read and one write copy u
    into v; three reads and one write store u + s + v into v."""
    g = orc.Grid(20, 22, 64, 1, 1, 1)
    u, s, v = (orc.fill_uniform(g, 9, f) for f in range(3))
    with gpu.Stencil(20, 22, 64, npts=7) as ctx:
        nbytes = ctx.layout().bytes // 16 * 16
        for name, x in (("u", u), ("s", s), ("v", v)):
            ctx.upload(name, x, 1)
        assert ctx.bw_probe(1, 1, lane, nbytes, reps=1) > 0
        y = ctx.download("v", 1)
        assert np.array_equal(y[g.interior], u[g.interior])
        ctx.upload("v", v, 1)
        ctx.bw_probe(3, 1, lane, nbytes, reps=1)  # warm-up + 1: v <- u + s + v twice
        y = ctx.download("v", 1)
        w = u + s + v
        w = u + s + w
        assert np.array_equal(y[g.interior], w[g.interior])
        assert ctx.bw_probe(2, 2, lane, nbytes, reps=1) > 0
        assert ctx.bw_probe(3, 0, lane, nbytes, reps=1) > 0
        with pytest.raises(gpu.SthError):
            ctx.bw_probe(0, 0, lane, nbytes)

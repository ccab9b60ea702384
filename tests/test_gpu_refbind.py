"""The reference's own driver on the GPU: main.cpp, grid.cpp, gilbert2d.cpp and
aux.cpp of the reference, compiled with stencil3d_amd/csrc/refbind/
iterators_hip.cpp in place of iterators.cpp and linked with libstencil_hip
(stencil3d_amd/_refbind, built where the reference's sources are present).

Its run is the reference's: init_arrays (u = s = 1, v = 0), one warm-up
ITERATE and its check, the timed ITERATE of --iter sweeps and the check
against iter + 1 (the known answer: each REF sweep adds the stencil's
constant 1 to v), and the same report line; the compiled CPU reference
(oracle/_ref) prints the same check lines on the same arguments."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "stencil3d_amd", "_refbind")
REF = os.path.join(ROOT, "oracle", "_ref")


def binary(npts, f64=False):
    path = os.path.join(BIN, f"hstencil3d_gpu_{npts}" + ("_f64" if f64 else ""))
    if not os.path.exists(path):
        pytest.skip(f"{path} is built only where the reference's sources are present")
    return path


def run(path, *args, env=None):
    return subprocess.run([path, *args], capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.parametrize("f64", [False, True], ids=["float", "double"])
@pytest.mark.parametrize("npts", [7, 13, 25])
def test_reference_driver_runs_on_the_gpu(gpu, npts, f64):
    r = run(binary(npts, f64), "--xs", "48", "--ys", "40", "--zs", "64", "--iter", "10")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"], r.stdout[-2000:]
    m = re.search(r"finish compute in ([0-9.]+) secs \(([0-9.]+) GFlop/s, mem: ([0-9.]+) GB/s",
                  r.stdout)
    assert m and float(m.group(1)) > 0


@pytest.mark.parametrize("npts", [7, 25])
def test_reference_driver_matches_the_cpu_reference(gpu, npts):
    """Same arguments (halos 7, 7, 16 as in the reference's README run, a
    non-default --iter): the GPU build and the CPU reference print the same
    grid description and check lines."""
    cpu = os.path.join(REF, f"hstencil3d_{npts}")
    if not os.path.exists(cpu):
        pytest.skip("the compiled CPU reference is built only where its sources are present")
    args = ["--xs", "64", "--ys", "64", "--zs", "96", "--xh", "7", "--yh", "7", "--zh", "16",
            "--iter", "7"]
    env = dict(os.environ, OMP_NUM_THREADS="8")
    g = run(binary(npts), *args, env=env)
    c = run(cpu, *args, env=env)
    assert g.returncode == 0 and c.returncode == 0, (g.stderr[-2000:], c.stderr[-2000:])

    def lines(out):
        return [l for l in out.splitlines()
                if l.startswith(("check:", "... grid dimensions", "... with ", " -- using a grid"))]

    assert lines(g.stdout) == lines(c.stdout)
    assert re.findall(r"check: (\d+) errors found", g.stdout) == ["0", "0"]


def test_reference_driver_reports_a_missing_argument_like_the_reference(gpu):
    """The CLI is the reference's own (aux.cpp): a value out of range exits
    with its message before any GPU work."""
    r = run(binary(25), "--iter", "0")
    assert r.returncode != 0
    assert "--iter must be larger or equal to 1" in (r.stdout + r.stderr)

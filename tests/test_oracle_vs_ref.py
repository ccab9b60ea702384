"""Live comparison of the oracle with the compiled reference (oracle/_ref,
built from the reference's own sources by oracle/Makefile).  Skipped where the
reference build is absent (it never travels to GPU hosts as source)."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_equals_reference_iterate_3d(orc, npts, seed):
    R = orc.radius(npts)
    rng = np.random.default_rng(seed)
    n = int(rng.integers(6, 20))
    halo = (R + int(rng.integers(0, 3)), R + int(rng.integers(0, 3)), R + int(rng.integers(0, 5)))
    g = orc.Grid(n, n, int(rng.integers(5, 40)), *halo)
    u = orc.fill_uniform(g, seed, 0, halo=0.75)
    s = orc.fill_uniform(g, seed, 1, halo=-0.5)
    v = orc.fill_uniform(g, seed, 2, halo=0.25)
    v2 = v.copy()
    orc.ref_sweep_f32(npts, g, v, s, u, 7)
    orc.ref_iterate(npts, g, v2, s, u, 7, kind="3d")
    assert np.array_equal(v.view(np.uint32), v2.view(np.uint32))


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_oracle_equals_reference_blocked_noncubic(orc, npts):
    """Non-cubic grids through iterate_2d_jk_blocked (iterate_3d bounds j by
    xs, iterators.cpp:146, so it is only a valid reference when xs == ys)."""
    R = orc.radius(npts)
    g = orc.Grid(7, 23, 41, R, R, R)
    u, s, v = (orc.fill_uniform(g, 5, f) for f in range(3))
    v2 = v.copy()
    orc.ref_sweep_f32(npts, g, v, s, u, 4)
    orc.ref_iterate(npts, g, v2, s, u, 4, kind="2d", zb=5)
    assert np.array_equal(v.view(np.uint32), v2.view(np.uint32))


def test_reference_iterate_3d_is_wrong_for_non_cubic(orc):
    """Documents the reference defect the oracle does not reproduce."""
    g = orc.Grid(6, 9, 8, 1, 1, 1)
    u, s, v = (orc.fill_uniform(g, 5, f) for f in range(3))
    v2 = v.copy()
    orc.ref_sweep_f32(7, g, v, s, u, 1)
    orc.ref_iterate(7, g, v2, s, u, 1, kind="3d")
    assert not np.array_equal(v, v2)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_reference_binary_kat(npts):
    """The compiled hstencil3d passes its own check (main.cpp:302, 312)."""
    r = subprocess.run([O.ref_binary(npts), "--xs", "20", "--ys", "20", "--zs", "20", "--iter",
                        "4"], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0
    assert re.findall(r"check: (\d+) errors found", r.stdout) == ["0", "0"]

"""CPU tests of the library's host-only helpers against the oracle: the Ricker
source coefficients (configs 4 and 5 inject them as c2[n]) and the
constant-preserving centre coefficient c0."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("f,dt,t0", [(15.0, 1e-3, 1.0 / 15.0), (8.0, 2e-3, 0.1), (25.0, 5e-4, 0.0)])
def test_ricker_matches_oracle(sth, orc, f, dt, t0):
    steps = 1000
    got = sth.ricker(steps, dt, f, t0, dt * dt)
    want = np.array([dt * dt * orc.ricker(f, n * dt - t0) for n in range(steps)], np.float32)
    assert got.dtype == np.float32 and got.tobytes() == want.tobytes()
    # the wavelet itself: 1 at tau = 0, zero crossings at tau = +-1/(pi f sqrt 2)
    assert orc.ricker(f, 0.0) == 1.0
    assert abs(orc.ricker(f, 1.0 / (math.pi * f * math.sqrt(2.0)))) < 1e-15
    # its peak sits at n = t0/dt
    if t0 > 0:
        assert abs(int(np.argmax(got)) - round(t0 / dt)) <= 1


def test_ricker_defaults(sth):
    a = sth.ricker(100)
    b = sth.ricker(100, 1e-3, 15.0, 1.0 / 15.0, 1e-6)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("npts", [7, 13, 25])
@pytest.mark.parametrize("c1", [0.25, 0.16, 0.1])
def test_leapfrog_c0_keeps_constants(sth, orc, npts, c1):
    c0 = sth.leapfrog_c0(npts, c1)
    assert c0 == np.float32(2.0 - orc.lap_gain(npts) * float(np.float32(c1)))
    # the gain is 6 for 7 and 13 points, 12 for 25 (to FP32 rounding of the weights)
    gain = {7: 6.0, 13: 6.0, 25: 12.0}[npts]
    assert abs(orc.lap_gain(npts) - gain) < 1e-6


def test_c0_rejects_bad_stencil(sth):
    with pytest.raises(sth.SthError):
        sth.leapfrog_c0(9, 0.25)

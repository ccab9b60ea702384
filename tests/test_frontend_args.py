"""CPU tests of the hstencil3d_hip front end's argument checks: the reference
driver validates every flag before it allocates anything and exits with 2
and a range message (aux.cpp:15-81, main.cpp:149-150); so does the front end,
before its first HIP call."""
import os
import subprocess

import pytest


@pytest.fixture(scope="module")
def frontend(sth):
    if not os.path.exists(sth.FRONTEND_PATH):
        pytest.skip("hstencil3d_hip is not built")
    return sth.FRONTEND_PATH


@pytest.mark.parametrize("args,message", [
    (["--iter", "0"], "--iter must be larger or equal to 1"),
    (["--xs", "2000"], "--xs must be smaller or equal to 1264"),
    (["--steps-per-pass", "6"], "--steps-per-pass must be smaller or equal to 5"),
    (["--sweeps-per-launch", "0"], "--sweeps-per-launch must be larger or equal to 1"),
    (["--sweeps-per-launch", "5000"], "--sweeps-per-launch must be smaller or equal to 4096"),
])
def test_out_of_range_flag_exits_2(frontend, args, message):
    r = subprocess.run([frontend, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert message in r.stdout + r.stderr

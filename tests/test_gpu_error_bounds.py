"""The stated error bounds of the storage forms, asserted (config 4's medium:
25 points, 256^3, c dt/dx = 0.4, dt = 1 ms, a 15 Hz Ricker source at the
centre, 1000 steps), each form against the FP64 GPU run of the same setup
with the same weights (the FP32 constants promoted).  Measured on MI355X:
FP32 max-rel 2.1e-5; FP16 6.8e-3 (L2-rel 4.4e-3); BF16 6.8e-2 (3.8e-2).  The
bounds below leave room for run-to-run identity only; any regression in
the arithmetic, the scale choice or the rounding moves these figures by
far more."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUNDS = {"fp32": (5e-5, 3e-5), "fp16": (1e-2, 6e-3), "bf16": (1e-1, 6e-2)}


def run(gpu, dtype, n=256, steps=1000):
    c1 = 0.16
    c0 = gpu.leapfrog_c0(25, c1)
    c2 = gpu.ricker(steps, 1e-3, 15.0)
    with gpu.Stencil(n, n, n, npts=25, dtype=dtype) as ctx:
        ctx.init_point("u", 0, 0, 0, 0.0)
        ctx.init_point("v", 0, 0, 0, 0.0)
        ctx.init_point("s", n // 2, n // 2, n // 2, 1.0)
        ctx.leapfrog(steps, c0, c1, c2)
        return ctx.download("u", 4, dtype=np.float64)


@pytest.fixture(scope="module")
def reference(gpu):
    return run(gpu, "fp64")


@pytest.mark.parametrize("dtype", ["fp32", "fp16", "bf16"])
def test_stated_error_bound(gpu, reference, dtype):
    x = run(gpu, dtype)
    r = reference
    err = np.abs(x - r)
    maxrel = err.max() / np.abs(r).max()
    l2rel = np.sqrt((err ** 2).sum() / (r ** 2).sum())
    bmax, bl2 = BOUNDS[dtype]
    print(f"{dtype}: max-rel {maxrel:.3e}, L2-rel {l2rel:.3e}")
    assert np.abs(r).max() > 0
    assert maxrel <= bmax and l2rel <= bl2, (maxrel, l2rel)

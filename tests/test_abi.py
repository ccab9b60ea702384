"""CPU tests of the C ABI boundary: the library loads, exports every symbol
include/stencil_hip.h declares, its host-only layout logic is right, errors
are statuses (never exits), and the front end keeps the reference's CLI
validation (exit code 2, main.cpp:122-174, aux.cpp:15-81)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stencil_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sth_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(sth):
    names = declared_symbols()
    assert len(names) >= 25
    L = sth.lib()
    for n in names:
        assert hasattr(L, n), n
    # the Python binding declares a signature for every one of them
    assert sorted(sth.SIGNATURES) == names


def test_header_constants_match_the_binding(sth):
    """The batch method codes and status codes the Python binding uses are
    the header's."""
    text = open(HEADER).read()
    defs = {k: int(v, 0) for k, v in re.findall(r"#define (STH_[A-Z0-9_]+)\s+(0x[0-9a-fA-F]+|-?\d+)",
                                                 text)}
    for name in ("PERSISTENT", "SEQUENTIAL", "TWO_STEP", "THREE_STEP", "FUSED", "GRAPH"):
        assert getattr(sth, "BATCH_" + name) == defs["STH_BATCH_" + name], name
    assert len({defs["STH_BATCH_" + n] for n in ("PERSISTENT", "SEQUENTIAL", "TWO_STEP",
                                                  "THREE_STEP", "FUSED")}) == 5


def test_nm_exports_c_symbols(sth):
    out = subprocess.run(["nm", "-D", "--defined-only", sth.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sth_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_symbols()) <= exported


def test_version_and_strerror(sth):
    L = sth.lib()
    assert L.sth_version() == 1
    assert L.sth_strerror(0) == b"ok"
    assert L.sth_strerror(1) == b"invalid argument"
    assert L.sth_strerror(99) == b"unknown status"


@pytest.mark.parametrize("npts,dtype", [(7, "fp32"), (13, "fp32"), (25, "fp32"), (25, "fp16"),
                                        (13, "bf16")])
@pytest.mark.parametrize("shape", [(1040, 1040, 1040), (5, 7, 129), (128, 128, 128), (1, 1, 1)])
def test_layout(sth, npts, dtype, shape):
    L = sth.layout(*shape, npts, dtype)
    R = sth.radius(npts)
    eb = 4 if dtype == "fp32" else 2
    align = 256 // eb
    xs, ys, zs = shape
    assert L.radius == R and L.elem_bytes == eb
    assert L.pitch % align == 0 and L.pitch >= zs + 2 * R
    assert L.pitch < zs + 2 * R + align
    assert L.plane == L.pitch * (ys + 2 * R)
    assert (L.origin * eb) % 256 == 0          # interior k = 0 of row 0 on 256 B
    assert (L.pitch * eb) % 256 == 0           # ... and of every row
    assert L.origin >= R + R * L.pitch + R * L.plane
    # the last cell (xs+R-1, ys+R-1, zs+R-1) is inside the allocation
    last = L.origin + (zs + R - 1) + (ys + R - 1) * L.pitch + (xs + R - 1) * L.plane
    assert last < L.elems and L.bytes >= L.elems * eb and L.bytes % 256 == 0
    # the first cell (-R,-R,-R) too
    assert L.origin - R - R * L.pitch - R * L.plane >= 0


def test_layout_1040_figures(sth):
    L = sth.layout(1040, 1040, 1040, 25, "fp32")
    assert (L.pitch, L.plane) == (1088, 1088 * 1048)
    assert L.bytes == ((1088 * 1048 * 1048 + 64) * 4 + 255) // 256 * 256


def test_layout_rejects_bad_geometry(sth):
    import ctypes as C
    L = sth.lib()
    out = sth.LayoutInfo()
    for g in [sth.Geom(0, 4, 4, 7, 0), sth.Geom(4, 4, 4, 9, 0), sth.Geom(4, 4, 4, 7, 5),
              sth.Geom(70000, 4, 4, 7, 0), sth.Geom(4, 23200, 23200, 7, 0)]:
        assert L.sth_layout(C.byref(g), C.byref(out)) == sth.STH_ERR_ARG
    assert L.sth_layout(None, C.byref(out)) == sth.STH_ERR_ARG


def test_layout_plane_limit(sth):
    """A device plane must stay below 2 GiB (32-bit lane offsets)."""
    L = sth.layout(4, 23000, 23100, 7, "fp32")  # 23002 rows x 23104 (pitch) x 4 B < 2^31
    assert L.plane * 4 < 2 ** 31
    with pytest.raises(sth.SthError):
        sth.layout(4, 23200, 23200, 7, "fp32")


def test_null_handles_are_errors(sth):
    L = sth.lib()
    assert L.sth_iterate(None, 1) == sth.STH_ERR_ARG
    assert L.sth_destroy(None) == sth.STH_ERR_ARG
    assert L.sth_sync(None) == sth.STH_ERR_ARG
    assert L.sth_timer_read(None, None, None) == sth.STH_ERR_STATE


def test_every_entry_point_rejects_null_arguments(sth):
    """Every entry point called with null handles and pointers (and zero
    sizes) returns an error status, a null handle or zero, and touches no
    memory: no call may crash a caller that passes a bad handle."""
    import ctypes as C
    L = sth.lib()
    for name, (res, args) in sth.SIGNATURES.items():
        if not args or name == "sth_strerror":
            continue
        vals = [0.0 if a in (C.c_float, C.c_double) else
                0 if a in (C.c_int, C.c_uint32, C.c_uint64) else None for a in args]
        r = getattr(L, name)(*vals)
        if name == "sth_ricker":
            assert r == 0  # zero steps: nothing to write
        elif res is C.c_int:
            assert r != 0, name
        else:
            assert not r, name


def test_reference_driver_without_gpu_reports_the_library_error(sth):
    """The reference's main.cpp linked with the GPU iterators (built where the
    reference's sources are present) loads libstencil_hip and, with no GPU,
    stops at its first ITERATE with the library's message and exit code 3."""
    if sth.device_count() > 0:
        pytest.skip("a GPU is visible")
    exe = os.path.join(os.path.dirname(sth.LIB_PATH), "_refbind", "hstencil3d_gpu_7")
    if not os.path.exists(exe):
        pytest.skip("built only where the reference's sources are present")
    r = subprocess.run([exe, "--xs", "16", "--ys", "16", "--zs", "32", "--iter", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert "libstencil_hip: sth_create:" in r.stderr
    assert "Phase #3: initial warmup" in r.stdout  # init_arrays ran on the host arrays


def test_create_without_gpu_is_a_status(sth):
    if sth.device_count() > 0:
        pytest.skip("a GPU is visible")
    import ctypes as C
    h = C.c_void_p()
    g = sth.Geom(8, 8, 8, 7, 0)
    assert sth.lib().sth_create(C.byref(g), C.byref(h)) == sth.STH_ERR_NODEV
    with pytest.raises(sth.SthError):
        sth.Stencil(8, 8, 8)


# ------------------------------------------------------------ front end CLI
def run_fe(sth, *args):
    return subprocess.run([sth.FRONTEND_PATH, *args], capture_output=True, text=True,
                          timeout=60)


@pytest.mark.parametrize("args,msg", [
    (["--iter", "0"], "--iter must be larger or equal to 1"),
    (["--iter", "100001"], "--iter must be smaller or equal to 100000"),
    (["--xh", "3"], "--xh must be larger or equal to 4"),
    (["--xs", "2000"], "--xs must be smaller or equal to 1264"),
    (["--bogus"], "unsupported argument --bogus"),
    (["--ys", "10", "--yb", "3"], "--ys must by an exact multiple of --yb"),
    (["--iter"], "--iter requires a numeric argument"),
    (["--sbm", "20"], "--sbm must be smaller or equal to"),
    (["--mode", "ref", "--dtype", "fp16"], "--mode ref runs FP32 or FP64 storage only"),
    (["--host-fields", "--mode", "leapfrog", "--dtype", "bf16"],
     "--host-fields needs FP32 or FP64 storage"),
    (["--stencil", "9"], "--stencil must be 7, 13 or 25"),
    (["--stencil", "25", "--zh", "2"], "halos must be at least the stencil radius 4"),
])
def test_frontend_rejects_bad_arguments(sth, args, msg):
    r = run_fe(sth, *args)
    assert r.returncode == 2
    assert msg in r.stderr


def test_frontend_echoes_settings_like_the_reference(sth):
    r = run_fe(sth, "--xs", "32", "--iter", "3", "--train")
    assert "--xs set to 32" in r.stdout and "--iter set to 3" in r.stdout
    assert "--train set to ENABLED" in r.stdout
    if sth.device_count() == 0:
        assert r.returncode == 3 and "no usable GPU" in r.stderr

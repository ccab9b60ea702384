"""Field files (sth_save / sth_load) against the oracle: the raw files are
read with numpy as manifest.json describes them (dense box with an R-wide
halo, C order, k fastest, little-endian; 16-bit files hold q with value =
q * 2^exp) and compared with the oracle's fields after the same steps --
FP32 and FP64 bit for bit, 16-bit storage against the oracle's emulation
(stored numbers, exponents and maxima).  Files written here by numpy load
into a context and step on to the oracle's result, so save and load are each
checked against the layout, not only against each other."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP_DTYPE = {"fp32": np.float32, "fp64": np.float64, "fp16": np.uint16, "bf16": np.uint16}


def read_fields(directory):
    with open(os.path.join(directory, "manifest.json")) as f:
        man = json.load(f)
    h = man["halo"]
    shape = (man["xs"] + 2 * h, man["ys"] + 2 * h, man["zs"] + 2 * h)
    out = {}
    for name, meta in man["fields"].items():
        a = np.fromfile(os.path.join(directory, meta["file"]), dtype=NP_DTYPE[man["dtype"]])
        assert a.size == np.prod(shape), name
        out[name] = (a.reshape(shape), meta["exp"], meta["maxabs_q"])
    return man, out


def write_fields(directory, xs, ys, zs, npts, R, dtype, fields):
    """fields: name -> (array, exp); the manifest keys sth_load reads."""
    eb = np.dtype(NP_DTYPE[dtype]).itemsize
    man = {"format": "stencil3d_amd fields", "version": 1, "xs": xs, "ys": ys, "zs": zs,
           "npts": npts, "halo": R, "dtype": dtype, "elem_bytes": eb, "fields": {}}
    for name, (a, e) in fields.items():
        np.ascontiguousarray(a).tofile(os.path.join(directory, name + ".raw"))
        man["fields"][name] = {"file": name + ".raw", "exp": int(e), "maxabs_q": 0.0}
    with open(os.path.join(directory, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1)


def seeded(orc, g, seed):
    u = orc.fill_uniform(g, seed, 0, halo=0.375)
    v = orc.fill_uniform(g, seed, 1, halo=-0.25)
    s = orc.fill_uniform(g, seed, 2, halo=0.5)
    return u, v, s


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
@pytest.mark.parametrize("npts", [7, 25])
def test_saved_files_hold_the_oracle_fields(gpu, orc, tmp_path, dtype, npts):
    shape = (19, 26, 70)
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 21)
    c1 = float(np.float32(0.16))  # the FP32 value the library receives
    c0 = float(np.float32(2.0 - orc.lap_gain(npts) * c1))
    c2 = np.linspace(1e-3, -1e-3, 5).astype(np.float32)
    ft = NP_DTYPE[dtype]
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        for name, x in (("u", u), ("v", v), ("s", s)):
            ctx.upload(name, x.astype(ft), R)
        ctx.leapfrog(5, c0, c1, c2)
        ctx.save(str(tmp_path))
    man, got = read_fields(str(tmp_path))
    assert (man["xs"], man["ys"], man["zs"], man["npts"], man["halo"], man["dtype"]) == \
        (*shape, npts, R, dtype)
    a, b, ss = u.astype(ft), v.astype(ft), s.astype(ft)
    if dtype == "fp32":
        newest = orc.leapfrog_f32(npts, g, a, b, ss, c0, c1, c2, 5)
    else:
        newest = orc.leapfrog_f64(npts, g, a, b, ss, c0, c1, c2.astype(np.float64), 5)
    want_u, want_v = (a, b) if newest == 0 else (b, a)
    for name, want in (("u", want_u), ("v", want_v), ("s", ss)):
        arr, e, _ = got[name]
        assert e == 0
        assert arr.tobytes() == want.tobytes(), name


@pytest.mark.parametrize("fmt,dtype", [(1, "fp16"), (2, "bf16")])
def test_saved_16bit_files_hold_the_emulated_fields(gpu, orc, tmp_path, fmt, dtype):
    shape, npts = (15, 18, 67), 13
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    v = orc.fill_uniform(g, 6, 0)
    s = orc.fill_uniform(g, 6, 1) * np.float32(1e-3)
    u = orc.fill_uniform(g, 6, 2)
    qs, exps, maxs = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, mx = orc.quantize(fmt, g, x, e)
        qs.append(q)
        exps.append(e)
        maxs.append(mx)
    c1, c0 = 0.16, 2.0 - orc.lap_gain(npts) * 0.16
    c2 = np.linspace(0.5, -0.5, 7).astype(np.float32)
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        for name, q, e in (("u", qs[0], exps[0]), ("v", qs[1], exps[1]), ("s", qs[2], exps[2])):
            ctx.upload_raw16(name, q, e, R)
        ctx.leapfrog(7, c0, c1, c2)
        ctx.save(str(tmp_path))
    _, got = read_fields(str(tmp_path))
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], exps, maxs, np.float32(c0),
                              np.float32(c1), c2, 7)
    iu, iv = (0, 1) if newest == 0 else (1, 0)
    want = {"u": ((a, b)[iu], exps[iu], maxs[iu]), "v": ((a, b)[iv], exps[iv], maxs[iv]),
            "s": (qs[2], exps[2], maxs[2])}
    I = g.interior
    for name in ("u", "v", "s"):
        arr, e, mx = got[name]
        wq, we, wm = want[name]
        assert e == we, name
        assert np.array_equal(arr[I], wq[I]), name
        assert np.float32(mx) == np.float32(wm), name


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_files_written_by_numpy_load_and_step_on(gpu, orc, tmp_path, dtype):
    shape, npts = (21, 17, 66), 25
    R = orc.radius(npts)
    g = orc.Grid(*shape, R, R, R)
    u, v, s = seeded(orc, g, 22)
    ft = NP_DTYPE[dtype]
    write_fields(str(tmp_path), *shape, npts, R, dtype,
                 {"u": (u.astype(ft), 0), "v": (v.astype(ft), 0), "s": (s.astype(ft), 0)})
    c1 = float(np.float32(0.16))
    c0 = float(np.float32(2.0 - orc.lap_gain(npts) * c1))
    with gpu.Stencil(*shape, npts=npts, dtype=dtype) as ctx:
        ctx.load(str(tmp_path))
        ctx.leapfrog(3, c0, c1, 2e-3)
        got = ctx.download("u", R, dtype=ft)
    a, b, ss = u.astype(ft), v.astype(ft), s.astype(ft)
    c2 = np.full(3, 2e-3, np.float32)
    if dtype == "fp32":
        newest = orc.leapfrog_f32(npts, g, a, b, ss, c0, c1, c2, 3)
    else:
        newest = orc.leapfrog_f64(npts, g, a, b, ss, c0, c1, c2.astype(np.float64), 3)
    want = a if newest == 0 else b
    assert got.tobytes() == want.tobytes()

"""CPU tests of the oracle (oracle/stencil_oracle.c) against what pins it:
golden vectors from the compiled reference, the reference's own known-answer
test, published SplitMix64 outputs, IEEE/numpy/torch conversions, and the
closed forms and invariants of the leapfrog scheme (which has no reference
implementation)."""
import hashlib
import json
import math
import os
from fractions import Fraction as F

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


# ------------------------------------------------------------------ golden
CASES = manifest()["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_ref_sweep_matches_reference_golden(orc, case):
    """REF sweep bit-identical to the compiled reference's ITERATE output."""
    g = orc.Grid(*case["shape"], *case["halo"])
    f64 = case["kind"].endswith("f64")
    hu, hs, hv = case["halos"]
    u = orc.fill_uniform(g, case["seed"], 0, halo=hu)
    s = orc.fill_uniform(g, case["seed"], 1, halo=hs)
    v = orc.fill_uniform(g, case["seed"], 2, halo=hv)
    if f64:
        u, s, v = u.astype(np.float64), s.astype(np.float64), v.astype(np.float64)
    assert {"u": sha(u), "s": sha(s), "v": sha(v)} == case["inputs_sha256"]
    want = np.fromfile(os.path.join(GOLDEN, case["file"]), dtype=case["dtype"])
    if f64:
        orc.ref_sweep_f64(case["npts"], g, v, s, u, case["sweeps"], coeff_mode=1)
    else:
        orc.ref_sweep_f32(case["npts"], g, v, s, u, case["sweeps"])
    got = np.ascontiguousarray(v[g.interior]).ravel()
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_golden_kat_lines_recorded():
    kat = manifest()["reference_kat"]
    assert [k["npts"] for k in kat] == [7, 13, 25]
    for k in kat:
        assert k["check_lines"] == ["check: 0 errors found"] * 2


# ------------------------------------------------------------------ generator
def test_splitmix64_published_vectors(orc):
    """First outputs of SplitMix64 (Vigna's splitmix64.c) for seed 1234567."""
    G, M = 0x9E3779B97F4A7C15, (1 << 64) - 1
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [orc.splitmix64_mix((1234567 + k * G) & M) for k in range(1, 6)] == want


def test_uniform_counter_definition(orc):
    G, M = 0x9E3779B97F4A7C15, (1 << 64) - 1
    seed, field, (i, j, k) = 42, 3, (-2, 5, 1000)
    key = orc.splitmix64_mix((seed + (field + 1) * G) & M)
    c = ((i + (1 << 20)) << 42) | ((j + (1 << 20)) << 21) | (k + (1 << 20))
    x = orc.splitmix64_mix((key + (c + 1) * G) & M)
    assert orc.uniform01(seed, field, i, j, k) == (x >> 40) / 2.0 ** 24


def test_fill_uniform_layout_independent(orc):
    """Values depend on the coordinate only, not on halo widths; [-1,1) values
    are exact multiples of 2^-23."""
    a = orc.fill_uniform(orc.Grid(5, 6, 7, 1, 1, 1), 9, 2)
    b = orc.fill_uniform(orc.Grid(5, 6, 7, 4, 2, 16), 9, 2)
    ia, ib = orc.Grid(5, 6, 7, 1, 1, 1).interior, orc.Grid(5, 6, 7, 4, 2, 16).interior
    assert np.array_equal(a[ia], b[ib])
    x = a[ia].astype(np.float64)
    assert np.all((x >= -1) & (x < 1))
    assert np.all((x + 1) * 2 ** 23 == np.round((x + 1) * 2 ** 23))


# ------------------------------------------------------------------ REF KAT
@pytest.mark.parametrize("npts", [7, 13, 25])
def test_reference_kat_on_oracle(orc, npts):
    """init_arrays + 1 + 20 sweeps -> check() finds 0 errors (main.cpp:289-312),
    and every interior v equals the scalar FP32 recurrence."""
    R = orc.radius(npts)
    g = orc.Grid(9, 10, 11, 4, 4, R)
    v, s, u = orc.init_arrays(g)
    orc.ref_sweep_f32(npts, g, v, s, u, 1)
    assert orc.check(g, v, 1.0) == 0
    orc.ref_sweep_f32(npts, g, v, s, u, 20)
    assert orc.check(g, v, 21.0) == 0
    assert orc.check(g, v, 22.0) == 9 * 10 * 11

    def rnd(x):
        return F(float(np.float32(float(x))))
    coef0, A, w, _ = orc.coeffs(npts)
    t = orc.terms(npts)
    p = rnd(F(float(coef0)) + rnd(F(float(w[t[0][0], t[0][1] - 1])) * 2))
    for ax, r in t[1:]:
        p = rnd(F(float(w[ax, r - 1])) * 2 + p)
    vn = F(0)
    for _ in range(21):
        vn = rnd(p + rnd(vn - F(float(A))))
    assert np.all(v[g.interior] == np.float32(float(vn)))
    if npts == 7:
        assert float(vn) == 21.0


def test_coefficients_are_the_reference_literals(orc):
    """stencil_13.h:13-22 / stencil_25.h:19-34: cy = cx*1.1f, cz = cx*0.9f in FP32."""
    for npts, cx in ((13, [0.3, 0.2]), (25, [0.4, 0.3, 0.2, 0.1])):
        _, _, w, R = orc.coeffs(npts)
        for r, c in enumerate(cx):
            f = np.float32(c)
            assert w[0, r] == f
            assert w[1, r] == np.float32(f * np.float32(1.1))
            assert w[2, r] == np.float32(f * np.float32(0.9))
    assert orc.coeffs(7)[:2] == (1.5, 3.5) and orc.coeffs(25)[:2] == (1.25, 6.25)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_ref_closed_form_f64(orc, npts):
    """u and s are read-only in REF, so v_N = v_0 + N*(s*P(u) - A*u) exactly;
    evaluated with exact rationals on a tiny grid by explicit neighbour
    enumeration."""
    R = orc.radius(npts)
    g = orc.Grid(3, 4, 5, R, R, R)
    u = orc.fill_uniform(g, 4, 0, halo=0.5).astype(np.float64)
    s = orc.fill_uniform(g, 4, 1, halo=0.5).astype(np.float64)
    v = orc.fill_uniform(g, 4, 2).astype(np.float64)
    v0 = v.copy()
    N = 6
    orc.ref_sweep_f64(npts, g, v, s, u, N)
    coef0, A, w, _ = orc.coeffs(npts)
    for i in range(3):
        for j in range(4):
            for k in range(5):
                c = (i + R, j + R, k + R)
                P = F(float(coef0)) * F(u[c])
                for ax in range(3):
                    for r in range(1, R + 1):
                        d = [0, 0, 0]
                        d[ax] = r
                        lo = (c[0] - d[0], c[1] - d[1], c[2] - d[2])
                        hi = (c[0] + d[0], c[1] + d[1], c[2] + d[2])
                        P += F(float(w[ax, r - 1])) * (F(u[lo]) + F(u[hi]))
                exact = F(v0[c]) + N * (F(s[c]) * P - F(float(A)) * F(u[c]))
                assert abs(float(exact) - v[c]) <= 1e-13 * max(1.0, abs(float(exact)))


# ------------------------------------------------------------------ LEAPFROG
def test_leapfrog_brute_force_fractions(orc):
    """README.md:10-24 scheme, v <- c0 u - v + c1 sum_n wh (u[-r]+u[+r]) + c2 s,
    evaluated with exact rationals by explicit neighbour enumeration on a tiny
    grid (13 points, R = 2), against the FP64 oracle."""
    npts, R = 13, 2
    g = orc.Grid(4, 3, 5, R, R, R)
    a = orc.fill_uniform(g, 1, 0).astype(np.float64)
    b = orc.fill_uniform(g, 1, 1).astype(np.float64)
    s = orc.fill_uniform(g, 1, 2).astype(np.float64)
    c0, c1 = 2.0 - 6 * 0.2, 0.2
    c2 = [0.125, -0.25, 0.5]
    _, _, w, _ = orc.coeffs(npts)
    U = {k: F(x) for k, x in np.ndenumerate(a)}
    V = {k: F(x) for k, x in np.ndenumerate(b)}
    for n in range(3):
        new = dict(V)
        for i in range(4):
            for j in range(3):
                for k in range(5):
                    c = (i + R, j + R, k + R)
                    L = F(0)
                    for ax in range(3):
                        for r in range(1, R + 1):
                            d = [0, 0, 0]
                            d[ax] = r
                            lo = (c[0] - d[0], c[1] - d[1], c[2] - d[2])
                            hi = (c[0] + d[0], c[1] + d[1], c[2] + d[2])
                            L += 2 * F(float(w[ax, r - 1])) * (U[lo] + U[hi])
                    new[c] = F(c0) * U[c] - V[c] + F(c1) * L + F(c2[n]) * F(s[c])
        U, V = new, U
    newest = orc.leapfrog_f64(npts, g, a, b, s, c0, c1, np.array(c2), 3)
    got = a if newest == 0 else b
    for idx in np.ndindex(4, 3, 5):
        c = (idx[0] + R, idx[1] + R, idx[2] + R)
        assert abs(got[c] - float(U[c])) <= 1e-12 * max(1.0, abs(float(U[c])))


def test_leapfrog_sine_mode_is_exact(orc):
    """7 points, zero halo: a discrete sine mode is an eigenvector of the
    neighbour sum (eigenvalue 2 sum cos theta), so the scheme reduces to the
    scalar recurrence A_{n+1} = mu A_n - A_{n-1}, mu = c0 + c1*lambda, with
    solution A_n = cos(n w), cos w = mu/2 (u_0 = phi, u_-1 = cos(w) phi)."""
    xs, ys, zs = 11, 9, 13
    g = orc.Grid(xs, ys, zs, 1, 1, 1)
    ma, mb, mc = 2, 3, 1
    th = [math.pi * ma / (xs + 1), math.pi * mb / (ys + 1), math.pi * mc / (zs + 1)]
    i, j, k = np.meshgrid(np.arange(xs), np.arange(ys), np.arange(zs), indexing="ij")
    phi = (np.sin(th[0] * (i + 1)) * np.sin(th[1] * (j + 1)) * np.sin(th[2] * (k + 1)))
    c1 = 0.25
    c0 = 2 - 6 * c1
    lam = 2 * sum(math.cos(t) for t in th)
    mu = c0 + c1 * lam
    w = math.acos(mu / 2)
    a, b, s = g.zeros(np.float64), g.zeros(np.float64), g.zeros(np.float64)
    a[g.interior] = phi
    b[g.interior] = math.cos(w) * phi
    N = 40
    newest = orc.leapfrog_f64(7, g, a, b, s, c0, c1, None, N)
    got = (a if newest == 0 else b)[g.interior]
    assert np.abs(got - math.cos(N * w) * phi).max() < 1e-12


@pytest.mark.parametrize("npts", [13, 25])
def test_leapfrog_plane_wave_dispersion(orc, npts):
    """Away from the boundary (distance > R*n) a plane wave cos(theta . x)
    evolves by the exact discrete dispersion relation
    cos w = (c0 + c1 * sum_axis,r 2 wh cos(r theta_axis)) / 2."""
    R = orc.radius(npts)
    n = 3
    N = 2 * R * n + 6
    g = orc.Grid(N, N, N, R, R, R)
    th = [0.3, 0.7, 1.1]
    _, _, w, _ = orc.coeffs(npts)
    c1 = 0.05
    c0 = 2 - orc.lap_gain(npts) * c1
    lam = sum(2 * 2 * float(w[ax, r - 1]) * math.cos(r * th[ax])
              for ax in range(3) for r in range(1, R + 1))
    mu = c0 + c1 * lam
    om = math.acos(mu / 2)
    i, j, k = np.meshgrid(*(np.arange(-R, N + R),) * 3, indexing="ij")
    phase = th[0] * i + th[1] * j + th[2] * k
    a = np.cos(phase)
    b = np.cos(phase) * math.cos(om) - np.sin(phase) * math.sin(om)  # u_{-1} = cos(x + w)
    s = np.zeros_like(a)
    newest = orc.leapfrog_f64(npts, g, a, b, s, c0, c1, None, n)
    got = a if newest == 0 else b
    want = np.cos(phase - n * om)
    inner = (slice(R + R * n, R + N - R * n),) * 3
    assert np.abs(got[inner] - want[inner]).max() < 1e-11


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_leapfrog_energy_conserved(orc, npts):
    """With s = 0 and zero halos, Q_n = <u_n,u_n> + <u_{n-1},u_{n-1}>
    - <u_n, M u_{n-1}> (M = c0 + c1 L, symmetric) is conserved exactly."""
    R = orc.radius(npts)
    g = orc.Grid(10, 12, 14, R, R, R)
    a = orc.fill_uniform(g, 2, 0).astype(np.float64)
    b = orc.fill_uniform(g, 2, 1).astype(np.float64)
    z = g.zeros(np.float64)
    c1 = 0.1
    c0 = 2 - orc.lap_gain(npts) * c1

    def M(x):  # one step with v = 0, s = 0 gives c0 x + c1 L(x)
        y, t = x.copy(), g.zeros(np.float64)
        orc.leapfrog_f64(npts, g, y, t, z, c0, c1, None, 1)
        return t

    def Q(un, um):
        I = g.interior
        return float((un[I] ** 2).sum() + (um[I] ** 2).sum() - (un[I] * M(um)[I]).sum())

    q0 = Q(a, b)
    newest = orc.leapfrog_f64(npts, g, a, b, z, c0, c1, None, 25)
    un, um = (a, b) if newest == 0 else (b, a)
    assert abs(Q(un, um) - q0) <= 1e-11 * abs(q0)


def test_leapfrog_f32_tracks_f64(orc):
    g = orc.Grid(16, 16, 16, 4, 4, 4)
    a = orc.fill_gaussian(g, 1.0, 3.0, 1e-3, 1, 0)
    b = a.copy()
    s = g.zeros()
    ad, bd = a.astype(np.float64), b.astype(np.float64)
    c1 = 0.16
    c0 = 2 - orc.lap_gain(25) * c1
    n32 = orc.leapfrog_f32(25, g, a, b, s, np.float32(c0), np.float32(c1), None, 30)
    n64 = orc.leapfrog_f64(25, g, ad, bd, s.astype(np.float64), float(np.float32(c0)),
                           float(np.float32(c1)), None, 30)
    x32 = a if n32 == 0 else b
    x64 = ad if n64 == 0 else bd
    mr, lr = orc.error_vs_f64(g, x32, x64)
    assert mr < 1e-5 and lr < 1e-5


# ------------------------------------------------------------------ 16-bit
def test_fp16_conversion_matches_numpy(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(100),
        (rng.standard_normal(2000) * 1e-6).astype(np.float32),  # fp16 subnormals
        (rng.standard_normal(1000) * 1e-8).astype(np.float32),  # underflow region
        np.array([65504, 65519.99, 65520, 1e6, -1e6, 0.0, -0.0, 2 ** -24, 2 ** -25,
                  3 * 2 ** -26, 2 ** -25 * 1.0000001, np.inf, -np.inf], np.float32),
    ])
    # exact ties between fp16 neighbours (round to even)
    h = np.arange(0x0001, 0x7bff, 97, dtype=np.uint16)
    lo = h.view(np.float16).astype(np.float64)
    hi = (h + 1).view(np.float16).astype(np.float64)
    xs = np.concatenate([xs, ((lo + hi) / 2).astype(np.float32)])
    for x in xs:
        with np.errstate(over="ignore"):  # values past 65520 round to inf on purpose
            want = int(np.float32(x).astype(np.float16).view(np.uint16))
        assert orc.to16(1, float(x)) == want, x
    # every finite fp16 value round-trips
    for q in range(0, 0x10000, 7):
        f = orc.from16(1, q)
        if np.isfinite(np.uint16(q).view(np.float16)):
            assert f == float(np.uint16(q).view(np.float16))
            assert orc.to16(1, f) == q or (q == 0x8000 and f == 0.0)


def test_bf16_conversion_matches_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    xs = rng.standard_normal(5000).astype(np.float32) * np.float32(1e3)
    base = rng.integers(0, 2 ** 31, 500, dtype=np.uint32) & np.uint32(0x7fff0000)
    ties = (base | np.uint32(0x8000)).view(np.float32)  # exact halfway cases
    xs = np.concatenate([xs, ties[np.isfinite(ties)]])
    want = torch.from_numpy(xs).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([orc.to16(2, float(x)) for x in xs], dtype=np.uint16)
    assert np.array_equal(got, want)
    for q in (0x3f80, 0xbf80, 0x0001, 0x7f7f):
        assert orc.from16(2, q) == float(np.uint32(q << 16).view(np.float32))


def test_scale_exponent_rule(orc):
    for B in [1e-30, 3e-7, 0.5, 1.0, 1.5, 32767.0, 32768.0, 65504.0, 1e20]:
        e = orc.scale_exp(B)
        assert B * 2.0 ** -e < 2 ** 15 <= B * 2.0 ** -(e - 1)
    assert orc.scale_exp(0.0) == 0


@pytest.mark.parametrize("fmt", [1, 2])
def test_q16_leapfrog_emulation_invariants(orc, fmt):
    """16-bit storage never overflows (max stored value < 2^15 * 2^e), the
    recorded max equals the max of the decoded field, and the result tracks
    the FP32 scheme within the format's precision."""
    npts, R = 25, 4
    g = orc.Grid(12, 13, 14, R, R, R)
    u = orc.fill_gaussian(g, 1.0, 2.5, 1e-3, 1, 0)
    v = orc.fill_gaussian(g, 1.0, 2.5, 1e-3, 1, 1)
    s = g.zeros()
    qs, ex, mx = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, m = orc.quantize(fmt, g, x, e)
        qs.append(q)
        ex.append(e)
        mx.append(m)
    c1 = 0.16
    c0 = np.float32(2 - orc.lap_gain(npts) * c1)
    a, b = qs[0].copy(), qs[1].copy()
    newest = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], ex, mx, c0, np.float32(c1), None, 20)
    q = a if newest == 0 else b
    e = ex[newest]
    x = orc.dequantize(fmt, g, q, e)
    qv = orc.dequantize(fmt, g, q, 0)
    assert np.all(np.isfinite(x))
    assert float(np.abs(qv).max()) == mx[newest]
    assert float(np.abs(qv).max()) < 2 ** 15
    # FP32 scheme from the same (quantized) start
    f0 = orc.dequantize(fmt, g, qs[0], orc.scale_exp(float(np.abs(u).max())))
    f1 = orc.dequantize(fmt, g, qs[1], orc.scale_exp(float(np.abs(v).max())))
    n32 = orc.leapfrog_f32(npts, g, f0, f1, s, c0, np.float32(c1), None, 20)
    ref = f0 if n32 == 0 else f1
    mr, _ = orc.error_vs_f64(g, x, ref.astype(np.float64))
    assert mr < (4e-3 if fmt == 1 else 4e-2)


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_q16_pass2_against_one_step_rule(orc, npts):
    """Passes of two steps (the second step's exponent from the bound of the
    first's output) against the one-step rule from the same start: the first
    step of a pass is the same step, bit for bit; the second uses an exponent
    at least the one-step rule's, and since scaling by a power of two
    commutes with rounding its decoded values equal the one-step rule's
    wherever both stored numbers are normal binary16, within half a
    subnormal step elsewhere.  A step left over runs with the one-step rule."""
    R = orc.radius(npts)
    g = orc.Grid(9, 10, 11, R, R, R)
    fmt = 1
    u = orc.fill_gaussian(g, 1.0, 2.0, 1e-3, 1, 0)
    v = orc.fill_gaussian(g, 1.0, 2.0, 1e-3, 1, 1)
    s = orc.fill_uniform(g, 3, 2, -1e-3, 1e-3)  # halos 0 (the 16-bit form's boundary)
    qs, ex, mx = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, m = orc.quantize(fmt, g, x, e)
        qs.append(q)
        ex.append(e)
        mx.append(m)
    c1 = np.float32(0.16)
    c0 = np.float32(2 - orc.lap_gain(npts) * 0.16)
    c2 = np.array([0.3, -0.2, 0.5], np.float32)
    # one pass (two steps)
    a1, b1 = qs[0].copy(), qs[1].copy()
    e1, m1 = list(ex), list(mx)
    n1 = orc.leapfrog_q16(npts, fmt, g, a1, b1, qs[2], e1, m1, c0, c1, c2, 2)
    a2, b2 = qs[0].copy(), qs[1].copy()
    e2, m2 = list(ex), list(mx)
    n2 = orc.leapfrog_q16(npts, fmt, g, a2, b2, qs[2], e2, m2, c0, c1, c2, 2, pass2=True)
    assert n1 == n2 == 0
    I = g.interior
    assert np.array_equal(b1[I], b2[I]) and e1[1] == e2[1] and m1[1] == m2[1]  # u_{n+1}
    assert e2[0] >= e1[0]
    x1 = orc.dequantize(fmt, g, a1, e1[0])[I]
    x2 = orc.dequantize(fmt, g, a2, e2[0])[I]
    normal = np.abs(orc.dequantize(fmt, g, a2, 0)[I]) >= 2.0 ** -14
    assert np.array_equal(x1[normal], x2[normal])
    assert np.all(np.abs(x1 - x2) <= 2.0 ** (-25 + e2[0]))
    # a pass and a step left over: the last step is the one-step rule's
    a3, b3 = qs[0].copy(), qs[1].copy()
    e3, m3 = list(ex), list(mx)
    n3 = orc.leapfrog_q16(npts, fmt, g, a3, b3, qs[2], e3, m3, c0, c1, c2, 3, pass2=True)
    a4, b4 = a2.copy(), b2.copy()
    e4, m4 = list(e2), list(m2)
    n4 = orc.leapfrog_q16(npts, fmt, g, a4, b4, qs[2], e4, m4, c0, c1, c2[2:], 1)
    assert n3 == 1 and n4 == 1
    assert np.array_equal(b3[I], b4[I]) and e3 == e4 and m3 == m4


@pytest.mark.parametrize("npts", [7, 13, 25])
def test_q16_passes_bound_recursion(orc, npts):
    """Passes of L steps (orc_leapfrog_q16_passes).  A pass of one step is the
    one-step rule, bit for bit.  The exponents of a pass's last two outputs
    are scale_exp of B_{L-1} and B_L from the recursion
    B_M = (|c0| + |c1| gain) B_{M-1} + B_{M-2} + |c2[M-1]| M_s, B_0 = M_u,
    B_{-1} = M_v, recomputed here from the inputs' maxima (a wrong index or
    coefficient in the recursion moves an exponent); each output's maximum
    stays at or below its bound (no overflow); 2L + r steps are two passes of
    L and one of r."""
    import math
    R = orc.radius(npts)
    g = orc.Grid(8, 9, 10, R, R, R)
    fmt = 1
    u = orc.fill_gaussian(g, 1.0, 2.0, 1e-3, 2, 0)
    v = orc.fill_gaussian(g, 0.7, 2.0, 1e-3, 2, 1)
    s = orc.fill_uniform(g, 4, 2, -1e-3, 1e-3)
    qs, ex, mx = [], [], []
    for x in (u, v, s):
        e = orc.scale_exp(float(np.abs(x).max()))
        q, m = orc.quantize(fmt, g, x, e)
        qs.append(q)
        ex.append(e)
        mx.append(m)
    c1 = np.float32(0.16)
    c0 = np.float32(2 - orc.lap_gain(npts) * 0.16)
    c2 = np.array([0.3, -0.2, 0.5, -0.4, 0.25, 0.1, -0.3, 0.2, 0.6, -0.1, 0.35, 0.15, -0.25],
                  np.float32)
    I = g.interior

    def run(steps, **kw):
        a, b = qs[0].copy(), qs[1].copy()
        e, m = list(ex), list(mx)
        n = orc.leapfrog_q16(npts, fmt, g, a, b, qs[2], e, m, c0, c1, c2, steps, **kw)
        return n, a, b, e, m

    n1, a1, b1, e1, m1 = run(5)
    n2, a2, b2, e2, m2 = run(5, plen=1)
    assert n1 == n2 and np.array_equal(a1[I], a2[I]) and np.array_equal(b1[I], b2[I])
    assert e1 == e2 and m1 == m2
    gmul = abs(float(c0)) + abs(float(c1)) * orc.lap_gain(npts)
    Ms = math.ldexp(float(mx[2]), ex[2])
    for L in (2, 3, 4, 5):
        bm2, bm1 = math.ldexp(float(mx[1]), ex[1]), math.ldexp(float(mx[0]), ex[0])
        bounds = []
        for M in range(1, L + 1):
            B = gmul * bm1 + bm2 + abs(float(c2[M - 1])) * Ms
            bounds.append(B)
            bm2, bm1 = bm1, B
        n, a, b, e, m = run(L, plen=L)
        assert n == L % 2
        last, prev = (a, b) if n == 0 else (b, a)
        il, ip = n, 1 - n  # buffer indices of u_{n+L} and u_{n+L-1}
        assert e[il] == orc.scale_exp(bounds[-1]) and e[ip] == orc.scale_exp(bounds[-2])
        assert math.ldexp(float(m[il]), e[il]) <= bounds[-1]
        assert math.ldexp(float(m[ip]), e[ip]) <= bounds[-2]
        assert np.all(np.isfinite(orc.dequantize(fmt, g, last, e[il])[I]))
        # 2L + r steps: two passes of L, then the remainder as one pass
        for r in (0, 1, L - 1):
            nA, aA, bA, eA, mA = run(2 * L + r, plen=L)
            a, b = qs[0].copy(), qs[1].copy()
            e, m = list(ex), list(mx)
            cur = 0
            for n0, ln in ((0, L), (L, L), (2 * L, r)):
                if ln == 0:
                    continue
                x, y = (a, b) if cur == 0 else (b, a)
                ee = e if cur == 0 else [e[1], e[0], e[2]]
                mm = m if cur == 0 else [m[1], m[0], m[2]]
                k = orc.leapfrog_q16(npts, fmt, g, x, y, qs[2], ee, mm, c0, c1, c2[n0:], ln,
                                     plen=ln)
                if cur == 1:
                    e[:] = [ee[1], ee[0], ee[2]]
                    m[:] = [mm[1], mm[0], mm[2]]
                else:
                    e[:], m[:] = ee, mm
                cur = cur ^ k
            assert nA == cur and eA == e and mA == m
            assert np.array_equal(aA[I], a[I]) and np.array_equal(bA[I], b[I])


# ------------------------------------------------------------------ misc
def test_gaussian_init(orc):
    g = orc.Grid(7, 8, 9, 2, 2, 2)
    a = orc.fill_gaussian(g, 2.0, 1.5, 0.0, 0, 0)
    i, j, k = np.meshgrid(np.arange(7), np.arange(8), np.arange(9), indexing="ij")
    d2 = (i - 3.0) ** 2 + (j - 3.5) ** 2 + (k - 4.0) ** 2
    want = (2.0 * np.exp(-d2 / (2 * 1.5 ** 2))).astype(np.float32)
    got = a[g.interior]
    assert np.abs(got.view(np.int32) - want.view(np.int32)).max() <= 1
    assert np.all(np.delete(a.ravel(), np.ravel_multi_index(
        np.nonzero(np.pad(np.ones((7, 8, 9)), 2)), a.shape)) == 0)


def test_ricker(orc):
    f = 15.0
    for tau in (0.0, 0.01, -0.02, 0.1):
        a = (math.pi * f * tau) ** 2
        assert orc.ricker(f, tau) == pytest.approx((1 - 2 * a) * math.exp(-a), rel=1e-15)
    assert orc.ricker(f, 0.0) == 1.0


def test_check_counts_planted_errors(orc):
    g = orc.Grid(6, 6, 6, 1, 1, 1)
    v = g.zeros()
    v[g.interior] = 5.0
    v[1, 1, 1] = 5.2
    v[3, 4, 5] = 4.85
    v[0, 0, 0] = 100.0  # halo: not counted
    assert orc.check(g, v, 5.0) == 2

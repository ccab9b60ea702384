"""Writes the golden vectors of tests/golden/ from the compiled reference.

Run in the build container only (needs oracle/_ref, which oracle/Makefile
compiles from the reference's own sources):

    make -C oracle ref && python tests/golden/gen_golden.py

Inputs are seeded fields from the oracle's counter-based generator (their
SHA-256 is recorded so a generator change is caught); outputs are the
reference's ITERATE results (iterate_3d, or iterate_2d_jk_blocked for
non-cubic grids since iterate_3d bounds j by xs, iterators.cpp:146), stored
as raw little-endian interiors with a JSON manifest.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import pyoracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fields(g, seed, halos, dtype=np.float32):
    hu, hs, hv = halos
    u = O.fill_uniform(g, seed, 0, halo=hu)
    s = O.fill_uniform(g, seed, 1, halo=hs)
    v = O.fill_uniform(g, seed, 2, halo=hv)
    if dtype == np.float64:
        u, s, v = u.astype(np.float64), s.astype(np.float64), v.astype(np.float64)
    return v, s, u


CASES = []
for n in (7, 13, 25):
    R = O.radius(n)
    CASES += [
        dict(name=f"ref{n}_cubic_s1", npts=n, shape=(16, 16, 16), halo=(R, R, R), seed=0,
             sweeps=1, kind="3d", halos=(0.0, 0.0, 0.0)),
        dict(name=f"ref{n}_cubic_s20", npts=n, shape=(16, 16, 16), halo=(R, R, R), seed=0,
             sweeps=20, kind="3d", halos=(0.0, 0.0, 0.0)),
        dict(name=f"ref{n}_noncubic_s3", npts=n, shape=(12, 14, 20), halo=(R, R, R), seed=1,
             sweeps=3, kind="2d", halos=(0.0, 0.0, 0.0)),
        dict(name=f"ref{n}_widehalo_s2", npts=n, shape=(10, 10, 24), halo=(4, 4, 16), seed=2,
             sweeps=2, kind="3d", halos=(1.0, 1.0, 0.0)),
        dict(name=f"ref{n}_f64_s4", npts=n, shape=(12, 12, 12), halo=(R, R, R), seed=3,
             sweeps=4, kind="3d_f64", halos=(0.0, 0.0, 0.0)),
    ]


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    manifest = {"generator": "tests/golden/gen_golden.py",
                "reference_build": "oracle/Makefile: clang++ -g -O3 -fopenmp -march=x86-64-v4 "
                                   "-DSCHED_STATIC=1 -DSTENCIL_SIZE=N [-DPAR2D=1|-DUSE_DOUBLE=1]",
                "layout": "interior only, C order (i, j, k), little-endian",
                "cases": []}
    for c in CASES:
        g = O.Grid(*c["shape"], *c["halo"])
        f64 = c["kind"].endswith("f64")
        dt = np.float64 if f64 else np.float32
        v, s, u = fields(g, c["seed"], c["halos"], dt)
        ins = {"u": sha(u), "s": sha(s), "v": sha(v)}
        O.ref_iterate(c["npts"], g, v, s, u, c["sweeps"], kind=c["kind"], zb=8)
        out = np.ascontiguousarray(v[g.interior])
        fn = c["name"] + (".f64" if f64 else ".f32")
        out.tofile(os.path.join(HERE, fn))
        manifest["cases"].append(dict(c, file=fn, dtype=str(out.dtype), inputs_sha256=ins,
                                      output_sha256=sha(out)))
    # the reference binary's own known-answer check (u=s=1, v=0; main.cpp:289-312)
    kat = []
    for n in (7, 13, 25):
        cmd = [O.ref_binary(n), "--xs", "24", "--ys", "24", "--zs", "24", "--xh", "4", "--yh",
               "4", "--zh", "4", "--iter", "20"]
        res = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ,
                             OMP_NUM_THREADS="4"), check=True)
        checks = [ln.strip() for ln in res.stdout.splitlines() if ln.startswith("check:")]
        kat.append({"npts": n, "cmd": " ".join(["hstencil3d_%d" % n] + cmd[1:]),
                    "check_lines": checks})
    manifest["reference_kat"] = kat
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(f"wrote {len(CASES)} cases")


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""bench.py -- throughput of the stencil3d sweep on one MI355X.

Workload (the headline metric): 25-point leapfrog sweep over a 1040^3 interior
(halo 4 = 0), u = v = Gaussian pulse (sigma 8 cells) at the centre plus
U[-1e-3, 1e-3) noise from seed 1, s = 0, c1 = 0.16, c0 = 2 - 12 c1.  A step is
one sweep over the whole grid (one kernel launch); fields are resident in HBM
before the timed region starts.  FP32 storage gives `value`; the same run with
FP16 and BF16 storage (FP32 arithmetic, per-field power-of-two scale) is
reported beside it, with its error against an FP64 run of the Ricker-source
medium (256^3, c dt/dx = 0.4, 1000 steps).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints ONE JSON line.  The CPU baseline is the reference's own hstencil3d
(25-point, OpenMP static) timed on this host's cores on a bounded number of
sweeps of the same grid, when oracle/_ref holds it (built from the reference's
sources in the build container), else the repo's CPU restatement.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpts/s + HBM GB/s (% of peak), 25-pt 1040^3: FP16-store vs FP32 @ stated err"
PEAK_GBS = 8000.0
NPTS = 25
FLOPS_PT = {7: 12, 13: 29, 25: 41}  # config.h:75,81,87


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def leap_coeffs(npts, c1):
    import stencil3d_amd as S  # noqa: F401
    gain = {7: 6.0, 13: 6.0, 25: 12.0}[npts]
    return 2.0 - gain * c1, c1


def run_throughput(S, n, dtype, steps, warmup, tuning=None, copy_bytes=0):
    """Returns wall seconds of `steps` timed steps, the HIP-event kernel time
    and launch count, the tuning used, the u norms and (copy_bytes > 0) the
    mean time of a plain copy kernel of copy_bytes."""
    c0, c1 = leap_coeffs(NPTS, 0.16)
    with S.Stencil(n, n, n, npts=NPTS, dtype=dtype) as ctx:
        ctx.init_gaussian("u", 1.0, 8.0, 1e-3, 1, 0)
        ctx.init_gaussian("v", 1.0, 8.0, 1e-3, 1, 1)
        ctx.init_point("s", 0, 0, 0, 0.0)
        if tuning:
            ctx.set_tuning(**tuning)
        ctx.leapfrog(warmup, c0, c1)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.timer_start()
        ctx.leapfrog(steps, c0, c1)
        kms, launches = ctx.timer_read()  # synchronises the stream
        t1 = time.perf_counter()
        mx, l2 = ctx.norms("u")
        used = ctx.get_tuning()
        copy_ms = ctx.stream_copy(copy_bytes, 5) if copy_bytes else None
    return t1 - t0, kms, launches, used, (mx, l2), copy_ms


def ricker_c2(steps, dt=1e-3, f=15.0):
    t0 = 1.0 / f
    out = []
    for n in range(steps):
        a = (math.pi * f * (n * dt - t0)) ** 2
        out.append(dt * dt * (1.0 - 2.0 * a) * math.exp(-a))
    return out


def run_error(S, dtype, n=256, steps=1000):
    """Config-4 medium: 25-point, c dt/dx = 0.4, Ricker source (15 Hz, dt 1 ms)
    at the centre; error of the `dtype` run against the FP64 GPU run."""
    import numpy as np
    c0, c1 = leap_coeffs(NPTS, 0.16)
    c2 = ricker_c2(steps)
    res = {}
    for dt in (dtype, "fp64"):
        with S.Stencil(n, n, n, npts=NPTS, dtype=dt) as ctx:
            ctx.init_point("u", 0, 0, 0, 0.0)
            ctx.init_point("v", 0, 0, 0, 0.0)
            ctx.init_point("s", n // 2, n // 2, n // 2, 1.0)
            ctx.leapfrog(steps, c0, c1, c2)
            res[dt] = ctx.download("u", 4, dtype=np.float64)
    x, r = res[dtype], res["fp64"]
    err = np.abs(x - r)
    return float(err.max() / np.abs(r).max()), float(np.sqrt((err ** 2).sum() / (r ** 2).sum()))


def run_physics_7pt(S, n=128, steps=20):
    """Small physics check: 7-point LEAPFROG on a 128^3 interior, halo 1 = 0,
    u, v, s ~ U[-1, 1) from seed 0 (bit-identical in FP32 and FP64: 24-bit
    values), c1 = 0.25, c0 = 2 - 6 c1, c2 = 1e-6; FP32 storage against the
    FP64 run of the same fields.  33.6 MB per step stays in the 256 MB
    last-level cache, so this is a numerics line, not an HBM measurement."""
    import numpy as np
    c1 = 0.25
    c0 = 2.0 - 6.0 * c1
    res = {}
    for dt in ("fp32", "fp64"):
        with S.Stencil(n, n, n, npts=7, dtype=dt) as ctx:
            for fid, f in enumerate(("v", "s", "u")):
                ctx.init_uniform(f, 0, fid, -1.0, 1.0, 0.0)
            ctx.leapfrog(steps, c0, c1, 1e-6)
            res[dt] = ctx.download("u", 1, dtype=np.float64)
    x, r = res["fp32"], res["fp64"]
    err = np.abs(x - r)
    return {"maxrel": float(err.max() / np.abs(r).max()),
            "l2rel": float(np.sqrt((err ** 2).sum() / (r ** 2).sum())),
            "setup": f"7-pt LEAPFROG {n}^3, halo 1 = 0, U[-1,1) seed 0, c1=0.25, "
                     f"c2=1e-6, {steps} steps, FP32 vs FP64"}


def run_batch(S, steps=100, nshots=64, seed=2):
    """Batching form: nshots independent 13-point FP32 shots, sides drawn from
    seed 2 uniformly among {96,128,192,256,320,384} per axis, a Ricker source at
    a random interior point of each, c dt/dx = 0.4; `steps` steps in one call,
    timed for the persistent launch (in a hipGraph and eager), for one plain
    launch per shot per step (eager and in a hipGraph), and for the persistent
    two- and three-steps-per-pass launches.  Tile and planes per item of the
    one-step persistent launch come from the autotuner."""
    import numpy as np
    rng = np.random.default_rng(seed)
    dims = [tuple(int(x) for x in rng.choice([96, 128, 192, 256, 320, 384], 3))
            for _ in range(nshots)]
    srcs = [tuple(int(rng.integers(0, d)) for d in dd) for dd in dims]
    pts = float(sum(a * b * c for a, b, c in dims))
    c1 = 0.16
    c0 = 2.0 - 6.0 * c1
    c2 = ricker_c2(steps)
    out = {"shots": nshots, "npts": 13, "steps": steps, "points_per_step": int(pts)}
    with S.Batch(dims, npts=13) as bt:
        tuned, tms = bt.autotune(reps=2)
        out["tuning"] = bt.get_tuning()
        for name, method, graph in (("persistent_graph", S.BATCH_PERSISTENT, True),
                                    ("persistent_eager", S.BATCH_PERSISTENT, False),
                                    ("sequential_eager", S.BATCH_SEQUENTIAL, False),
                                    ("sequential_graph", S.BATCH_SEQUENTIAL, True),
                                    ("two_step_graph", S.BATCH_TWO_STEP, True),
                                    ("two_step_eager", S.BATCH_TWO_STEP, False),
                                    ("three_step_graph", S.BATCH_THREE_STEP, True),
                                    ("three_step_eager", S.BATCH_THREE_STEP, False)):
            for q, (d, sp) in enumerate(zip(dims, srcs)):
                sh = bt.shot(q)
                sh.init_point("u", 0, 0, 0, 0.0)
                sh.init_point("v", 0, 0, 0, 0.0)
                sh.init_point("s", *sp, 1.0)
            bt.leapfrog(6, c0, c1, None, method=method, graph=graph)  # warm-up
            bt.sync()
            bt.timer_start()
            t0 = time.perf_counter()
            bt.leapfrog(steps, c0, c1, c2, method=method, graph=graph)
            bt.sync()
            wall = time.perf_counter() - t0
            ms, launches = bt.timer_read()
            out[name] = {"gpu_ms_per_step": round(ms / steps, 4),
                         "wall_ms_per_step": round(wall * 1e3 / steps, 4),
                         "gpts": round(pts * steps / (ms * 1e-3) / 1e9, 2),
                         "hbm_gbs_alg": round(16.0 * pts * steps / (ms * 1e-3) / 1e9, 1),
                         "launches": launches}
            log(f"[bench] batch {name}: {ms / steps:.3f} ms/step, "
                f"{out[name]['gpts']} Gpts/s")
    p, q = out["persistent_graph"]["gpu_ms_per_step"], out["sequential_eager"]["gpu_ms_per_step"]
    out["persistent_speedup_vs_sequential"] = round(q / p, 3)
    out["two_step_speedup_vs_persistent"] = round(p / out["two_step_graph"]["gpu_ms_per_step"], 3)
    out["three_step_speedup_vs_persistent"] = round(
        p / out["three_step_graph"]["gpu_ms_per_step"], 3)
    return out


def run_other_stencils(S, n, steps, warmup):
    """The plain FP32 sweep for every stencil size at n^3: REF (the reference's
    update, what hstencil3d runs) and LEAPFROG, Gpts/s from HIP events."""
    out = {}
    for npts in (7, 13, 25):
        c0, c1 = leap_coeffs(npts, 0.16)
        row = {}
        with S.Stencil(n, n, n, npts=npts, dtype="fp32") as ctx:
            ctx.init_gaussian("u", 1.0, 8.0, 1e-3, 1, 0)
            ctx.init_gaussian("v", 1.0, 8.0, 1e-3, 1, 1)
            ctx.init_point("s", 0, 0, 0, 0.0)
            for mode in ("ref", "leapfrog", "ref_10_per_launch"):
                if mode == "ref_10_per_launch":
                    # opt-in: ten REF sweeps per launch (bit-identical to ten
                    # launches; u and s are read once per ten sweeps)
                    ctx.set_tuning(ref_sweeps=10)
                run = (lambda k: ctx.leapfrog(k, c0, c1)) if mode == "leapfrog" else \
                    (lambda k: ctx.iterate(k))
                run(warmup)
                ctx.sync()
                ctx.timer_start()
                run(steps)
                ms, nl = ctx.timer_read()
                per = ms / steps  # per step (a pass counts as its steps)
                g = float(n) ** 3 / (per * 1e-3) / 1e9
                row[mode] = {"kernel_ms_per_step": round(per, 4), "gpts": round(g, 2),
                             "hbm_gbs_alg": round(16.0 * g, 1),
                             "gflops_nominal": round(FLOPS_PT[npts] * g, 1),
                             "steps_per_launch": round(steps / nl, 3)}
                if mode == "leapfrog":
                    row[mode]["tuning"] = ctx.get_tuning()
        out[str(npts)] = row
        log(f"[bench] fp32 {npts}-pt: REF {row['ref']['gpts']} Gpts/s, "
            f"LEAPFROG {row['leapfrog']['gpts']} Gpts/s, "
            f"REF 10 sweeps per launch {row['ref_10_per_launch']['gpts']} Gpts/s")
    return out


def cpu_baseline(n, sweeps, npts=NPTS):
    """hstencil3d (the reference, iterate_3d, OpenMP static) on this host, or
    the repo's CPU restatement when the reference build is absent."""
    cores = os.cpu_count() or 1
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="spread",
               OMP_PLACES="cores")
    binary = os.path.join(ROOT, "oracle", "_ref", f"hstencil3d_{npts}")
    if os.path.exists(binary):
        cmd = [binary, "--xs", str(n), "--ys", str(n), "--zs", str(n), "--xh", "4", "--yh", "4",
               "--zh", "4", "--iter", str(sweeps)]
        out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
        m = re.search(r"finish compute in ([0-9.]+) secs", out)
        if not m:
            raise RuntimeError("hstencil3d printed no timing")
        secs = float(m.group(1))
        kind = "reference"
        what = (f"hstencil3d_{npts} (reference iterate_3d + stencil_{npts}_ref_1, REF update, "
                f"OpenMP static) {n}^3, {sweeps} timed sweeps after 1 warm-up")
    else:
        import numpy as np  # noqa: F401
        from oracle import pyoracle as O
        g = O.Grid(n, n, n, 4, 4, 4)
        a = O.fill_gaussian(g, 1.0, 8.0, 1e-3, 1, 0)
        b = O.fill_gaussian(g, 1.0, 8.0, 1e-3, 1, 1)
        s = g.zeros()
        c0, c1 = leap_coeffs(npts, 0.16)
        O.leapfrog_f32(npts, g, a, b, s, c0, c1, None, 1, cores)
        t0 = time.perf_counter()
        O.leapfrog_f32(npts, g, a, b, s, c0, c1, None, sweeps, cores)
        secs = time.perf_counter() - t0
        kind = "port"
        what = f"oracle leapfrog_f32 (OpenMP static) {n}^3, {sweeps} timed sweeps"
    gpts = n ** 3 * sweeps / secs / 1e9
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                       "")
    except OSError:
        pass
    return {"value": round(gpts, 4), "unit": "Gpts/s", "cores": cores, "kind": kind,
            "sample": what, "secs": round(secs, 3), "cpu": cpu}


def measured_traffic(n, dtype_code, tune):
    """HBM bytes per launch of this sweep schedule from the committed rocprofv3
    summary (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE runs scaled by a
    copy kernel of known size); None when that schedule was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("grid") != [n, n, n]:
        return None, None
    if tune.get("steps_per_pass", 0) >= 2:
        prefix = (f"void sth::sweep2_kernel<{NPTS}, {tune['rows_per_wave']}, "
                  f"{tune['waves']}, {tune['steps_per_pass']}")
    else:
        # the last template argument (fold of the last k-tile) follows the grid
        prefix = (f"void sth::sweep_kernel<{NPTS}, 1, {dtype_code}, {tune['rows_per_wave']}, "
                  f"{tune['waves']}, {tune['vec']}")
    hits = [v for kname, v in t["kernels"].items()
            if kname.startswith(prefix) and kname[len(prefix)] in ",>"]
    return (hits[0]["traffic_bytes"], t["source"]) if len(hits) == 1 else (None, None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=1040)
    ap.add_argument("--cpu-sweeps", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-16bit", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-other", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the sweep is not sharded across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    import stencil3d_amd as S
    if S.device_count() < 1:
        sys.exit("no GPU visible")
    n, K, W = args.size, args.steps, args.warmup
    pts = float(n) ** 3

    log(f"[bench] fp32 {NPTS}-pt {n}^3: {W} warm-up + {K} timed steps")
    copy_bytes = (n * n * n * 4) // 16 * 16
    wall, kms, launches, tune, nrm, copy_ms = run_throughput(S, n, "fp32", K, W,
                                                             copy_bytes=copy_bytes)
    copy_gbs = 2.0 * copy_bytes / (copy_ms * 1e-3) / 1e9
    gpts = pts * K / wall / 1e9
    # A launch advances `units` steps (2 to 5 for a fused pass):
    # algorithmic bytes per launch = 16 B x points x units.
    units = K / launches
    per_launch_s = kms / launches * 1e-3
    bytes_launch = 16.0 * pts * units
    achieved = bytes_launch / per_launch_s / 1e9
    log(f"[bench] fp32: {gpts:.2f} Gpts/s, kernel {kms / launches:.4f} ms per launch "
        f"({units:g} steps), {achieved:.0f} GB/s")
    traffic, traffic_src = measured_traffic(n, 0, tune)
    one_step = None
    if tune.get("steps_per_pass", 0) >= 2:
        # the one-step sweep on the same fields: the baseline of the FP16 ratio
        w1, k1, l1, t1, _, _ = run_throughput(S, n, "fp32", K, W, tuning={"steps_per_pass": 1})
        g1 = pts * K / w1 / 1e9
        one_step = {"gpts": round(g1, 3), "ms_per_step": round(w1 / K * 1e3, 4),
                    "kernel_ms": round(k1 / l1, 4),
                    "hbm_gbs_alg": round(16.0 * pts / (k1 / l1 * 1e-3) / 1e9, 1), "tuning": t1}
        log(f"[bench] fp32 one step per pass: {g1:.2f} Gpts/s")

    extra = {}
    if not args.no_16bit:
        for dt in ("fp16", "bf16"):
            log(f"[bench] {dt}: {W} warm-up + {K} timed steps")
            w2, k2, l2, t2, _, _ = run_throughput(S, n, dt, K, W)
            g2 = pts * K / w2 / 1e9
            a2 = 8.0 * pts / (k2 / l2 * 1e-3) / 1e9
            extra[dt] = {"gpts": round(g2, 3), "ms_per_step": round(w2 / K * 1e3, 4),
                         "kernel_ms": round(k2 / l2, 4), "hbm_gbs_alg": round(a2, 1),
                         "frac_of_peak": round(a2 / PEAK_GBS, 4),
                         "speedup_vs_fp32": round(g2 / gpts, 3), "tuning": t2}
            if one_step:
                extra[dt]["speedup_vs_fp32_one_step"] = round(g2 / one_step["gpts"], 3)
            log(f"[bench] {dt}: {g2:.2f} Gpts/s ({g2 / gpts:.2f}x fp32)")
        if not args.no_error:
            log("[bench] error of fp32 vs FP64 (256^3, 1000 steps, Ricker source)")
            mr, lr = run_error(S, "fp32")
            extra["fp32_err_vs_fp64"] = {"maxrel": mr, "l2rel": lr}
            for dt in ("fp16", "bf16"):
                log(f"[bench] error of {dt} vs FP64 (256^3, 1000 steps, Ricker source)")
                mr, lr = run_error(S, dt)
                extra[dt]["err_vs_fp64"] = {"maxrel": mr, "l2rel": lr,
                                            "setup": "25-pt 256^3, c dt/dx=0.4, 1000 steps, "
                                                     "Ricker 15 Hz at the centre"}
                log(f"[bench] {dt}: max-rel {mr:.3e}, L2-rel {lr:.3e}")

    if not args.no_other:
        extra["fp32_by_stencil"] = run_other_stencils(S, n, 24, 6)
    if not args.no_error:
        extra["physics_7pt_128"] = run_physics_7pt(S)
        log(f"[bench] 7-pt 128^3 physics: {extra['physics_7pt_128']}")

    if not args.no_batch:
        log("[bench] batch: 64 shots x 13 points, 100 steps")
        extra["batch"] = run_batch(S)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_sweeps} sweeps")
        cpu = cpu_baseline(n, args.cpu_sweeps)
        log(f"[bench] cpu: {cpu['value']} Gpts/s ({cpu['kind']}, {cpu['cores']} cores)")
        cpu7 = cpu_baseline(n, args.cpu_sweeps, npts=7)
        log(f"[bench] cpu 7-pt: {cpu7['value']} Gpts/s")
        extra["cpu_baseline_7pt"] = cpu7

    line = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Gaussian pulse (sigma 8) + U[-1e-3,1e-3) noise, seed 1; s = 0",
        "config": {"workload": f"{NPTS}-point leapfrog sweep, {n}^3 interior, FP32 storage",
                   "npts": NPTS, "grid": [n, n, n], "halo": 4, "c1": 0.16,
                   "tuning": tune},
        "hbm_gbs_alg": round(16.0 * gpts, 1),
        "gflops_nominal": round(FLOPS_PT[NPTS] * gpts, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel_ms": round(kms / launches, 5),
                     "bytes_per_launch_alg": bytes_launch, "steps_per_launch": units,
                     "copy_kernel_gbs": round(copy_gbs, 1),
                     "copy_kernel_bytes": 2 * copy_bytes},
        "cpu_baseline": cpu,
        "norms_u": {"max": nrm[0], "l2": nrm[1]},
    }
    if one_step:
        extra["fp32_one_step_per_pass"] = one_step
    line.update(extra)
    # The stated targets: 25-point FP32 at >= 80 % of 8 TB/s; FP16 storage
    # >= 1.6x the FP32 Gpts/s at max-rel <= 1e-3 against FP64 (1000 steps).
    tg = {"fp32_frac_of_peak": round(achieved / PEAK_GBS, 4), "fp32_frac_target": 0.8,
          "fp32_met": achieved / PEAK_GBS >= 0.8}
    if "fp16" in extra:
        # against the FP32 headline (two steps per pass when that is the
        # default); the ratio against the one-step FP32 sweep, the same
        # schedule as the 16-bit sweep, is reported beside it
        sp = extra["fp16"]["speedup_vs_fp32"]
        tg.update({"fp16_speedup": sp, "fp16_speedup_target": 1.6, "fp16_speedup_met": sp >= 1.6})
        if "speedup_vs_fp32_one_step" in extra["fp16"]:
            tg["fp16_speedup_vs_fp32_one_step"] = extra["fp16"]["speedup_vs_fp32_one_step"]
        if "err_vs_fp64" in extra["fp16"]:
            mr = extra["fp16"]["err_vs_fp64"]["maxrel"]
            tg.update({"fp16_maxrel": mr, "fp16_maxrel_target": 1e-3,
                       "fp16_maxrel_met": mr <= 1e-3})
    line["targets"] = tg
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

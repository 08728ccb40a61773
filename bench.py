"""Benchmark of the FFT / propagation hot path on one MI355X.

Workload (the batched-propagation configuration): 16 complex FP32 fields of
2048 x 2048, amplitude a Gaussian of waist 0.3*N*dx times a seeded uniform
phase, pitch 4 um, lambda 633 nm, per-field z uniform in [10, 100] mm.

A step is one fft2d (forward) over the whole batch, inputs resident in HBM.
`value` is the step's algorithmic HBM throughput: 32 B per point (FP32
planar complex read + write in the row pass and in the column pass) x points
/ step time.  The same run times the rest of the configuration's pipeline
(band-limited angular spectrum forward + inverse, single-FFT Fresnel) and
reports it under "pipeline".  `roofline` covers the dominant kernel of the
step, timed with HIP events on the library's stream; `cpu_baseline` is the
FP64 JavaScript oracle (the reference's algorithm: radix-2 FFT) on one node
thread, on a bounded sample.

"gs_batch_uneven" runs the batched uneven GS configuration: 256 far-field
WGS designs of sides {256, 512, 1024, 2048} (spot arrays, filled polygons),
200 graph-captured iterations.

"storage_variants" times the same fft2d step with the batch stored in FP64,
FP16 and BF16 (the half-width ones compute in FP32).

"targets" times makeTarget + autoRoi on the device for 64 mixed 1024^2
targets (text and image rasters, disc, ring, spot grid, zone plate, checker,
lines) and planeToRGB (viridis, percentile 0.99) of a 2048^2 field;
"fabrication" times errorDiffusionQuantize (4 levels) on one 1024^2 image
with plain rows (wavefront) and with serpentine rows (one serial chain), and
on 64 images of 512^2 with serpentine rows (one chain per image).

The second half of the metric, WGS time to screen RMSE <= 1 % at 1024^2
(learned warm start vs plain random start, and a restart of WGS from its own
best iterate after 25 iterations), is reported under "wgs": the network
trained at the learned-warm-start configuration (20k pairs of 500-iteration
labels, output 1024^2; parameters in diffractiveopticalelement_amd/data)
starts 32 held-out 1024^2 far-field designs (spots, blobs, lines, glyphs),
which run 500 WGS iterations from each start; per family it gives the
designs reaching 1 %, the GPU time to it (inference included), the best rmse
reached, and plain's GPU time over the start's at the common threshold
max(1 %, best plain, best start).  "labels@1024" starts from each design's
own best plain phase: the bound on any network of that output resolution.

"c1_projector" times the small reference configuration (ProjectorSystem
n 256, 100 image-plane WGS iterations, seed 0, text 'DOE') through the
JavaScript drop-in; "errors" states the error of the timed FP32 results
against the FP64 path and naive-DFT crops (and the 4096^2 precision sweep's
figures); "checksum" is the batch's round trip after the timed loop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-wgs]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import shutil
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
# per-dispatch HBM byte counters of the same kernels (rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs of this script), scaled to bytes by the copy calibration
COUNTERS = ROOT / "profiles" / "fft_passes_16x2048sq_f32_counters.json"
B, N = 16, 2048
DX, LAM = 4e-6, 633e-9


C1_CFG = {"n": 256, "pitch": 8e-6, "lambda": 532e-9, "distance": 0.03, "mode": "image-plane",
          "algorithm": "wgs", "iterations": 100, "seed": 0, "target": {"kind": "text", "text": "DOE"}}


def host_info() -> dict:
    """CPU model, core count and node version of the host running the baseline."""
    model = platform.processor() or platform.machine()
    try:
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    node = None
    if shutil.which("node"):
        node = subprocess.run(["node", "--version"], capture_output=True, text=True).stdout.strip()
    return {"cpu_model": model, "nproc": os.cpu_count(), "node": node,
            "pinned": "taskset -c 0" if shutil.which("taskset") else None}


def cpu_baseline(n: int = 2048, reps: int = 4) -> dict | None:
    """The reference's node path (the FP64 JavaScript oracle: radix-2 FFT,
    per-call transfer function, shifts) on ONE host core (taskset -c 0), on a
    bounded sample of each workload: fft2d 2048^2 (the metric), band-limited
    ASM and Fresnel at 2048^2, one far-field WGS iteration at 1024^2, the
    small reference configuration (256^2, 100 WGS iterations, simulate +
    screenMetrics), target + ROI and planeToRGB at 1024^2."""
    if shutil.which("node") is None:
        return None
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_js as oj  # the oracle is the CPU baseline's code, nothing else

    def t(args, timeout=900):
        return oj.run("time", args, timeout=timeout, pin_cpu=0)[1]

    res = t({"what": "fft2d", "n": n, "reps": reps})
    ms = res["ms"]
    gbs = 32 * n * n / (ms * 1e-3) / 1e9
    ra = t({"what": "asm", "n": n, "reps": 1, "dx": DX, "lambda": LAM, "z": 0.055})
    rf = t({"what": "fresnel", "n": n, "reps": 1, "dx": DX, "lambda": LAM, "z": 0.055})
    rw = t({"what": "wgs_ff", "n": 1024, "reps": 1, "iterations": 1})
    _, rc1 = oj.run("projector", {"cfg": C1_CFG}, timeout=900, pin_cpu=0)
    rt = t({"what": "target", "kind": "text", "n": 1024, "reps": 2})
    ri = t({"what": "image", "n": 1024, "reps": 2, "opts": {"map": "viridis", "percentile": 0.99}})
    return {
        "value": round(gbs, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{reps} forward fft2d of {n}x{n} complex FP64 by the JS oracle (radix-2 DIT, "
                   f"strided columns) on one node {res.get('node')} thread pinned to core 0, 1 warm-up; "
                   f"{ms:.1f} ms per transform; GB/s counts the same 32 B/pt as the GPU value"),
        "host": host_info(),
        "per_unit_ms": {
            "fft2d_2048sq": round(ms, 1),
            "asm_bandlimited_2048sq": round(ra["ms"], 1),
            "fresnel_2048sq": round(rf["ms"], 1),
            "wgs_far_field_iteration_1024sq": round(rw["ms"], 1),
            "c1_design_100_iterations_256sq": round(rc1["design_ms"], 1),
            "c1_simulate_screen_metrics_256sq": round(rc1["simulate_metrics_ms"], 1),
            "make_target_text_1024sq": round(rt["ms"], 2),
            "plane_to_rgb_p99_1024sq": round(ri["ms"], 2),
        },
        "per_unit_sample": ("one unit each (1 warm-up + 1 timed; fft2d 4, targets/images 2), JS oracle on "
                            "one pinned node thread; the batched configurations extrapolate linearly: "
                            "16 x 2048^2 ASM fwd + inv + Fresnel = 16 x (2 asm + 1 fresnel); "
                            "C1 design = 100 WGS iterations at 256^2 (ProjectorSystem, seed 0)"),
        "c1_metrics": {k: rc1["metrics"][k] for k in ("rmse", "efficiency", "correlation", "ssim")},
    }


def traffic_of(tag: str) -> int | None:
    """HBM bytes per launch of kernel `tag` from the committed counter run, or None."""
    if not COUNTERS.exists():
        return None
    d = json.loads(COUNTERS.read_text())
    k = d["kernels"].get(tag)
    if k is None:
        return None
    return int(round((k["fetch_kib"] * d["scale"]["fetch"] + k["write_kib"] * d["scale"]["write"]) * 1024))


def storage_variants(ctx, field, steps: int) -> dict:
    """The same fft2d step with the batch stored in FP64, FP16 and BF16 (FP32
    arithmetic for the half-width ones): ms per step and algorithmic GB/s at
    the storage's own bytes (re + im read and written per pass)."""
    out = {}
    re, im = None, None
    for dt, size in (("f64", 8), ("f16", 2), ("bf16", 2)):
        g = ctx.field(N, N, B, dt)
        for b in range(B):  # convert field by field through the host
            src = ctx.wrap(field.ptrs()[0] + b * N * N * 4, field.ptrs()[1] + b * N * N * 4, N, N, 1, "f32")
            re, im = src.download()
            src.free()
            dst_re = g.ptrs()[0] + b * N * N * size
            dst_im = g.ptrs()[1] + b * N * N * size
            dst = ctx.wrap(dst_re, dst_im, N, N, 1, dt)
            dst.upload(re, im)
            dst.free()
        ctx.autotune_fft2d(N, N, B, dt, reps=3)  # the configurations that exist for this storage
        # forward and inverse alternate (as in the main step), so half-width
        # storage stays finite
        ctx.fft2d(g)
        ctx.fft2d(g, inverse=True)
        ctx.sync()
        reps = 2 * max(5, steps // 2)
        ctx.timer_start()
        for i in range(reps):
            ctx.fft2d(g, inverse=bool(i & 1))
        ms = ctx.timer_stop() / reps
        # per-kernel durations from a separate profiled pair of calls
        ctx.set_flags(asynchronous=True, profile=True)
        for i in range(4):
            ctx.fft2d(g, inverse=bool(i & 1))
        ctx.sync()
        prof = ctx.profile_report()
        ctx.set_flags(asynchronous=True, profile=False)
        # two passes, each reading and writing both planes
        out[dt] = {"ms_per_step": round(ms, 4), "gbs_alg": round(8 * size * B * N * N / (ms * 1e-3) / 1e9, 1),
                   "bytes_per_pt": 8 * size, "tuning": ctx.tune_get(N, N, B, dt),
                   "kernels_avg_us": {t: round(tot / c * 1e3, 2) for t, (c, tot, _) in prof.items()}}
        g.free()
    return out


def _rel(a, b) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def dft_crop(x: np.ndarray, k0: int, k1: int, size: int) -> np.ndarray:
    """Naive DFT (the reference's dft2dNaive convention, exp(-2 pi i n k / N),
    unscaled) of an ny x nx grid at the output rows k0..k0+size and columns
    k1..k1+size, in FP64: separable partial sums with exactly reduced
    twiddle angles (n k mod N)."""
    ny, nx = x.shape
    ky = np.arange(k0, k0 + size)
    kx = np.arange(k1, k1 + size)
    wy = np.exp(-2j * np.pi * (np.outer(ky, np.arange(ny)) % ny) / ny)
    wx = np.exp(-2j * np.pi * (np.outer(np.arange(nx), kx) % nx) / nx)
    return wy @ (x @ wx)


def errors_leg(ctx, field, zs, picks=(0, 15)) -> dict:
    """Error of the timed FP32 results against the FP64 path, on fields
    `picks` of the batch (their FP32 values promoted to FP64 on both sides):
    fft2d vs the device's FP64 check mode, numpy's FP64 FFT and a naive DFT
    on a 256^2 crop; band-limited ASM forward (+ inverse) and Fresnel vs the
    FP64 check mode; FP64 check mode vs numpy; Parseval of the Fresnel
    transform (FP32)."""
    xs = []
    for b in picks:
        v = ctx.wrap(field.ptrs()[0] + b * N * N * 4, field.ptrs()[1] + b * N * N * 4, N, N, 1, "f32")
        re, im = v.download()
        v.free()
        xs.append(re[0] + 1j * im[0])
    x = np.stack(xs)
    z = np.asarray([zs[b] for b in picks])
    out = {"fields": list(picks), "z_m": [round(float(v), 6) for v in z]}

    def run(dtype, op):
        f = ctx.upload(x.real, x.imag, dtype)
        g = ctx.field(N, N, len(picks), dtype)
        op(f, g)
        ctx.sync()
        re, im = g.download()
        f.free()
        g.free()
        return re + 1j * im

    def fft_op(f, g):
        ctx.lib.doe_field_copy(ctx.h, f.h, g.h)
        ctx.fft2d(g)

    y32 = run("f32", fft_op)
    y64 = run("f64", fft_op)
    ynp = np.fft.fft2(x)
    crop = dft_crop(x[0], 896, 896, 256)  # centre of the spectrum
    out["fft2d"] = {"f32_vs_f64_device": _rel(y32, y64), "f32_vs_numpy_f64": _rel(y32, ynp),
                    "f64_device_vs_numpy": _rel(y64, ynp),
                    "f32_vs_naive_dft_crop256": _rel(y32[0, 896:1152, 896:1152], crop),
                    "f64_device_vs_naive_dft_crop256": _rel(y64[0, 896:1152, 896:1152], crop)}
    del y32, y64, ynp

    def asm_rt(f, g):
        ctx.asm(f, g, DX, DX, LAM, z)
    a32 = run("f32", asm_rt)
    a64 = run("f64", asm_rt)

    def asm_back(f, g):
        ctx.asm(f, g, DX, DX, LAM, z)
        ctx.asm(g, g, DX, DX, LAM, -z)
    b32 = run("f32", asm_back)
    b64 = run("f64", asm_back)
    out["asm_band_limited"] = {"forward_f32_vs_f64_device": _rel(a32, a64),
                               "round_trip_f32_vs_f64_device": _rel(b32, b64),
                               "round_trip_f64_vs_input": _rel(b64, x)}
    del a32, a64, b32, b64

    def fres(f, g):
        ctx.fresnel(f, g, DX, DX, LAM, z)
    r32 = run("f32", fres)
    r64 = run("f64", fres)
    odx = LAM * np.abs(z) / (N * DX)
    p_in = np.sum(np.abs(x) ** 2, axis=(1, 2)) * DX * DX
    p_out = np.sum(np.abs(r32) ** 2, axis=(1, 2)) * odx * odx
    out["fresnel"] = {"f32_vs_f64_device": _rel(r32, r64),
                      "parseval_f32_max_abs_dev": float(np.max(np.abs(p_out / p_in - 1)))}
    out["c4_4096sq"] = c4_errors(ctx)
    for k, v in out.items():
        if isinstance(v, dict):
            out[k] = {kk: (float(f"{vv:.3e}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
    out["note"] = ("relative L2 over the picked fields; the device FP64 check mode is pinned to the FP64 "
                   "oracle (<= 1e-12) by the tests; gates: FP32 <= 1e-5, Parseval <= 1e-6")
    return out


def c4_errors(ctx, n: int = 4096) -> dict:
    """The precision sweep's figures on its complex N(0, 1) field at 4096^2:
    fft2d and band-limited ASM (z = 50 mm, 532 nm, 2 um) stored in FP32,
    FP16, BF16 against the FP64 check mode, and fft2d against a naive DFT on
    a 256^2 crop of the spectrum."""
    rng = np.random.default_rng(44)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2)
    dx, lam, z = 2e-6, 532e-9, 0.05

    def run(dt, asm):
        f = ctx.upload(x.real, x.imag, dt)
        if asm:
            ctx.asm(f, f, dx, dx, lam, z)
        else:
            ctx.fft2d(f)
        re, im = f.download()
        f.free()
        return re[0] + 1j * im[0]

    out = {"field": "complex N(0,1), seed 44"}
    crop = dft_crop(x, 1920, 1920, 256)
    for asm in (False, True):
        ref = run("f64", asm)
        name = "asm" if asm else "fft2d"
        if not asm:
            out["fft2d_f64_vs_naive_dft_crop256"] = _rel(ref[1920:2176, 1920:2176], crop)
        for dt in ("f32", "f16", "bf16"):
            y = run(dt, asm)
            out[f"{name}_{dt}_vs_f64"] = _rel(y, ref)
            if not asm:
                out[f"fft2d_{dt}_vs_naive_dft_crop256"] = _rel(y[1920:2176, 1920:2176], crop)
    return out


C1_NODE = r"""
import * as d from './diffractiveopticalelement_amd/js/index.js';
d.setPrecision('f32');
const cfg = __CFG__;
const now = () => Number(process.hrtime.bigint()) / 1e6;
const sys = new d.ProjectorSystem(cfg);
let r = sys.design();
let m = sys.screenMetrics(sys.simulate());
const td = [], ts = [];
for (let i = 0; i < __REPS__; i++) {
  const a = now();
  r = sys.design();
  const b = now();
  m = sys.screenMetrics(sys.simulate());
  ts.push(now() - b);
  td.push(b - a);
}
console.log(JSON.stringify({ design_ms: td, simulate_metrics_ms: ts, rmse: m.rmse, efficiency: m.efficiency,
  correlation: m.correlation, ssim: m.ssim, final_error: r.error[r.error.length - 1] }));
"""


def c1_leg(reps: int = 5) -> dict | None:
    """The small reference configuration through the JavaScript drop-in
    (ProjectorSystem -> N-API addon -> this library): design (100 image-plane
    WGS iterations at 256^2, seed 0, text 'DOE'), then simulate +
    screenMetrics; wall-clock per call, host arrays in and out (the drop-in's
    contract), 1 warm-up call."""
    if shutil.which("node") is None:
        return None
    src = C1_NODE.replace("__CFG__", json.dumps(C1_CFG)).replace("__REPS__", str(reps))
    proc = subprocess.run(["node", "--input-type=module", "-e", src], cwd=str(ROOT), capture_output=True,
                          text=True, timeout=600)
    if proc.returncode != 0:
        return {"error": proc.stderr[-1500:]}
    r = json.loads(proc.stdout.strip().splitlines()[-1])
    return {
        "config": {"workload": "ProjectorSystem n 256, pitch 8 um, 532 nm, z 30 mm, image-plane WGS, "
                               "100 iterations, seed 0, target text 'DOE'; FP32 on the device",
                   "path": "JS drop-in -> N-API -> libdoehip"},
        "design_ms_median": round(float(np.median(r["design_ms"])), 2),
        "simulate_metrics_ms_median": round(float(np.median(r["simulate_metrics_ms"])), 2),
        "design_ms": [round(v, 2) for v in r["design_ms"]],
        "metrics": {k: r[k] for k in ("rmse", "efficiency", "correlation", "ssim")},
        "reference_published": "4.5 s for 60 iterations at 256^2 on 2 vCPU (README.md:204)",
    }


def c3_targets(n: int, rng, polygon: bool) -> np.ndarray:
    """binary target irradiance of the batched uneven GS configuration: 4-64
    Gaussian spots (sigma 1.5 px, uniform weights), or one random filled
    convex polygon (3-8 vertices), inside the central half of the plane"""
    I = np.zeros((n, n))
    lo, hi = n // 4, 3 * n // 4
    if not polygon:
        k = int(rng.integers(4, 65))
        d = np.arange(-4, 5)
        g = np.exp(-(d[None, :] ** 2 + d[:, None] ** 2) / (2 * 1.5 ** 2))
        for y, x in zip(rng.integers(lo + 4, hi - 4, k), rng.integers(lo + 4, hi - 4, k)):
            I[y - 4:y + 5, x - 4:x + 5] += g
        return (I > 0.5).astype(float)
    m = int(rng.integers(3, 9))
    c = rng.uniform(0.4 * n, 0.6 * n, 2)
    r = rng.uniform(0.05 * n, 0.2 * n)
    ang = np.sort(rng.uniform(0, 2 * np.pi, m))
    vy, vx = c[0] + r * np.sin(ang), c[1] + r * np.cos(ang)
    yy, xx = np.mgrid[lo:hi, lo:hi].astype(float)
    inside = np.ones_like(yy, dtype=bool)
    for a in range(m):  # convex: inside every edge's half plane (counter-clockwise)
        b = (a + 1) % m
        inside &= (vx[b] - vx[a]) * (yy - vy[a]) - (vy[b] - vy[a]) * (xx - vx[a]) >= 0
    I[lo:hi, lo:hi] = inside
    return I


def c3_leg(ctx, count: int = 256, iters: int = 200, seed: int = 7, graph: bool = True,
           per_class: bool = False) -> dict:
    """Batched uneven far-field WGS: `count` designs, sides drawn from
    {256, 512, 1024, 2048}, half spot arrays and half filled polygons, binary,
    power-matched to a top-hat illumination; `iters` WGS iterations in one
    graph-captured loop.  value: designs per second; GB/s at 52 B per point
    and iteration (two passes + illumination, target, weights)."""
    from diffractiveopticalelement_amd import warmstart as ws

    rng = np.random.default_rng(seed)
    sides = rng.choice([256, 512, 1024, 2048], count).tolist()
    t0 = time.perf_counter()
    b = nat_batch(ctx, sides)
    ills = {n: ws.illumination(n) for n in set(sides)}
    for i, n in enumerate(sides):
        I = c3_targets(n, rng, polygon=bool(i % 2))
        t = np.sqrt(I)
        t *= np.sqrt(np.sum(ills[n] ** 2) / max(np.sum(t ** 2), 1e-300))
        b.set(i, ills[n], t)
    b.random_init(seed)
    setup_s = time.perf_counter() - t0
    # warm-up: the same run, which captures the iters-iteration graph (each
    # run starts the designs afresh from their initial phases); the timed run
    # replays it
    b.run(iters, graph=graph, per_class=per_class)
    ctx.sync()
    ctx.timer_start()
    b.run(iters, graph=graph, per_class=per_class)
    ms = ctx.timer_stop()
    pts = sum(n * n for n in sides)
    best = [b.get(i)["error"][-1] for i in range(0, count, max(1, count // 8))]
    b.close()
    return {
        "config": {"workload": f"{count} far-field WGS designs, sides from {{256, 512, 1024, 2048}}^2 "
                               f"(seed {seed}), spot arrays and filled polygons, {iters} iterations, "
                               "FP32, " + ("hipGraph-captured" if graph else "eager launches")
                               + (", one launch per pass per size class" if per_class
                                  else ", one launch per pass for the whole batch"),
                   "points": pts,
                   "sides": {str(n): sides.count(n) for n in sorted(set(sides))}},
        "ms": round(ms, 2), "designs_per_s": round(count / (ms * 1e-3), 1),
        "us_per_design_iteration_mean": round(ms * 1e3 / iters / count, 2),
        "gbs_alg": round(52 * pts * iters / (ms * 1e-3) / 1e9, 1), "setup_s": round(setup_s, 1),
        "sample_best_rmse": [round(float(e), 4) for e in best],
    }


def nat_batch(ctx, sides):
    from diffractiveopticalelement_amd import native as nat

    return nat.GsBatch(ctx, sides, nat.ALG_WGS, 0.9)


WSNET = ROOT / "diffractiveopticalelement_amd" / "data" / "wsnet_c5_1024.npy"


def wgs_leg(ctx, count: int = 32, iters: int = 500) -> dict:
    """WGS time to screen RMSE <= 1 % at 1024^2, learned warm start vs plain
    random start, at the learned-warm-start configuration: the network
    (1.75 M parameters, output 1024^2) trained on 20k (target, 500-iteration
    WGS phase) pairs by warmstart.py --stream (the trained parameters ship in
    diffractiveopticalelement_amd/data; training is not repeated here), then
    `count` held-out far-field designs (8 per family: spots, blobs, lines,
    glyphs) run `iters` WGS iterations from each start in the batched,
    graph-captured engine; the warm start's GPU time includes its inference."""
    from diffractiveopticalelement_amd import native as nat
    from diffractiveopticalelement_amd import warmstart as ws

    if not WSNET.exists():
        return {"skipped": f"{WSNET.name} not found"}
    meta = json.loads(WSNET.with_suffix(".json").read_text())
    n, thr = 1024, 0.01
    log = lambda s: print(f"bench.py wgs: {s}", file=sys.stderr, flush=True)  # noqa: E731
    net = nat.WsNet(ctx, n, 32, 8, 0)
    net.set_params(np.load(WSNET))
    r = ws.evaluate(ctx, net, count, n, iters, seed=99999, thr=thr, log=log,
                    modes=("plain", "warm", "restart@25"), label_res=(n,))
    net.close()
    summ = ws.summarize(r, thr)
    spots = {m: summ[m]["families"]["spots"] for m in summ}
    return {
        "config": {"workload": f"{count} held-out far-field WGS designs at {n}^2 (8 per family: spots, blobs, "
                               f"lines, glyphs), top-hat illumination, power-matched targets, dilated ROI; "
                               f"{iters} iterations, beta 0.9, FP32, hipGraph-captured",
                   "threshold": thr, "network_res": n, "params": net.nparams,
                   "training": meta["training"]},
        "summary": summ,
        "spots_gpu_ms_to_thr": {m: spots[m]["median_gpu_ms_to_thr"] for m in spots},
        "spots_plain_over_start": {m: spots[m].get("median_plain_over_mode_gpu_time") for m in spots},
        "note": ("plain WGS reaches 1 % on spot arrays; on the other families it stalls near or above it "
                 "(best rmse per family in the summary; the ratios then use the common threshold "
                 "max(1 %, best plain, best start)). warm: the trained network's start, its inference "
                 "included. labels@1024: each design's own best plain phase at full resolution, the bound "
                 "on a network of this output resolution without inference cost. restart@25: 25 plain "
                 "iterations, then WGS restarted from each design's best iterate"),
    }


TGT_KINDS = [("raster", None), ("disc", {}), ("ring", {"inner": 0.15, "outer": 0.35}),
             ("dots", {"cols": 12, "rows": 9, "missing": [[2, 3], [7, 1]]}), ("zone", {"period": 40}),
             ("checker", {"cells": 10}), ("lines", {"period": 24}), ("raster", "image")]


def targets_leg(ctx, n: int = 1024, count: int = 64, reps: int = 10, img_n: int = 2048) -> dict:
    """Device makeTarget + autoRoi (one doe_target_make_roi call) over a mixed
    batch, and planeToRGB of a field.

    Algorithmic bytes: a target's irradiance written (8 B/pt), read back for the
    ROI and the ROI written (16 B/pt): 24 B/pt.  An image reads the FP32 field
    (8 B/pt) per scan (min/max, 8 radix-select passes, the mapping) and writes
    3 B/pt; its time includes the copy of the pixels to the host."""
    rng = np.random.default_rng(3)
    text = np.kron((rng.random((7, 17)) > 0.5).astype(np.float64), np.ones((6, 6)))
    photo = rng.random((300, 400)) ** 2
    specs = []
    for q in range(count):
        kind, opts = TGT_KINDS[q % len(TGT_KINDS)]
        if kind == "raster":
            specs.append(ctx.target_spec("raster", raster=photo if opts == "image" else text,
                                         **({"gamma": 2.2} if opts == "image" else {})))
        else:
            specs.append(ctx.target_spec(kind, **opts))
    out = ctx.field(n, n, count, "f64")
    ctx.set_flags(asynchronous=True, profile=False)
    ctx.target_make_roi(specs, n, n, out)
    ctx.sync()
    ctx.set_flags(asynchronous=True, profile=True)
    ctx.timer_start()
    for _ in range(reps):
        ctx.target_make_roi(specs, n, n, out)
    call_ms = ctx.timer_stop() / reps
    prof = ctx.profile_report()
    ctx.set_flags(asynchronous=True, profile=False)
    # device time: the bracketed launches of both calls (the call time adds the
    # host's per-call work, e.g. copying 1 MB of rasters into the parameter block)
    ms = sum(tot for _, tot, _ in prof.values()) / reps
    out.free()
    pts = count * n * n
    # planeToRGB (viridis, percentile 0.99) of an FP32 field's intensity, PNG rows
    f = ctx.field(img_n, img_n, 1, "f32")
    re = rng.standard_normal((img_n, img_n))
    f.upload(re, re[::-1].copy())
    viridis = [(0, (68, 1, 84)), (0.25, (59, 82, 139)), (0.5, (33, 145, 140)), (0.75, (94, 201, 98)),
               (1, (253, 231, 37))]
    ctx.set_flags(asynchronous=False, profile=False)
    ctx.field_image(f, quantity="intensity", stops=viridis, percentile=0.99, png_rows=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.field_image(f, quantity="intensity", stops=viridis, percentile=0.99, png_rows=True)
    img_ms = (time.perf_counter() - t0) * 1e3 / reps
    f.free()
    ctx.set_flags(asynchronous=True, profile=False)
    return {
        "targets": count, "n": n, "kinds": sorted({k if o != "image" else "image" for k, o in TGT_KINDS}),
        "ms": round(ms, 4), "call_ms": round(call_ms, 4), "us_per_target": round(ms * 1e3 / count, 2),
        "kernels_ms": {t: round(tot / reps, 4) for t, (_, tot, _) in prof.items()},
        "gbs_alg": round(24 * pts / (ms * 1e-3) / 1e9, 1), "bytes_per_pt": 24,
        "image": {"n": img_n, "quantity": "intensity", "map": "viridis", "percentile": 0.99,
                  "ms_wall_incl_d2h": round(img_ms, 3)},
    }


def fab_leg(ctx, n: int = 1024, batch: int = 64, nb: int = 512) -> dict:
    """errorDiffusionQuantize (4 levels) on the device: one n^2 image with
    plain rows (diagonal wavefront, one workgroup), the same image with
    serpentine rows (one serial chain, rows staged through LDS) and a batch
    of nb^2 serpentine images (one chain per image, one workgroup each)."""
    rng = np.random.default_rng(5)
    out = {}
    f = ctx.upload(rng.uniform(-np.pi, np.pi, (1, n, n)), None, "f64")
    o = ctx.field(n, n, 1, "f64")
    ctx.set_flags(asynchronous=False, profile=False)
    ctx.error_diffusion(f, o, 4, snake=False)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.error_diffusion(f, o, 4, snake=False)
    out["wavefront_ms"] = round((time.perf_counter() - t0) * 1e3 / 3, 3)
    out["wavefront_n"] = n
    ctx.error_diffusion(f, o, 4, snake=True)
    t0 = time.perf_counter()
    ctx.error_diffusion(f, o, 4, snake=True)
    out["serpentine_single_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    f.free()
    o.free()
    fb = ctx.field(nb, nb, batch, "f64")
    ob = ctx.field(nb, nb, batch, "f64")
    for b in range(batch):
        v = ctx.wrap(fb.ptrs()[0] + b * nb * nb * 8, fb.ptrs()[1] + b * nb * nb * 8, nb, nb, 1, "f64")
        ph = rng.uniform(-np.pi, np.pi, (nb, nb))
        v.upload(ph, ph)
        v.free()
    t0 = time.perf_counter()
    ctx.error_diffusion(fb, ob, 4, snake=True)
    out["serpentine_batch"] = {"images": batch, "n": nb, "ms": round((time.perf_counter() - t0) * 1e3, 2)}
    fb.free()
    ob.free()
    ctx.set_flags(asynchronous=True, profile=False)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--chunk-mib", type=int, default=None,
                    help="batch chunk size in MiB of field planes (default: the library's, 0 = whole batch)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the fft2d autotuner (built-in configuration)")
    ap.add_argument("--no-wgs", action="store_true", help="skip the WGS warm-start leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the small reference configuration leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the error statements")
    ap.add_argument("--no-c3", action="store_true", help="skip the batched uneven GS leg")
    ap.add_argument("--c3-count", type=int, default=256, help="designs of the batched uneven GS leg")
    ap.add_argument("--c3-iters", type=int, default=200, help="iterations of the batched uneven GS leg")
    ap.add_argument("--c3-per-class", action="store_true",
                    help="run the batched uneven GS leg with one launch per pass per size class "
                         "instead of one per pass for the whole batch")
    ap.add_argument("--c3-eager", action="store_true",
                    help="launch the batched uneven GS iterations one by one instead of from a graph "
                         "(for profilers that count per dispatch)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP64 / FP16 / BF16 storage variants of the fft2d step")
    ap.add_argument("--no-targets", action="store_true",
                    help="skip the target / figure and fabrication legs")
    ap.add_argument("--wgs-eval", type=int, default=32)
    ap.add_argument("--wgs-iters", type=int, default=500)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy the 512 MiB batch once (device-to-device), so a profiler's "
                         "byte counters can be scaled against a copy of known size")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the hot path shards by instance "
              "with no communication); run with --gpus 1", file=sys.stderr)
        return 2

    from diffractiveopticalelement_amd import native as nat

    ctx = nat.Context(0)
    if args.chunk_mib is not None:
        ctx.set_chunk_bytes(args.chunk_mib << 20)
    rng = np.random.default_rng(0)
    c = (np.arange(N) - N / 2) * DX
    w = 0.3 * N * DX
    amp = np.exp(-(c[None, :] ** 2 + c[:, None] ** 2) / w ** 2)
    zs = rng.uniform(0.010, 0.100, B)
    field = ctx.field(N, N, B, "f32")
    scratch = ctx.field(N, N, B, "f32")
    for b in range(B):  # upload field by field (host memory stays small)
        ph = rng.uniform(0, 2 * np.pi, (N, N))
        view = ctx.wrap(field.ptrs()[0] + b * N * N * 4, field.ptrs()[1] + b * N * N * 4, N, N, 1, "f32")
        view.upload(amp * np.cos(ph), amp * np.sin(ph))
        view.free()
    pts = B * N * N
    if args.copy_calibration:
        ctx.sync()
        ctx.lib.doe_field_copy(ctx.h, field.h, scratch.h)
        ctx.sync()

    # steps alternate forward and inverse transforms (the same traffic), so
    # the batch stays finite and an even number of steps returns it to the
    # input: the round trip after the timed loop is the run's checksum
    v0 = ctx.wrap(field.ptrs()[0], field.ptrs()[1], N, N, 1, "f32")
    x0 = np.concatenate(v0.download())
    tuning = None
    if not args.no_tune:
        # pick the fastest configuration for this shape (outside the timed
        # region); the warm-up follows at once, as a GPU left idle for a
        # second runs the next few steps 5-8 % slower while its clock rises
        rep_t = ctx.autotune_fft2d(N, N, B, "f32", reps=3)
        tuning = ctx.tune_get(N, N, B, "f32")
        tuning["candidates"] = len(rep_t.strip().splitlines()) - 1
    ctx.set_flags(asynchronous=True, profile=False)
    for i in range(args.warmup):
        ctx.fft2d(field, inverse=bool(i & 1))
    ctx.sync()
    ctx.set_flags(asynchronous=True, profile=True)
    t0 = time.perf_counter()
    ctx.timer_start()
    for i in range(args.steps):
        ctx.fft2d(field, inverse=bool((args.warmup + i) & 1))
    ms_events = ctx.timer_stop()
    t1 = time.perf_counter()
    rep = ctx.profile_report()
    ms_step = ms_events / args.steps
    value = 32 * pts / (ms_step * 1e-3) / 1e9
    ctx.set_flags(asynchronous=True, profile=False)
    if (args.warmup + args.steps) & 1:
        ctx.fft2d(field, inverse=True)  # close the last round trip (untimed)
    ctx.sync()
    x1 = np.concatenate(v0.download())
    v0.free()
    checksum = {"round_trips": (args.warmup + args.steps + 1) // 2,
                "field0_rel_l2_vs_input": float(f"{_rel(x1, x0):.3e}"),
                "all_finite": bool(np.all(np.isfinite(x1)))}

    kernels = {}
    for tag, (cnt, tot, kpts) in rep.items():
        avg_us = tot / cnt * 1e3
        kernels[tag] = {"launches": cnt, "avg_us": round(avg_us, 2), "pts_per_launch": kpts // cnt,
                        "gbs_alg": round(16 * kpts / (tot * 1e-3) / 1e9, 1)}
    # dominant kernel: the largest share of the step's time
    dom_tag = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    dom = kernels[dom_tag]

    # rest of the configuration's pipeline: ASM fwd + inv (band-limited), Fresnel fwd
    # three untimed rounds first: the clock has dropped while the host
    # prepared the leg, and the first rounds run 5-10 % slower
    ctx.set_flags(asynchronous=True, profile=False)
    for _ in range(3):
        ctx.asm(field, scratch, DX, DX, LAM, zs)
        ctx.asm(scratch, scratch, DX, DX, LAM, -zs)
        ctx.fresnel(field, scratch, DX, DX, LAM, zs)
    ctx.sync()
    reps = max(5, args.steps // 2)
    ctx.set_flags(asynchronous=True, profile=True)
    ctx.timer_start()
    for _ in range(reps):
        ctx.asm(field, scratch, DX, DX, LAM, zs)
        ctx.asm(scratch, scratch, DX, DX, LAM, -zs)
        ctx.fresnel(field, scratch, DX, DX, LAM, zs)
    ms_pipe = ctx.timer_stop() / reps
    prep = ctx.profile_report()
    pipe_bytes = (48 + 48 + 32) * pts
    pipeline = {
        "ms": round(ms_pipe, 4),
        "gbs_alg": round(pipe_bytes / (ms_pipe * 1e-3) / 1e9, 1),
        "bytes_per_pt": {"asm": 48, "fresnel": 32},
        "kernels": {t: {"launches": c, "avg_us": round(s / c * 1e3, 2)} for t, (c, s, _) in prep.items()},
    }

    ctx.set_flags(asynchronous=True, profile=False)
    ctx.sync()
    scratch.free()
    errors = None if args.no_errors else errors_leg(ctx, field, zs)
    variants = None if args.no_variants else storage_variants(ctx, field, args.steps)
    field.free()
    c3 = None if args.no_c3 else c3_leg(ctx, args.c3_count, args.c3_iters, graph=not args.c3_eager,
                                       per_class=args.c3_per_class)
    targets = None if args.no_targets else targets_leg(ctx)
    fab = None if args.no_targets else fab_leg(ctx)
    c1 = None if args.no_c1 else c1_leg()
    wgs = None if args.no_wgs else wgs_leg(ctx, args.wgs_eval, args.wgs_iters)

    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": "fft2d HBM GB/s (16 x 2048^2 complex FP32 batch; WGS time-to-RMSE<=1% at 1024^2 reported separately)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round((t1 - t0) * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Gaussian amplitude x seeded uniform phase, seed 0)",
        "config": {"workload": "batched propagation: 16 x 2048^2 complex FP32, pitch 4 um, "
                               "lambda 633 nm, z ~ U[10,100] mm; step = one fft2d of the batch "
                               "(forward and inverse alternate)",
                   "batch": B, "n": N},
        "checksum": checksum,
        "errors": errors,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_tag,
            "achieved": dom["gbs_alg"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(dom["gbs_alg"] / PEAK_HBM_GBS, 4),
            "traffic": traffic_of(dom_tag),
            "alg_bytes_per_launch": 16 * dom["pts_per_launch"],
        },
        "chunk_bytes": ctx.chunk_bytes(),
        "tuning": tuning,
        "kernels": kernels,
        "pipeline": pipeline,
        "storage_variants": variants,
        "gs_batch_uneven": c3,
        "targets": targets,
        "fabrication": fab,
        "c1_projector": c1,
        "wgs": wgs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

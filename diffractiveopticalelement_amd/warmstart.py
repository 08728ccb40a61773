"""Learned warm start for batched far-field WGS: target families, training
pairs from the batched engine, training, and the warm-vs-plain comparison.

Target families (seeded, square side n, no counterpart in the reference; the
reference's own text/disc/spot targets live in js/projector.js):
  spots   4..64 Gaussian spots (sigma 1.5 px), uniform weights
  blobs   2..8 Gaussian blobs, sigma 0.01..0.04 n
  lines   1..6 line segments, 3 px wide
  glyphs  2..5 characters of the 5x7 glyph table, block-scaled
Each is an irradiance I >= 0 inside the central half of the plane (blob
tails reach up to 2 sigma past it); the
design target is t = sqrt(I) power-matched to the illumination (a top-hat
aperture of radius 0.45 n with a 2 px raised-cosine edge), and the ROI is
the support I > 0.05 max, dilated by 2 px.

Usage (GPU):
    python -m diffractiveopticalelement_amd.warmstart --train 512 --steps 1500 --eval 32
    python -m diffractiveopticalelement_amd.warmstart --stream --train 20000 --label-iters 500 \
        --res 1024 --eval 32 --eval-iters 500 --save wsnet.npy --out result.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import native as nat

# 5x7 glyphs as row bit masks (bit 4 = leftmost column), for the glyph family
_GLYPHS = {
    "A": [14, 17, 17, 31, 17, 17, 17], "B": [30, 17, 17, 30, 17, 17, 30],
    "C": [14, 17, 16, 16, 16, 17, 14], "D": [30, 17, 17, 17, 17, 17, 30],
    "E": [31, 16, 16, 30, 16, 16, 31], "F": [31, 16, 16, 30, 16, 16, 16],
    "H": [17, 17, 17, 31, 17, 17, 17], "L": [16, 16, 16, 16, 16, 16, 31],
    "O": [14, 17, 17, 17, 17, 17, 14], "P": [30, 17, 17, 30, 16, 16, 16],
    "T": [31, 4, 4, 4, 4, 4, 4], "X": [17, 17, 10, 4, 10, 17, 17],
}
FAMILIES = ("spots", "blobs", "lines", "glyphs")


def illumination(n: int) -> np.ndarray:
    c = np.arange(n) - n / 2 + 0.5
    r = np.hypot(c[None, :], c[:, None])
    rad, soft = 0.45 * n, 2.0
    v = 0.5 * (1 + np.cos(np.pi * (r - rad + soft) / soft))
    return np.where(r <= rad - soft, 1.0, np.where(r >= rad, 0.0, v))


def make_target(kind: str, n: int, rng: np.random.Generator) -> np.ndarray:
    """irradiance of one target of family `kind` (float64, n x n)"""
    I = np.zeros((n, n))
    lo, hi = n // 4, 3 * n // 4
    if kind == "spots":
        k = int(rng.integers(4, 65))
        ys = rng.integers(lo + 4, hi - 4, k)
        xs = rng.integers(lo + 4, hi - 4, k)
        d = np.arange(-4, 5)
        g = np.exp(-(d[None, :] ** 2 + d[:, None] ** 2) / (2 * 1.5 ** 2))
        for y, x in zip(ys, xs):
            I[y - 4:y + 5, x - 4:x + 5] += g
    elif kind == "blobs":
        k = int(rng.integers(2, 9))
        yy, xx = np.mgrid[0:n, 0:n]
        for _ in range(k):
            s = rng.uniform(0.01, 0.04) * n
            cy, cx = rng.uniform(lo + 2 * s, hi - 2 * s, 2)
            r0, r1 = int(max(0, cy - 4 * s)), int(min(n, cy + 4 * s + 1))
            c0, c1 = int(max(0, cx - 4 * s)), int(min(n, cx + 4 * s + 1))
            I[r0:r1, c0:c1] += rng.uniform(0.5, 1.0) * np.exp(
                -((yy[r0:r1, c0:c1] - cy) ** 2 + (xx[r0:r1, c0:c1] - cx) ** 2) / (2 * s * s))
    elif kind == "lines":
        k = int(rng.integers(1, 7))
        yy, xx = np.mgrid[lo:hi, lo:hi].astype(float)
        sub = np.zeros((hi - lo, hi - lo))
        for _ in range(k):
            p = rng.uniform(lo + 8, hi - 8, 2)
            q = rng.uniform(lo + 8, hi - 8, 2)
            d = q - p
            t = np.clip(((yy - p[0]) * d[0] + (xx - p[1]) * d[1]) / max(d @ d, 1e-9), 0, 1)
            dist = np.hypot(yy - (p[0] + t * d[0]), xx - (p[1] + t * d[1]))
            sub = np.maximum(sub, (dist <= 1.5).astype(float))
        I[lo:hi, lo:hi] = sub
    elif kind == "glyphs":
        text = "".join(rng.choice(sorted(_GLYPHS), int(rng.integers(2, 6))))
        raster = np.zeros((7, 6 * len(text) - 1))
        for i, ch in enumerate(text):
            for r, bits in enumerate(_GLYPHS[ch]):
                for c in range(5):
                    if bits >> (4 - c) & 1:
                        raster[r, 6 * i + c] = 1.0
        s = max(1, int((hi - lo) * 0.9 / raster.shape[1]))
        big = np.kron(raster, np.ones((s, s)))
        y0 = (n - big.shape[0]) // 2
        x0 = (n - big.shape[1]) // 2
        I[y0:y0 + big.shape[0], x0:x0 + big.shape[1]] = big
    else:
        raise ValueError(kind)
    return I


def design_inputs(I: np.ndarray, ill: np.ndarray):
    """target amplitude (power-matched to the illumination) and binary ROI"""
    t = np.sqrt(I)
    t *= np.sqrt(np.sum(ill ** 2) / np.sum(t ** 2))
    roi = I > 0.05 * I.max()
    for _ in range(2):
        p = np.pad(roi, 1)
        roi = (p[1:-1, 1:-1] | p[:-2, 1:-1] | p[2:, 1:-1] | p[1:-1, :-2] | p[1:-1, 2:])
    return t, roi.astype(float)


def fill_batch(ctx, count: int, n: int, seed: int, iters_unused=None):
    rng = np.random.default_rng(seed)
    b = nat.GsBatch(ctx, [n] * count, nat.ALG_WGS, 0.9, with_roi=True)
    ill = illumination(n)
    kinds = []
    for i in range(count):
        kind = FAMILIES[i % len(FAMILIES)]
        t, roi = design_inputs(make_target(kind, n, rng), ill)
        b.set(i, ill, t, roi)
        kinds.append(kind)
    return b, kinds


def quadratic_phasors(x: np.ndarray, n: int, spread: float = 0.25) -> np.ndarray:
    """Target-plane phasors exp(i a |u - u0|^2) per pooled input x[d] (res x res),
    u0 the irradiance centroid, a chosen so the target's RMS radius r maps to
    spread * n in the element plane (a chirp of curvature a moves a target
    point at radius r to element radius a r n / pi).  Returns
    (count, res, res, 2) float32."""
    count, res, _ = x.shape
    f = n / res
    v, u = np.mgrid[0:res, 0:res].astype(float)
    out = np.zeros((count, res, res, 2), np.float32)
    for d in range(count):
        w = x[d] / max(x[d].sum(), 1e-30)
        cu, cv = (w * u).sum(), (w * v).sum()
        r2 = (w * ((u - cu) ** 2 + (v - cv) ** 2)).sum()
        r = max(np.sqrt(r2) * f, 4.0)  # full-resolution pixels
        a = spread * np.pi / r  # rad per full-resolution pixel^2
        ph = a * f * f * ((u - cu) ** 2 + (v - cv) ** 2)
        out[d, ..., 0] = np.cos(ph)
        out[d, ..., 1] = np.sin(ph)
    return out


def _phasor_field(ctx, z: np.ndarray):
    """device copy of (count, res, res, 2) interleaved phasors"""
    count, res = z.shape[0], z.shape[1]
    zf = ctx.field(2 * res, res, count, "f32")
    zf.upload(z.reshape(count, res, 2 * res), np.zeros((count, res, 2 * res)))
    return zf


def start_quadratic(ctx, b, count: int, n: int, res: int, spread: float = 0.25):
    """replace every design's initial phase by the quadratic target-plane
    start, computed on the device (quadratic_phasors is its host form)"""
    scratch = ctx.field(res, res, 3 * count, "f32")
    b.warm_apply_quadratic(res, spread, scratch.ptrs()[0])
    scratch.free()


def make_pairs(ctx, count: int, n: int, res: int, iters: int, seed: int, chunk: int = 64):
    """(inputs, label_re, label_im, weights) of `count` designs, res x res each;
    labels are the best target-plane phases of `iters` WGS iterations started
    from the quadratic target-plane phase, a deterministic function of the
    target (a random start would give every target many unrelated labels)"""
    X, LR, LI, W = [], [], [], []
    for c0 in range(0, count, chunk):
        m = min(chunk, count - c0)
        b, _ = fill_batch(ctx, m, n, seed + c0)
        start_quadratic(ctx, b, m, n, res)
        b.run(iters, graph=True)
        inp = ctx.field(res, res, m, "f32")
        lab = ctx.field(res, res, m, "f32")
        wts = ctx.field(res, res, m, "f32")
        b.warm_inputs(res, inp.ptrs()[0])
        b.warm_labels(res, lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
        x, _ = inp.download()
        lr, li = lab.download()
        w, _ = wts.download()
        X.append(x)
        LR.append(lr)
        LI.append(li)
        W.append(w)
        for f in (inp, lab, wts):
            f.free()
        b.close()
    return np.concatenate(X), np.concatenate(LR), np.concatenate(LI), np.concatenate(W)


def train(ctx, data, res: int, base: int = 32, batch: int = 8, steps: int = 1000,
          lr: float = 1e-3, seed: int = 0, log=print):
    X, LR, LI, W = data
    net = nat.WsNet(ctx, res, base, batch, seed)
    inp = ctx.field(res, res, batch, "f32")
    lab = ctx.field(res, res, batch, "f32")
    wts = ctx.field(res, res, batch, "f32")
    rng = np.random.default_rng(seed)
    z = np.zeros((batch, res, res))
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = rng.choice(len(X), batch, replace=False)
        inp.upload(X[idx], z)
        lab.upload(LR[idx], LI[idx])
        wts.upload(W[idx], z)
        loss, _ = net.loss_grad(batch, inp.ptrs()[0], lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
        step_lr = lr * min(1.0, (s + 1) / 50) * (0.5 * (1 + np.cos(np.pi * s / steps)))
        net.adam_step(step_lr)
        hist.append(loss)
        if s % max(1, steps // 10) == 0 or s == steps - 1:
            log(f"  step {s}: loss {np.mean(hist[-20:]):.4f} ({time.perf_counter() - t0:.1f} s)")
    for f in (inp, lab, wts):
        f.free()
    return net, hist


def _gen_chunk(job):
    """host side of a training chunk: (target amplitude float32, ROI uint8)
    of `count` designs of side n from `seed`, in fill_batch's order (runs in
    a worker process)"""
    seed, count, n = job
    rng = np.random.default_rng(seed)
    ill = illumination(n)
    out = []
    for i in range(count):
        t, roi = design_inputs(make_target(FAMILIES[i % len(FAMILIES)], n, rng), ill)
        out.append((t.astype(np.float32), roi.astype(np.uint8)))
    return out


def train_stream(ctx, pool, total: int, n: int, res: int, label_iters: int, chunk: int = 64,
                 batch: int = 8, base: int = 32, lr: float = 1e-3, seed: int = 1000, log=print,
                 deadline: float | None = None):
    """Train on `total` (target, label) pairs generated and consumed chunk by
    chunk, each pair used once: a worker pool builds the next chunks' targets
    on the host while the GPU runs the current chunk's `label_iters` WGS
    iterations from the quadratic start and trains on its pairs (res x res
    inputs, labels and weights stay on the device).  Stops early at
    `deadline` (time.perf_counter()); returns the network, the loss history
    and the pairs used."""
    from collections import deque

    net = nat.WsNet(ctx, res, base, batch, seed)
    ill = illumination(n)
    steps_total = max(1, total // batch)
    jobs = [(seed + c0, min(chunk, total - c0), n) for c0 in range(0, total, chunk)]
    pending = deque(pool.apply_async(_gen_chunk, (j,)) for j in jobs[:4])
    nxt = 4
    hist, used, step = [], 0, 0
    t0 = time.perf_counter()
    for q, job in enumerate(jobs):
        data = pending.popleft().get()
        if nxt < len(jobs):
            pending.append(pool.apply_async(_gen_chunk, (jobs[nxt],)))
            nxt += 1
        m = job[1]
        b = nat.GsBatch(ctx, [n] * m, nat.ALG_WGS, 0.9, with_roi=True)
        for i, (t, roi) in enumerate(data):
            b.set(i, ill, t, roi)
        start_quadratic(ctx, b, m, n, res)
        b.run(label_iters, graph=True)
        inp = ctx.field(res, res, m, "f32")
        lab = ctx.field(res, res, m, "f32")
        wts = ctx.field(res, res, m, "f32")
        b.warm_inputs(res, inp.ptrs()[0])
        b.warm_labels(res, lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
        b.close()
        stride = batch * res * res * 4
        for k in range(m // batch):
            o = k * stride
            loss, _ = net.loss_grad(batch, inp.ptrs()[0] + o, lab.ptrs()[0] + o, lab.ptrs()[1] + o,
                                    wts.ptrs()[0] + o)
            sl = lr * min(1.0, (step + 1) / 50) * (0.5 * (1 + np.cos(np.pi * min(step, steps_total) / steps_total)))
            net.adam_step(sl)
            hist.append(loss)
            step += 1
        used += m
        for f in (inp, lab, wts):
            f.free()
        if q % 10 == 0 or q == len(jobs) - 1:
            log(f"  pairs {used}/{total}, step {step}: loss {np.mean(hist[-40:]):.4f} "
                f"({time.perf_counter() - t0:.0f} s)")
        if deadline is not None and time.perf_counter() > deadline:
            log(f"  stopping at the time limit after {used} pairs")
            break
    return net, hist, used


def first_below(h, thr):
    k = np.nonzero(np.asarray(h) <= thr)[0]
    return int(k[0]) + 1 if k.size else None


def evaluate(ctx, net, count: int, n: int, iters: int, seed: int, thr: float = 0.01, log=print,
             modes=("plain", "quadratic", "warm"), label_res=()):
    """WGS from several starts on the same `count` targets.

    plain      random element phase (the method without the network)
    quadratic  the quadratic target-plane phase of start_quadratic
    warm       the network's target-plane phase
    labels@r   each design's own best plain-run far-field phase, pooled to r^2
               and handed off like the network's output: what a network that
               reproduced its training labels exactly at resolution r would
               give (an upper bound for the learned start at that resolution)
    restart@k  plain for k iterations, then a restart from each design's best
               iterate (doe_gs_batch_restart_best) for the remaining
               iterations; its history and GPU time cover both stages

    Per design: the screen rmse history, the first iteration at or below
    `thr`, the best rmse, and for every mode but plain the iterations to the
    common threshold max(thr, best plain, best of the mode) in both runs."""
    out = {}
    labels = {}
    scr_plain = None
    order = ["plain"] + [m for m in modes if m != "plain"]
    for r in label_res:
        order.append(f"labels@{r}")
    for mode in order:
        b, kinds = fill_batch(ctx, count, n, seed)
        t_inf = 0.0
        if mode == "plain":
            b.random_init(seed)
        elif mode == "quadratic":
            start_quadratic(ctx, b, count, n, net.res)
        elif mode.startswith("restart@"):
            pass  # the stages run below
        elif mode == "warm":
            scratch = ctx.field(net.res, net.res, 4 * count, "f32")  # >= 3 count res^2 floats
            ctx.sync()
            ctx.timer_start()
            b.warm_apply(net, scratch.ptrs()[0])
            t_inf = ctx.timer_stop()
            scratch.free()
        else:
            r = int(mode.split("@")[1])
            zf = _phasor_field(ctx, labels[r])
            b.warm_apply_phasor(r, zf.ptrs()[0])
            zf.free()
        if mode.startswith("restart@"):
            k = int(mode.split("@")[1])
            stages = [k, iters - k]
            b.random_init(seed)
            ms, parts = 0.0, []
            for q, m in enumerate(stages):
                ctx.timer_start()
                if q:
                    b.restart_best()
                b.run(m, graph=True)
                ms += ctx.timer_stop()
                parts.append(np.array([b.get_screen(i) for i in range(count)]))  # the next run overwrites them
            scr = np.concatenate(parts, axis=1)
        else:
            ctx.timer_start()
            b.run(iters, graph=True)
            ms = ctx.timer_stop()
            scr = np.array([b.get_screen(i) for i in range(count)])
        if mode == "plain":
            scr_plain = scr
            for r in label_res:
                lab = ctx.field(r, r, count, "f32")
                wts = ctx.field(r, r, count, "f32")
                b.warm_labels(r, lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
                lr_, li_ = lab.download()
                labels[r] = np.stack([lr_, li_], axis=-1).astype(np.float32)
                lab.free()
                wts.free()
        b.close()
        its = [first_below(h, thr) for h in scr]
        per_iter = ms / iters / count
        rec = {
            "ms_batch": ms, "ms_per_design_iteration": per_iter, "inference_ms_batch": t_inf,
            "iterations_to_thr": its, "final_screen_rmse": scr[:, -1].tolist(),
            "best_screen_rmse": scr.min(axis=1).tolist(), "kinds": kinds,
            "screen_rmse_median": np.median(scr, axis=0)[[0, 9, 49, 99, iters - 1]].tolist()
            if iters >= 100 else None,
        }
        if mode != "plain":
            common = np.maximum(np.maximum(scr_plain.min(axis=1), scr.min(axis=1)), thr)
            rec["common_threshold"] = common.tolist()
            rec["iterations_to_common"] = [first_below(h, c) for h, c in zip(scr, common)]
            rec["plain_iterations_to_common"] = [first_below(h, c) for h, c in zip(scr_plain, common)]
        out[mode] = rec
        reached = [k for k in its if k is not None]
        log(f"  {mode}: {len(reached)}/{count} designs reach screen rmse <= {thr}; median "
            f"iterations {np.median(reached) if reached else None}; final rmse median "
            f"{np.median(scr[:, -1]):.4f}; {per_iter * 1e3:.2f} us per design-iteration")
        for kind in FAMILIES:
            sel = [i for i, k in enumerate(kinds) if k == kind]
            log(f"    {kind}: iterations {[its[i] for i in sel]}, screen rmse at 10/50/final "
                f"{np.round(np.median(scr[sel][:, [min(9, iters - 1), min(49, iters - 1), -1]], axis=0), 4).tolist()}")
    return out


def summarize(res: dict, thr: float = 0.01) -> dict:
    """per mode and target family: designs reaching `thr`, median iterations and
    GPU ms per design to `thr`, median best rmse, and the median ratio of plain's
    GPU time to the mode's GPU time at the common threshold (> 1: the start helps)"""
    plain = res["plain"]
    out = {}
    for mode, r in res.items():
        if not isinstance(r, dict) or "kinds" not in r:
            continue
        kinds = r["kinds"]
        count = len(kinds)
        inf_ms = r["inference_ms_batch"] / count
        fam = {}
        for kind in FAMILIES:
            sel = [i for i, k in enumerate(kinds) if k == kind]
            its = [r["iterations_to_thr"][i] for i in sel if r["iterations_to_thr"][i] is not None]
            e = {"designs": len(sel), "reach_thr": len(its),
                 "median_iterations_to_thr": float(np.median(its)) if its else None,
                 "median_gpu_ms_to_thr": round(float(np.median(its)) * r["ms_per_design_iteration"]
                                               + inf_ms, 4) if its else None,
                 "median_best_rmse": round(float(np.median([r["best_screen_rmse"][i] for i in sel])), 5)}
            if mode != "plain":
                ratios = []
                for i in sel:
                    a, b_ = r["plain_iterations_to_common"][i], r["iterations_to_common"][i]
                    if a is None or b_ is None:
                        continue
                    ratios.append(a * plain["ms_per_design_iteration"]
                                  / (b_ * r["ms_per_design_iteration"] + inf_ms))
                e["median_common_thr"] = round(float(np.median([r["common_threshold"][i] for i in sel])), 5)
                e["median_plain_over_mode_gpu_time"] = round(float(np.median(ratios)), 3) if ratios else None
            fam[kind] = e
        out[mode] = {"us_per_design_iteration": round(r["ms_per_design_iteration"] * 1e3, 2),
                     "inference_ms_per_design": round(inf_ms, 4), "families": fam}
    return out


def main_stream(args) -> int:
    """The learned warm start at full scale: `--train` pairs streamed through
    training (see train_stream), then the evaluation; the network's
    parameters are saved with --save."""
    import multiprocessing as mp

    t_start = time.perf_counter()
    # the worker processes start before this process touches the GPU
    pool = mp.get_context("fork").Pool(args.workers)
    ctx = nat.Context(0)
    deadline = t_start + args.max_seconds if args.max_seconds else None
    print(f"streaming {args.train} pairs at {args.n}^2 ({args.label_iters} WGS iterations each), "
          f"network output {args.res}^2", flush=True)
    net, hist, used = train_stream(ctx, pool, args.train, args.n, args.res, args.label_iters,
                                   log=lambda s: print(s, flush=True), deadline=deadline)
    pool.close()
    t_train = time.perf_counter() - t_start
    if args.save:
        np.save(args.save, net.params())
    print("evaluation", flush=True)
    res = evaluate(ctx, net, args.eval, args.n, args.eval_iters, seed=99999, thr=args.threshold,
                   log=lambda s: print(s, flush=True), modes=("plain", "quadratic", "warm", "restart@25"),
                   label_res=[args.n])
    res["training"] = {"pairs": used, "label_iterations": args.label_iters, "steps": len(hist),
                       "batch": 8, "network_res": args.res, "params": net.nparams,
                       "loss_first_last": [float(np.mean(hist[:40])), float(np.mean(hist[-40:]))],
                       "seconds": round(t_train, 1)}
    res["summary"] = summarize(res, args.threshold)
    print(json.dumps({"training": res["training"], "summary": res["summary"]}, indent=1), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--train", type=int, default=256, help="training designs")
    ap.add_argument("--label-iters", type=int, default=300)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--eval", type=int, default=32)
    ap.add_argument("--eval-iters", type=int, default=300)
    ap.add_argument("--threshold", type=float, default=0.01)
    ap.add_argument("--label-res", type=int, nargs="*", default=None,
                    help="resolutions of the label-bound starts (default: the network's and n)")
    ap.add_argument("--out", default=None, help="write the result JSON here")
    ap.add_argument("--stream", action="store_true",
                    help="stream the pairs through training (full-scale runs; see train_stream)")
    ap.add_argument("--workers", type=int, default=8, help="host processes building targets (--stream)")
    ap.add_argument("--max-seconds", type=float, default=None, help="stop streaming after this long")
    ap.add_argument("--save", default=None, help="save the trained parameters (.npy, --stream)")
    args = ap.parse_args(argv)
    if args.stream:
        return main_stream(args)
    label_res = args.label_res if args.label_res is not None else [args.res, args.n]
    ctx = nat.Context(0)
    t0 = time.perf_counter()
    print(f"pairs: {args.train} designs at {args.n}^2, {args.label_iters} WGS iterations", flush=True)
    data = make_pairs(ctx, args.train, args.n, args.res, args.label_iters, seed=1000)
    print(f"  {time.perf_counter() - t0:.1f} s", flush=True)
    print(f"training {args.steps} steps", flush=True)
    net, hist = train(ctx, data, args.res, steps=args.steps,
                      log=lambda s: print(s, flush=True))
    print("evaluation", flush=True)
    res = evaluate(ctx, net, args.eval, args.n, args.eval_iters, seed=99999, thr=args.threshold,
                   log=lambda s: print(s, flush=True), label_res=label_res)
    res["train_loss_first_last"] = [float(np.mean(hist[:20])), float(np.mean(hist[-20:]))]
    res["summary"] = summarize(res, args.threshold)
    print(json.dumps(res["summary"], indent=1), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())

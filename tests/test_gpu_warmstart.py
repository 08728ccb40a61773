"""What the learned warm start gains over plain random-start WGS, measured on
the GPU at a small size (256^2 designs, network at 64^2), next to the bound
any network of that output resolution could reach (each design's own best
phase pooled to 64^2) and the full-resolution bound (its own phase at 256^2).

The gains are reported, not assumed: the test checks that every start's run
is complete and finite, that the full-resolution bound does reach 1 % on
spot arrays sooner than plain WGS (the hand-off itself works), and prints the
per-family comparison.
"""
import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
from diffractiveopticalelement_amd import warmstart as ws

pytestmark = pytest.mark.gpu


def test_trained_start_against_plain_and_label_bounds():
    ctx = nat.Context(0)
    n, res, iters, count = 256, 64, 200, 16
    data = ws.make_pairs(ctx, 32, n, res, 150, seed=1000)
    net, hist = ws.train(ctx, data, res, base=16, steps=150, log=lambda s: None)
    assert np.all(np.isfinite(hist))
    out = ws.evaluate(ctx, net, count, n, iters, seed=4242, thr=0.01, log=print,
                      modes=("plain", "warm"), label_res=(res, n))
    summ = ws.summarize(out, 0.01)
    for mode in ("plain", "warm", f"labels@{res}", f"labels@{n}"):
        r = out[mode]
        assert len(r["best_screen_rmse"]) == count
        assert np.all(np.isfinite(r["best_screen_rmse"]))
        assert r["ms_per_design_iteration"] > 0
        print(mode, {k: (v["reach_thr"], v["median_iterations_to_thr"], v["median_best_rmse"],
                         v.get("median_plain_over_mode_gpu_time"))
                     for k, v in summ[mode]["families"].items()})
    assert out["warm"]["inference_ms_batch"] > 0
    spots = [i for i, k in enumerate(out["plain"]["kinds"]) if k == "spots"]
    full = out[f"labels@{n}"]
    plain_its = [out["plain"]["iterations_to_thr"][i] for i in spots]
    full_its = [full["iterations_to_thr"][i] for i in spots]
    assert all(k is not None for k in plain_its + full_its), (plain_its, full_its)
    assert np.median(full_its) < 0.75 * np.median(plain_its), (plain_its, full_its)
    net.close()
    ctx.close()

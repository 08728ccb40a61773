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


def test_restart_from_best_equals_the_full_resolution_handoff():
    """doe_gs_batch_restart_best = the labels@n hand-off (each design's best
    target-plane phasor at full resolution, applied through
    doe_warmstart_apply_phasor), done on the device in one call."""
    ctx = nat.Context(0)
    n, count, seed = 256, 8, 77
    a, kinds = ws.fill_batch(ctx, count, n, seed)
    a.random_init(seed)
    a.run(25)
    lab = ctx.field(n, n, count, "f32")
    wts = ctx.field(n, n, count, "f32")
    a.warm_labels(n, lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
    lr_, li_ = lab.download()
    z = np.stack([lr_, li_], axis=-1).astype(np.float32)
    lab.free()
    wts.free()
    a.restart_best()
    a.run(5)
    sa = np.array([a.get_screen(i) for i in range(count)])
    b, _ = ws.fill_batch(ctx, count, n, seed)
    b.random_init(seed)
    b.run(25)
    zf = ws._phasor_field(ctx, z)
    b.warm_apply_phasor(n, zf.ptrs()[0])
    zf.free()
    b.run(5)
    sb = np.array([b.get_screen(i) for i in range(count)])
    assert np.all(np.isfinite(sa))
    np.testing.assert_allclose(sa, sb, rtol=1e-3, atol=1e-5)
    a.close()
    b.close()
    # the evaluation's restart mode: both stages in the history and the time
    out = ws.evaluate(ctx, type("N", (), {"res": 64})(), count, n, 60, seed=seed, thr=0.01, log=print,
                      modes=("plain", "restart@25"))
    r = out["restart@25"]
    assert len(r["best_screen_rmse"]) == count and r["ms_per_design_iteration"] > 0
    assert np.all(np.isfinite(r["best_screen_rmse"]))
    ctx.close()

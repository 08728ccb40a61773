"""What the learned warm start gains over plain random-start WGS.

* The network trained at the learned-warm-start configuration (20k pairs of
  500-iteration labels, output 1024^2; parameters shipped in
  diffractiveopticalelement_amd/data) on the benchmark's 32 held-out 1024^2
  designs (seed 99999, 500 iterations): the measured iterations and GPU-time
  ratios are asserted.
* The training path at a small size (256^2 designs, network at 64^2), next
  to the bounds of the hand-off.
"""
from pathlib import Path

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
from diffractiveopticalelement_amd import warmstart as ws

pytestmark = pytest.mark.gpu


WSNET = Path(__file__).resolve().parent.parent / "diffractiveopticalelement_amd" / "data" / "wsnet_c5_1024.npy"


@pytest.mark.skipif(not WSNET.exists(), reason="trained parameters not shipped")
def test_c5_network_gain_on_the_fixed_evaluation_seed():
    """On spot arrays (the family where plain WGS reaches 1 %), plain needs a
    median 124.5 iterations, the network's start 148.5 (it does not help),
    and with the network's inference (~5.3 ms per 1024^2 design) plain's GPU
    time over the warm start's is 0.27; each design's own converged phase
    (the bound on a network of this output resolution) needs 28.5 iterations,
    4.6x less GPU time than plain.  So the 3x target is missed at this
    configuration, by the prediction and by the inference cost."""
    ctx = nat.Context(0)
    net = nat.WsNet(ctx, 1024, 32, 8, 0)
    net.set_params(np.load(WSNET))
    out = ws.evaluate(ctx, net, 32, 1024, 500, seed=99999, thr=0.01, log=print,
                      modes=("plain", "warm"), label_res=(1024,))
    summ = ws.summarize(out, 0.01)
    sp = {m: summ[m]["families"]["spots"] for m in summ}
    print({m: (v["reach_thr"], v["median_iterations_to_thr"], v["median_gpu_ms_to_thr"],
               v.get("median_plain_over_mode_gpu_time")) for m, v in sp.items()})
    assert sp["plain"]["reach_thr"] == 8 and abs(sp["plain"]["median_iterations_to_thr"] - 124.5) <= 2
    assert sp["warm"]["reach_thr"] == 8 and abs(sp["warm"]["median_iterations_to_thr"] - 148.5) <= 2
    assert abs(sp["labels@1024"]["median_iterations_to_thr"] - 28.5) <= 2
    assert 0.18 <= sp["warm"]["median_plain_over_mode_gpu_time"] <= 0.4
    assert sp["labels@1024"]["median_plain_over_mode_gpu_time"] >= 3.0
    assert summ["warm"]["inference_ms_per_design"] > 1.0
    net.close()
    ctx.close()


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


def test_device_quadratic_start_equals_host_phasors():
    """doe_warmstart_apply_quadratic computes the quadratic target-plane start
    on the device; it equals the host form (quadratic_phasors of the pooled
    inputs through doe_warmstart_apply_phasor): same screen rmse after 3
    iterations from either start, to FP32 rounding."""
    ctx = nat.Context(0)
    n, res, count, seed = 256, 64, 8, 31
    a, _ = ws.fill_batch(ctx, count, n, seed)
    ws.start_quadratic(ctx, a, count, n, res)
    a.run(3)
    sa = np.array([a.get_screen(i) for i in range(count)])
    b, _ = ws.fill_batch(ctx, count, n, seed)
    inp = ctx.field(res, res, count, "f32")
    b.warm_inputs(res, inp.ptrs()[0])
    x, _ = inp.download()
    inp.free()
    zf = ws._phasor_field(ctx, ws.quadratic_phasors(x, n))
    b.warm_apply_phasor(res, zf.ptrs()[0])
    zf.free()
    b.run(3)
    sb = np.array([b.get_screen(i) for i in range(count)])
    assert np.all(np.isfinite(sa))
    np.testing.assert_allclose(sa, sb, rtol=1e-4, atol=1e-6)
    a.close()
    b.close()
    ctx.close()

"""screenMetrics on the device (doe_screen_metrics) against the oracle's FP64
ProjectorSystem#screenMetrics (projector.js:554-614, metrics.js:10-321) on
the same screen field, target, ROI and element.

Gates: with the screen stored in FP64 every figure agrees to 1e-9 relative
(only the summation order differs); with the screen stored in FP32 (the
design path's storage) to 1e-3 absolute for the dimensionless figures and
1e-4 relative for the powers, the screenMetrics bar.
"""
import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
import oracle_js as oj

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]

CFGS = [
    {"n": 128, "pitch": 8e-6, "lambda": 532e-9, "distance": 0.03, "mode": "image-plane",
     "algorithm": "wgs", "iterations": 8, "seed": 3, "target": {"kind": "text", "text": "DOE"}},
    {"n": 128, "pitch": 8e-6, "lambda": 532e-9, "mode": "far-field", "algorithm": "wgs",
     "iterations": 12, "seed": 5, "target": {"kind": "disc"}},
    {"n": 128, "pitch": 8e-6, "lambda": 532e-9, "mode": "far-field", "algorithm": "wgs",
     "iterations": 12, "seed": 9, "target": {"kind": "grid", "cols": 4, "rows": 4}},
    {"n": 64, "pitch": 6e-6, "lambda": 633e-9, "mode": "far-field", "algorithm": "gs",
     "iterations": 5, "seed": 1, "target": {"kind": "dots", "cols": 3, "rows": 5}},
]
KEYS = {"efficiency": "efficiency", "fluxFraction": "flux_fraction", "uniformity": "uniformity",
        "spotUniformity": "spot_uniformity", "flatness": "flatness", "contrast": "contrast",
        "rmse": "rmse", "fidelity": "fidelity", "correlation": "correlation", "ssim": "ssim",
        "snr": "snr", "zeroOrder": "zero_order", "rmsRadius": "rms_radius",
        "meanIntensity": "mean_intensity", "peakIntensity": "peak_intensity",
        "powerInSignal": "power_in_signal", "powerAtScreen": "power_at_screen",
        "powerIncident": "power_incident"}
POWERS = {"rmsRadius", "meanIntensity", "peakIntensity", "powerInSignal", "powerAtScreen",
          "powerIncident", "snr"}


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def device_metrics(ctx, ra, r, cfg, dtype):
    n = cfg["n"]
    sh = (n, n)
    scr = ctx.upload(ra["screen_re"].reshape(sh), ra["screen_im"].reshape(sh), dtype)
    tgt = ctx.upload(ra["target"].reshape(sh), ra["roi"].reshape(sh), "f64")
    ill, ph = ra["illumination"].reshape(sh), ra["phase"].reshape(sh)
    el = ctx.upload(ill * np.cos(ph), ill * np.sin(ph), "f64")
    t = cfg["target"]
    grid = (t.get("cols", 8), t.get("rows", 8)) if t["kind"] in ("dots", "grid") else None
    sdx = r["screen_dx"]
    m = ctx.screen_metrics(scr, tgt, el, sdx, sdx, cfg["pitch"], cfg["pitch"], grid)
    for f in (scr, tgt, el):
        f.free()
    return m


@pytest.mark.parametrize("i", range(len(CFGS)))
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_screen_metrics_match_oracle(ctx, i, dtype):
    cfg = CFGS[i]
    print(f"screenMetrics {cfg['mode']} {cfg['target']['kind']} n={cfg['n']} {dtype}", flush=True)
    ra, r = oj.run("projector", {"cfg": cfg})
    ref = r["metrics"]
    got = device_metrics(ctx, ra, r, cfg, dtype)
    for k, g in KEYS.items():
        a, b = ref[k], got[g]
        if a is None:
            assert b is None, k
            continue
        if dtype == "f64":
            assert abs(a - b) <= 1e-9 * max(1.0, abs(a)), (k, a, b)
        elif k in POWERS:
            assert abs(a - b) <= 1e-4 * abs(a) + 1e-300, (k, a, b)
        else:
            assert abs(a - b) <= 1e-3, (k, a, b)


def test_screen_metrics_closed_forms(ctx):
    """A flat screen equal to the target: rmse 0, correlation 1 (fidelity 1),
    SSIM 1, uniformity/flatness 1, flux fraction = share inside the ROI; a
    flat element has zero-order fraction 1."""
    n = 64
    tgt = np.zeros((n, n))
    tgt[16:48, 16:48] = 2.0
    roi = np.zeros((n, n))
    roi[12:52, 12:52] = 1.0
    s = ctx.upload(np.sqrt(tgt), np.zeros((n, n)), "f64")
    t = ctx.upload(tgt, roi, "f64")
    e = ctx.upload(np.ones((n, n)), np.zeros((n, n)), "f64")
    m = ctx.screen_metrics(s, t, e, 1.0, 1.0, 1.0, 1.0)
    assert m["rmse"] == pytest.approx(0, abs=1e-15)
    assert m["correlation"] == pytest.approx(1, abs=1e-12)
    assert m["ssim"] == pytest.approx(1, abs=1e-12)
    assert m["uniformity"] == pytest.approx(1) and m["flatness"] == pytest.approx(1)
    assert m["flux_fraction"] == pytest.approx(1)
    assert m["zero_order"] == pytest.approx(1)
    assert m["efficiency"] == pytest.approx(2.0 * 32 * 32 / (n * n))
    assert m["spot_uniformity"] is None
    for f in (s, t, e):
        f.free()


def test_screen_metrics_argument_errors(ctx):
    s = ctx.field(32, 32, 1, "f32")
    t32 = ctx.field(32, 32, 1, "f32")
    t64 = ctx.field(16, 32, 1, "f64")
    e = ctx.field(32, 32, 1, "f32")
    with pytest.raises(nat.DoeError):
        ctx.screen_metrics(s, t32, e, 1, 1, 1, 1)  # target not FP64
    with pytest.raises(nat.DoeError):
        ctx.screen_metrics(s, t64, e, 1, 1, 1, 1)  # grids differ
    for f in (s, t32, t64, e):
        f.free()

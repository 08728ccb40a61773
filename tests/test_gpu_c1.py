"""The small reference configuration end to end on the device:
ProjectorSystem({n: 256, pitch: 8 um, lambda: 532 nm, distance: 30 mm,
mode: 'image-plane', algorithm: 'wgs', iterations: 100, seed: 0,
target: text 'DOE'}) -> design() -> simulate() -> screenMetrics()
(ref: src/projector.js:345-498, 554-614; src/doe.js:455-546, 327-437).

The oracle runs the same configuration in FP64.  The device gets the inputs
designDOE builds (illumination, power-matched target amplitude, ROI) and the
reference's xorshift start for seed 0 (state 1), runs the WGS loop, forms the
element illum * e^{i phase}, propagates it by z without band limit (the
simulate() default) and reduces the screen with doe_screen_metrics.

Gates: the first 10 iterates' error and efficiency within 1e-4 (FP32) and
1e-9 (FP64 check mode); the final screenMetrics within 1e-3 absolute for
the dimensionless figures, 1e-3 relative for the powers.
"""
import time

import numpy as np
import pytest

import oracle_js as oj
from diffractiveopticalelement_amd import native as nat

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]

C1 = {"n": 256, "pitch": 8e-6, "lambda": 532e-9, "distance": 0.03, "mode": "image-plane",
      "algorithm": "wgs", "iterations": 100, "seed": 0, "target": {"kind": "text", "text": "DOE"}}
DIMLESS = [("rmse", "rmse"), ("efficiency", "efficiency"), ("correlation", "correlation"),
           ("ssim", "ssim"), ("fluxFraction", "flux_fraction"), ("zeroOrder", "zero_order"),
           ("uniformity", "uniformity"), ("flatness", "flatness"), ("contrast", "contrast"),
           ("fidelity", "fidelity")]
POWERS = [("powerInSignal", "power_in_signal"), ("powerAtScreen", "power_at_screen"),
          ("powerIncident", "power_incident"), ("meanIntensity", "mean_intensity"),
          ("peakIntensity", "peak_intensity"), ("rmsRadius", "rms_radius")]


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle_c1():
    print("oracle: C1 projector design (FP64, node)", flush=True)
    ra, r = oj.run("projector", {"cfg": C1})
    phi0, _ = oj.run("random_phase", {"nx": 256, "ny": 256, "seed": 0})
    return ra, r, phi0["phase"]


def device_c1(ctx, ra, r, phi0, dtype):
    n, dx = C1["n"], C1["pitch"]
    plan = nat.GsPlan(mode=nat.MODE_IMAGE_PLANE, algorithm=nat.ALG_WGS, iterations=C1["iterations"],
                      nx=n, ny=n, dx=dx, dy=dx, lambda_=C1["lambda"], z=r["z"], beta=0.9, feedback=0.0,
                      escape=round(C1["iterations"] * 0.2), dtype=nat.DTYPES[dtype])
    ctx.sync()
    t0 = time.perf_counter()
    g = ctx.gs_run(plan, ra["illumination"], ra["target_amp"], ra["roi"], phi0)
    t1 = time.perf_counter()
    sh = (n, n)
    ill = ra["illumination"].reshape(sh)
    ph = g["phase"].reshape(sh)
    el = ctx.upload(ill * np.cos(ph), ill * np.sin(ph), dtype)
    scr = ctx.field(n, n, 1, dtype)
    ctx.asm(el, scr, dx, dx, C1["lambda"], r["z"], band_limit=False)  # simulate() default
    tgt = ctx.upload(ra["target"].reshape(sh), ra["roi"].reshape(sh), "f64")
    m = ctx.screen_metrics(scr, tgt, el, dx, dx, dx, dx, None)
    t2 = time.perf_counter()
    for f in (el, scr, tgt):
        f.free()
    return g, m, (t1 - t0) * 1e3, (t2 - t1) * 1e3


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_c1_design_simulate_metrics(ctx, oracle_c1, dtype):
    ra, r, phi0 = oracle_c1
    print(f"C1 device design {dtype}", flush=True)
    device_c1(ctx, ra, r, phi0, dtype)  # first call builds plans and tables
    g, m, design_ms, sim_ms = device_c1(ctx, ra, r, phi0, dtype)
    err_ref = np.array(r["error"])
    eff_ref = np.array(r["efficiency"])
    d10 = max(np.max(np.abs(g["error"][:10] - err_ref[:10])), np.max(np.abs(g["efficiency"][:10] - eff_ref[:10])))
    d100 = max(np.max(np.abs(g["error"] - err_ref)), np.max(np.abs(g["efficiency"] - eff_ref)))
    print(f"C1 {dtype}: design {design_ms:.2f} ms, simulate+metrics {sim_ms:.2f} ms (oracle "
          f"{r['design_ms']:.0f} / {r['simulate_metrics_ms']:.0f} ms); iterates: first 10 {d10:.2e}, "
          f"all 100 {d100:.2e}; rmse {m['rmse']:.5f} vs {r['metrics']['rmse']:.5f}", flush=True)
    tol_it = 1e-9 if dtype == "f64" else 1e-4
    assert d10 < tol_it
    # image-plane WGS is well conditioned: the whole history stays close
    assert d100 < (1e-8 if dtype == "f64" else 1e-4)
    mr = r["metrics"]
    for ko, kd in DIMLESS:
        assert abs(m[kd] - mr[ko]) < 1e-3, (ko, m[kd], mr[ko])
    for ko, kd in POWERS:
        assert abs(m[kd] - mr[ko]) <= 1e-3 * abs(mr[ko]), (ko, m[kd], mr[ko])
    if dtype == "f64":
        for ko, kd in DIMLESS + POWERS:
            assert abs(m[kd] - mr[ko]) <= 1e-9 * max(1.0, abs(mr[ko])), (ko, m[kd], mr[ko])

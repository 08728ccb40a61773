"""Multi-plane GS (doe_gs_multiplane_run) and the batched multi-distance
helpers (centred resample, slice copy, batched screen metrics) on the GPU,
against the oracle (gerchbergSaxtonMultiPlane, doe.js:557-626; focusSweep,
propagate.js:427-434), which tests/test_oracle_cpu.py pins to the
image-plane GS loop and the angular-spectrum semigroup.

Gates: per-plane errors of the first 10 iterations within 1e-9 (FP64) and
1e-4 (FP32, the GS iterate bar); final element phase within 1e-7 (FP64) as
unit phasors.
"""
import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
import oracle_js as oj
from test_oracle_cpu import disc, tophat

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]
LAM = 532e-9


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def case(n, dx):
    ill = tophat(n, dx, 0.45 * n * dx, 2 * dx)
    p = np.sum(ill ** 2)
    ta = np.sqrt(disc(n, 0.22 * n))
    tb = np.zeros((n, n))
    tb[n // 4:3 * n // 4, n // 3:2 * n // 3] = 1.0
    tc = np.sqrt(disc(n, 0.12 * n))
    ts = [t * np.sqrt(p / np.sum(t ** 2)) for t in (ta, tb, tc)]
    rois = [disc(n, 0.3 * n), None, disc(n, 0.2 * n)]
    planes = [(0.012, ts[0], rois[0], 1.0), (0.025, ts[1], rois[1], 0.5), (0.04, ts[2], rois[2], 2.0)]
    return ill, planes


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_multiplane_matches_oracle(ctx, n, dtype):
    dx, K = 8e-6, 10
    print(f"multiplane n={n} {dtype}", flush=True)
    ill, planes = case(n, dx)
    inputs = {"illum": ill}
    spec = []
    for q, (z, t, roi, w) in enumerate(planes):
        inputs[f"target{q}"] = t
        inputs[f"roi{q}"] = roi
        spec.append({"z": z, "weight": w})
    out, ref = oj.run("multiplane", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "iterations": K,
                                     "planes": spec}, inputs)
    g = ctx.gs_multiplane(ill, planes, n, n, dx, dx, LAM, K, dtype)
    tol = 1e-9 if dtype == "f64" else 1e-4
    pe = np.array(ref["perPlaneError"])
    assert np.max(np.abs(g["per_plane"] - pe)) < tol, np.max(np.abs(g["per_plane"] - pe))
    assert np.max(np.abs(g["error"] - np.array(ref["error"]))) < tol
    assert np.allclose(g["efficiency"], 1 - g["error"], rtol=0, atol=1e-15)
    if dtype == "f64":
        d = np.abs(np.exp(1j * g["phase"]) - np.exp(1j * out["phase"]))
        assert np.max(d) < 1e-7, np.max(d)


def test_multiplane_initial_phase_and_errors(ctx):
    n, dx = 64, 8e-6
    ill, planes = case(n, dx)
    phi0 = np.random.default_rng(3).uniform(-np.pi, np.pi, (n, n))
    inputs = {"illum": ill, "phase0": phi0}
    spec = []
    for q, (z, t, roi, w) in enumerate(planes[:2]):
        inputs[f"target{q}"] = t
        inputs[f"roi{q}"] = roi
        spec.append({"z": z, "weight": w})
    _, ref = oj.run("multiplane", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "iterations": 6,
                                   "planes": spec}, inputs)
    g = ctx.gs_multiplane(ill, planes[:2], n, n, dx, dx, LAM, 6, "f64", init_phase=phi0)
    assert np.max(np.abs(g["per_plane"] - np.array(ref["perPlaneError"]))) < 1e-9
    with pytest.raises(nat.DoeError):
        ctx.gs_multiplane(ill, [], n, n, dx, dx, LAM, 3)
    with pytest.raises(nat.DoeError):
        ctx.gs_multiplane(ill[:32, :32], planes, 32, 32, dx, dx, LAM, 3)  # sides below 64


def test_focus_sweep_batch_matches_oracle(ctx):
    """Broadcast + one batched ASM + slice copies = the oracle's per-z loop."""
    n, dx = 64, 8e-6
    rng = np.random.default_rng(4)
    re, im = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    out, r = oj.run("focus_sweep", {"nx": n, "ny": n, "dx": dx, "lambda": LAM, "z0": 0.005, "z1": 0.035,
                                    "steps": 5, "bandLimit": True}, {"re": re.ravel(), "im": im.ravel()})
    ref = (out["re"] + 1j * out["im"]).reshape(6, n, n)
    for dtype, tol in (("f64", 1e-12), ("f32", 1e-5)):
        src = ctx.upload(re, im, dtype)
        batch = ctx.field(n, n, 6, dtype)
        ctx.resample_centred(src, batch)
        ctx.asm(batch, batch, dx, dx, LAM, np.array(r["z"]), band_limit=True)
        one = ctx.field(n, n, 1, dtype)
        for s in range(6):
            ctx.copy_slices(batch, s, one, 0, 1)
            a, b = one.download()
            got = (a + 1j * b)[0]
            assert np.linalg.norm(got - ref[s]) / np.linalg.norm(ref[s]) < tol, (dtype, s)
        for f in (src, batch, one):
            f.free()


def test_resample_centred_pad_crop_broadcast(ctx):
    """Same-pitch resample (field.js:443-465): centred zero pad, centred crop,
    batch broadcast; argument errors."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((16, 32)) + 1j * rng.standard_normal((16, 32))
    src = ctx.upload(a.real, a.imag, "f64")
    big = ctx.field(64, 32, 3, "f64")
    ctx.resample_centred(src, big)
    br, bi = big.download()
    ref = np.zeros((32, 64), complex)
    ref[8:24, 16:48] = a
    for b in range(3):
        assert np.array_equal(br[b] + 1j * bi[b], ref)
    small = ctx.field(8, 8, 3, "f64")
    ctx.resample_centred(big, small)
    sr, si = small.download()
    assert np.array_equal(sr[1] + 1j * si[1], ref[12:20, 28:36])
    with pytest.raises(nat.DoeError):
        ctx.resample_centred(big, ctx.field(8, 8, 2, "f64"))  # batch 3 -> 2
    with pytest.raises(nat.DoeError):
        ctx.resample_centred(src, ctx.field(64, 32, 1, "f32"))  # dtypes differ
    with pytest.raises(nat.DoeError):
        ctx.copy_slices(big, 2, small, 2, 2)  # past the end
    for f in (src, big, small):
        f.free()


def test_screen_metrics_batch_equals_single(ctx):
    n = 64
    rng = np.random.default_rng(6)
    scr = ctx.upload(rng.standard_normal((3, n, n)), rng.standard_normal((3, n, n)), "f32")
    tg = np.zeros((n, n))
    tg[20:44, 20:44] = 1.0
    t = ctx.upload(tg, (tg > 0).astype(float), "f64")
    e = ctx.upload(np.ones((n, n)), np.zeros((n, n)), "f64")
    ms = ctx.screen_metrics_batch(scr, t, e, 1e-5, 1e-5, 8e-6, 8e-6, grid=(4, 4))
    one = ctx.field(n, n, 1, "f32")
    for b in range(3):
        ctx.copy_slices(scr, b, one, 0, 1)
        m1 = ctx.screen_metrics(one, t, e, 1e-5, 1e-5, 8e-6, 8e-6, grid=(4, 4))
        assert ms[b] == m1
    for f in (scr, t, e, one):
        f.free()

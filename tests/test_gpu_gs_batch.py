"""Batched far-field GS/WGS over designs of uneven sizes (the persistent
batch engine, `doe_gs_batch_*`).

Parity: each design of a batch is checked against the FP64 oracle's
fourierCgh / WGS loop directly (first ten iterations' error and efficiency
within 1e-4, the FP32 gate), and must reproduce the single-design far-field
loop (`doe_gs_run`, FP32) to rounding of the same operations (bitwise in
practice), over longer runs too.  The graph-replayed loop is checked against
the eagerly launched one on every replay.
"""
import numpy as np
import pytest

import oracle_js as oj
from diffractiveopticalelement_amd import native as nat
from test_gpu_gs import phase_err, problem

pytestmark = pytest.mark.gpu

LAM = 532e-9


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def single(ctx, n, alg, iters, ill, t, roi, phi0):
    plan = nat.GsPlan(mode=nat.MODE_FAR_FIELD, algorithm=alg, iterations=iters, nx=n, ny=n,
                      dx=8e-6, dy=8e-6, lambda_=LAM, z=0.03, beta=0.9, feedback=0, escape=0,
                      dtype=nat.DTYPES["f32"])
    return ctx.gs_run(plan, ill, t, roi, phi0)


def close(a, b, tol):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) <= tol


@pytest.mark.parametrize("alg", [nat.ALG_GS, nat.ALG_WGS])
@pytest.mark.parametrize("with_roi", [False, True])
def test_batch_matches_single_designs(ctx, alg, with_roi):
    sides = [256, 1024, 64, 512, 256, 2048, 128]
    iters = 12
    print(f"gs batch sides={sides} alg={alg} roi={with_roi}", flush=True)
    b = nat.GsBatch(ctx, sides, alg, 0.9, with_roi)
    inputs = []
    for i, n in enumerate(sides):
        ill, t, roi, phi0 = problem(n, 8e-6, 100 + i)
        roi = roi if with_roi else None
        b.set(i, ill, t, roi, phi0)
        inputs.append((ill, t, roi, phi0))
    b.run(iters, graph=True)
    for i, n in enumerate(sides):
        g = b.get(i)
        r = single(ctx, n, alg, iters, *inputs[i])
        assert close(g["error"], r["error"], 1e-6), i
        assert close(g["efficiency"], r["efficiency"], 1e-6), i
        assert close(np.exp(1j * g["phase"]), np.exp(1j * r["phase"]), 1e-4), i
        assert close(g["amplitude"], r["amplitude"], 1e-5 * np.max(r["amplitude"])), i
        if alg == nat.ALG_WGS:
            assert abs(g["rmse"] - r["rmse"]) <= 1e-7, i
            assert close(g["weights"], r["weights"], 1e-4 * np.max(r["weights"])), i
        else:
            assert np.isnan(g["rmse"])
    b.close()


@pytest.mark.parametrize("alg", ["gs", "wgs"])
def test_batch_designs_against_the_oracle(ctx, alg):
    # a mixed batch (sides 64, 128, 64, 256) against the oracle's far-field
    # loop on each design's inputs (doe.js:455-546, 639-730)
    sides = [64, 128, 64, 256]
    iters = 10
    print(f"gs batch vs oracle sides={sides} alg={alg}", flush=True)
    b = nat.GsBatch(ctx, sides, nat.ALG_WGS if alg == "wgs" else nat.ALG_GS, 0.9, True)
    inputs = []
    for i, n in enumerate(sides):
        ill, t, roi, phi0 = problem(n, 8e-6, 300 + i)
        b.set(i, ill, t, roi, phi0)
        inputs.append((ill, t, roi, phi0))
    b.run(iters, graph=True)
    for i, n in enumerate(sides):
        ill, t, roi, phi0 = inputs[i]
        args = {"mode": "far-field", "algorithm": alg, "nx": n, "ny": n, "dx": 8e-6, "dy": 8e-6,
                "lambda": LAM, "z": 0.03, "iterations": iters, "beta": 0.9}
        _, ref = oj.run("gs", args, {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
        g = b.get(i)
        assert close(g["error"], ref["error"], 1e-4), (i, n)
        assert close(g["efficiency"], ref["efficiency"], 1e-4), (i, n)
    b.close()


@pytest.mark.parametrize("alg", [nat.ALG_GS, nat.ALG_WGS])
def test_merged_launches_match_per_class_launches(ctx, alg):
    # one launch per pass for the whole batch (a table of size classes, the
    # default) gives the per-class launches' designs: the same line
    # transforms, other workgroup tilings (the FP64 sums fold in another order)
    sides = [256, 1024, 64, 512, 256, 2048, 128, 64]
    iters = 12
    print(f"gs batch merged vs per-class sides={sides} alg={alg}", flush=True)
    b = nat.GsBatch(ctx, sides, alg, 0.9, True)
    try:
        ills = []
        for i, n in enumerate(sides):
            ill, t, roi, phi0 = problem(n, 8e-6, 500 + i)
            b.set(i, ill, t, roi, phi0)
            ills.append(ill.ravel())
        b.run(iters, graph=True, per_class=True)
        ref = [b.get(i) for i in range(len(sides))]
        for graph in (False, True):
            b.run(iters, graph=graph)
            for i in range(len(sides)):
                g = b.get(i)
                assert close(g["error"], ref[i]["error"], 1e-6), (graph, i)
                assert close(g["efficiency"], ref[i]["efficiency"], 1e-6), (graph, i)
                # the element phase where the illumination carries it (outside
                # the aperture it multiplies zero and is not determined)
                assert phase_err(g["phase"], ref[i]["phase"], ills[i]) <= 1e-5, (graph, i)
                if alg == nat.ALG_WGS:
                    # weights compound per-point ratios (t/a)^0.9, clamped to
                    # [1/4, 4]: compared in relative L2 (a point whose measured
                    # amplitude is near zero may sit on either side of a clamp)
                    dw = np.linalg.norm(g["weights"] - ref[i]["weights"])
                    assert dw <= 1e-4 * np.linalg.norm(ref[i]["weights"]), (graph, i, dw)
    finally:
        b.close()


def test_graph_replays_match_eager(ctx):
    # the loop captured once and replayed several times, with eager runs and
    # host synchronisation in between, gives the eager result every time
    sides = [512, 256, 1024, 256]
    b = nat.GsBatch(ctx, sides, nat.ALG_WGS, 0.9)
    for i, n in enumerate(sides):
        ill, t, _, phi0 = problem(n, 8e-6, 7 + i)
        b.set(i, ill, t, None, phi0)
    b.run(10, graph=False)
    ref = [b.get(i) for i in range(len(sides))]
    for rep in range(3):
        b.run(10, graph=True)
        for i in range(len(sides)):
            g = b.get(i)
            assert np.array_equal(g["error"], ref[i]["error"]), (rep, i)
            assert np.array_equal(g["efficiency"], ref[i]["efficiency"]), (rep, i)
            assert np.array_equal(g["phase"], ref[i]["phase"]), (rep, i)
            assert np.array_equal(g["weights"], ref[i]["weights"]), (rep, i)
        b.run(4, graph=False)  # eager work of a different length between replays
    b.close()


def splitmix_phase(n, seed, index):
    """numpy restatement of the batch's counter-based FP64 initial phase."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    k = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)
             + np.uint64(index) * np.uint64(0x8CB92BA72F3D8DD7) + k) & M
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return ((u * 2 - 1) * np.pi).astype(np.float32)


def test_random_init_matches_restatement(ctx):
    sides = [64, 128, 64]
    b = nat.GsBatch(ctx, sides, nat.ALG_GS)
    for i, n in enumerate(sides):
        ill, t, _, _ = problem(n, 8e-6, i)
        b.set(i, ill, t)
    b.random_init(42)
    b.run(1, graph=False)
    # one GS iteration from phase phi0 equals the single-design loop started
    # from the restated phase (so the device phase equals the restatement)
    for i, n in enumerate(sides):
        ill, t, _, _ = problem(n, 8e-6, i)
        phi0 = splitmix_phase(n * n, 42, i).astype(np.float64).reshape(n, n)
        r = single(ctx, n, nat.ALG_GS, 1, ill, t, None, phi0)
        g = b.get(i)
        assert close(g["error"], r["error"], 1e-7), i
        assert close(np.exp(1j * g["phase"]), np.exp(1j * r["phase"]), 1e-5), i
    # statistics of the stream: uniform on [-pi, pi)
    ph = splitmix_phase(1 << 20, 42, 0).astype(np.float64)
    assert abs(np.mean(ph)) < 0.01 and abs(np.var(ph) - np.pi ** 2 / 3) < 0.01
    b.close()


def test_batch_argument_errors(ctx):
    with pytest.raises(nat.DoeError, match="powers of two"):
        nat.GsBatch(ctx, [256, 300], nat.ALG_GS)
    with pytest.raises(nat.DoeError, match="powers of two"):
        nat.GsBatch(ctx, [8192], nat.ALG_GS)
    b = nat.GsBatch(ctx, [64], nat.ALG_WGS, 0.9, with_roi=True)
    ill, t, roi, _ = problem(64, 8e-6, 0)
    with pytest.raises(nat.DoeError, match="needs an ROI"):
        b.set(0, ill, t)
    with pytest.raises(nat.DoeError, match="out of range"):
        b.set(1, ill, t, roi)
    b.close()


def test_screen_rmse_history_matches_numpy(ctx):
    # iteration 0 measures the far field of illum * exp(i phi0): judge it as
    # screenMetrics does (target power-matched to the ROI power)
    sides = [128, 256]
    b = nat.GsBatch(ctx, sides, nat.ALG_WGS, 0.9, with_roi=True)
    probs = []
    for i, n in enumerate(sides):
        ill, t, roi, phi0 = problem(n, 8e-6, 40 + i, soft_roi=False)
        b.set(i, ill, t, roi, phi0)
        probs.append((ill, t, roi, phi0))
    b.run(3, graph=False)
    for i, n in enumerate(sides):
        ill, t, roi, phi0 = probs[i]
        x = ill * np.exp(1j * phi0)
        Y = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x))) / n
        irr = np.abs(Y) ** 2
        T = t ** 2
        s = np.sum(irr * roi) / np.sum(T)
        m = roi > 0
        ref = np.sqrt(np.sum((irr[m] - s * T[m]) ** 2) / np.sum((s * T[m]) ** 2))
        scr = b.get_screen(i)
        assert abs(scr[0] - ref) < 1e-5 * max(ref, 1.0), (scr[0], ref)
        assert np.all(np.isfinite(scr))
    b.close()

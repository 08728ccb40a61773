"""Pins for the oracle's target, ROI, blur and plane-to-pixel functions (CPU).

Every check is independent of the oracle's code: numpy closed forms of each
target kind, scipy.ndimage's dilation and separable filters, exact
cumulative-sum box sums, np.interp colour ramps, and the reference's own
known answers (ref test/projector.test.js:39-60, test/doe.test.js:325-339,
test/png.test.js:91-115).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle_js as oj

pytestmark = pytest.mark.skipif(not oj.node_available(), reason="node is required for the oracle")

SIZES = [(64, 48), (37, 53), (128, 128)]


def grid(nx, ny):
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    return i, j, (nx - 1) / 2, (ny - 1) / 2


def target(kind, nx, ny, opts=None, raster=None):
    arr, res = oj.run("target", {"kind": kind, "nx": nx, "ny": ny, "opts": opts or {}},
                      {"raster": raster})
    return arr["irradiance"].reshape(ny, nx), arr["roi"].reshape(ny, nx), res


def box_roi(irr, dilate=2, soft=2, threshold=0.05):
    """autoRoi from library routines: threshold, scipy square dilation,
    clipped-window box mean from integral images (exact integers)."""
    mx = max(0.0, np.nanmax(irr))
    on = irr > threshold * mx
    r = max(1, math.floor(dilate + 0.5))
    dil = ndimage.binary_dilation(on, structure=np.ones((2 * r + 1, 2 * r + 1), bool)).astype(np.float64)
    if soft <= 0:
        return dil
    s = max(1, math.floor(soft + 0.5))
    ny, nx = dil.shape
    ii = np.zeros((ny + 1, nx + 1))
    ii[1:, 1:] = dil.cumsum(0).cumsum(1)
    j, i = np.mgrid[0:ny, 0:nx]
    i0, i1 = np.maximum(0, i - s), np.minimum(nx - 1, i + s)
    j0, j1 = np.maximum(0, j - s), np.minimum(ny - 1, j + s)
    acc = ii[j1 + 1, i1 + 1] - ii[j0, i1 + 1] - ii[j1 + 1, i0] + ii[j0, i0]
    return acc / ((j1 - j0 + 1) * (i1 - i0 + 1))


# ------------------------------------------------------------ targets -----
@pytest.mark.parametrize("nx,ny", SIZES)
def test_disc_ring_zone_closed_forms(nx, ny):
    i, j, cx, cy = grid(nx, ny)
    r = np.sqrt((i - cx) ** 2 + (j - cy) ** 2)  # (i-cx)^2 + (j-cy)^2 is exact here
    m = min(nx, ny)
    irr, _, _ = target("disc", nx, ny, {"radius": 0.4})
    np.testing.assert_array_equal(irr, (r <= 0.4 * m * 0.5).astype(float))
    assert abs(irr.sum() - math.pi * (0.2 * m) ** 2) < 2 * math.pi * 0.2 * m + 4
    irr, _, _ = target("ring", nx, ny)
    np.testing.assert_array_equal(irr, ((r >= 0.25 * m * 0.5) & (r <= 0.4 * m * 0.5)).astype(float))
    irr, _, _ = target("zone", nx, ny, {"period": 9})
    arg = np.pi * r * r / 9
    ref = (np.sin(arg) > 0).astype(float)
    edge = np.abs(np.sin(arg)) < 1e-9  # sign of sin at a zero crossing is library-dependent
    np.testing.assert_array_equal(irr[~edge], ref[~edge])


@pytest.mark.parametrize("nx,ny", SIZES)
def test_dots_grid_closed_form_and_dark_cells(nx, ny):
    i, j, _, _ = grid(nx, ny)
    cols, rows, sig = 5, 4, 0.3
    px, py = nx / cols, ny / rows
    sigma = sig * min(px, py)
    fx = np.fmod((i + 0.5) / px, 1) - 0.5
    fy = np.fmod((j + 0.5) / py, 1) - 0.5
    ref = np.exp(-((fx * px) ** 2 + (fy * py) ** 2) / (2 * sigma * sigma))
    missing = [[1, 2], [4, 0], [9, 9]]
    irr, roi, res = target("dots", nx, ny, {"cols": cols, "rows": rows, "sigma": sig, "missing": missing})
    dark = np.zeros_like(ref, bool)
    for c, rr in missing:
        dark |= ((i / px).astype(int) == c) & ((j / py).astype(int) == rr)
    assert dark.any()
    np.testing.assert_allclose(irr[~dark], ref[~dark], rtol=1e-14, atol=0)
    assert np.all(irr[dark] == 0)
    assert res["grid"] == {"cols": cols, "rows": rows}
    np.testing.assert_array_equal(roi, box_roi(irr))


@pytest.mark.parametrize("nx,ny", SIZES)
def test_checker_lines_closed_forms(nx, ny):
    i, j, _, _ = grid(nx, ny)
    irr, _, _ = target("checker", nx, ny, {"cells": 6})
    u = np.floor(i / nx * 6)
    v = np.floor(j / ny * 6)
    np.testing.assert_array_equal(irr, ((u + v) % 2 == 0).astype(float))
    irr, _, _ = target("lines", nx, ny, {"period": 7.5, "duty": 0.3})
    np.testing.assert_array_equal(irr, (np.fmod(i, 7.5) / 7.5 < 0.3).astype(float))


def test_render_text_and_contain_resample_known_answers():
    """ref test/projector.test.js:39-60, plus the 17x7 glyph strip of 'DOE'."""
    arr, res = oj.run("render_text", {"text": "DOE", "opts": {"scale": 6}})
    assert (res["width"], res["height"]) == (102, 42)
    d = arr["data"]
    assert set(np.unique(d)) == {0.0, 1.0} and 0 < d.sum() < d.size
    # the strip is 17 x 7 glyph cells upscaled 6x (nearest neighbour)
    img = d.reshape(42, 102)
    np.testing.assert_array_equal(img, np.kron(img[::6, ::6], np.ones((6, 6))))
    # a constant raster fills exactly the letterboxed window
    out, _ = oj.run("resample", {"sw": 8, "sh": 8, "dw": 32, "dh": 16, "opts": {"fit": "contain"}},
                    {"src": np.ones(64)})
    o = out["out"].reshape(16, 32)
    assert np.all(o[:, 8:24] == 1) and np.all(o[:, :8] == 0) and np.all(o[:, 24:] == 0)


def test_bilinear_resample_reproduces_a_linear_ramp():
    sw, sh, dw, dh = 10, 6, 47, 31
    src = (np.arange(sw)[None, :] * 2.0 + np.arange(sh)[:, None] * 0.5)
    out, _ = oj.run("resample", {"sw": sw, "sh": sh, "dw": dw, "dh": dh, "opts": {"gain": 3}},
                    {"src": src})
    o = out["out"].reshape(dh, dw)
    s = min(dw / sw, dh / sh)
    ox, oy = (dw - sw * s) / 2, (dh - sh * s) / 2
    x = (np.arange(dw) - ox) / s
    y = (np.arange(dh) - oy) / s
    inside_x = (x >= 0) & (x <= sw - 1)
    inside_y = (y >= 0) & (y <= sh - 1)
    ref = 3 * (x[None, :] * 2.0 + y[:, None] * 0.5)
    m = inside_y[:, None] & inside_x[None, :]
    np.testing.assert_allclose(o[m], ref[m], rtol=1e-13, atol=1e-13)
    outside = ~((np.floor(y) >= 0) & (np.floor(y) < sh))[:, None] | ~((np.floor(x) >= 0) & (np.floor(x) < sw))[None, :]
    assert np.all(o[outside] == 0)


def test_image_target_gamma():
    raster = np.linspace(-0.2, 1, 24).reshape(4, 6)
    irr1, _, res = target("image", 30, 20, {"width": 6, "height": 4, "name": "pic"}, raster)
    irr2, _, _ = target("image", 30, 20, {"width": 6, "height": 4, "gamma": 2.2}, raster)
    assert res["name"] == "pic"
    np.testing.assert_allclose(irr2, np.power(np.maximum(0, irr1), 2.2), rtol=1e-14, atol=0)


# ---------------------------------------------------------------- ROI -----
@pytest.mark.parametrize("dilate,soft,threshold", [(2, 2, 0.05), (0.4, 0, 0.05), (3.5, 1, 0.2), (6, 4, 0.01)])
def test_auto_roi_matches_scipy_dilation_and_box(dilate, soft, threshold):
    rng = np.random.default_rng(3)
    nx, ny = 45, 38
    irr = rng.random((ny, nx)) ** 6
    irr[5, 7] = np.nan  # skipped by the maximum, never above the threshold
    out, _ = oj.run("auto_roi", {"nx": nx, "ny": ny,
                                 "opts": {"dilate": dilate, "soft": soft, "threshold": threshold}},
                    {"irr": irr})
    np.testing.assert_array_equal(out["roi"].reshape(ny, nx), box_roi(irr, dilate, soft, threshold))


def test_text_target_roi_covers_the_glyphs():
    """ref test/projector.test.js:56-59"""
    irr, roi, res = target("text", 128, 128, {"text": "DOE"})
    assert res["name"] == "text:DOE"
    assert np.any(roi > 0.5) and np.all(roi[irr > 0.05 * irr.max()] > 0)
    np.testing.assert_array_equal(roi, box_roi(irr))


# --------------------------------------------------------------- blur -----
@pytest.mark.parametrize("sigma", [0.7, 2.0, 3.3])
def test_blur_matches_scipy_separable_nearest(sigma):
    rng = np.random.default_rng(5)
    nx, ny = 40, 29
    src = rng.random((ny, nx))
    out, _ = oj.run("blur", {"nx": nx, "ny": ny, "sigma": sigma}, {"src": src})
    r = max(1, math.ceil(3 * sigma))
    w = np.exp(-(np.arange(-r, r + 1) ** 2) / (2 * sigma * sigma))
    w /= w.sum()
    ref = ndimage.correlate1d(ndimage.correlate1d(src, w, axis=1, mode="nearest"), w, axis=0, mode="nearest")
    np.testing.assert_allclose(out["out"].reshape(ny, nx), ref, rtol=1e-13, atol=1e-15)


def test_blur_conserves_energy_known_answer():
    """ref test/doe.test.js:325-339"""
    n = 32
    src = np.zeros(n * n)
    src[16 + 16 * n] = 1
    out, _ = oj.run("blur", {"nx": n, "ny": n, "sigma": 2}, {"src": src})
    b = out["out"]
    assert abs(b.sum() - 1) < 0.02 and b[16 + 16 * n] < 1 and b[18 + 16 * n] > 0
    out, _ = oj.run("blur", {"nx": n, "ny": n, "sigma": 0}, {"src": src})
    np.testing.assert_array_equal(out["out"], src)


# ------------------------------------------------------------- pixels -----
VIRIDIS = [(0, (68, 1, 84)), (0.25, (59, 82, 139)), (0.5, (33, 145, 140)), (0.75, (94, 201, 98)),
           (1, (253, 231, 37))]


def ramp_np(stops, t):
    x = np.clip(t, 0, 1)
    pos = np.array([p for p, _ in stops])
    return np.stack([np.floor(np.interp(x, pos, [c[k] for _, c in stops]) + 0.5) for k in range(3)], -1)


def test_plane_to_rgb_linear_matches_interp():
    rng = np.random.default_rng(7)
    plane = rng.normal(size=500) * 3 + 1
    out, res = oj.run("image", {"opts": {"map": "viridis"}}, {"plane": plane})
    lo, hi = plane.min(), plane.max()
    assert (res["lo"], res["hi"]) == (lo, hi)
    ref = ramp_np(VIRIDIS, (plane - lo) / (hi - lo))
    np.testing.assert_array_equal(out["bytes"].reshape(-1, 3), ref)
    # stop positions give the stop colours exactly
    q = np.array([0, 0.25, 0.5, 0.75, 1.0])
    out, _ = oj.run("image", {"opts": {"min": 0, "max": 1}}, {"plane": q})
    np.testing.assert_array_equal(out["bytes"].reshape(-1, 3), [c for _, c in VIRIDIS])


def test_plane_to_rgb_percentile_and_gamma():
    rng = np.random.default_rng(9)
    plane = rng.random(777) ** 3
    out, res = oj.run("image", {"opts": {"map": "gray", "percentile": 0.9, "gamma": 0.5}}, {"plane": plane})
    srt = np.sort(plane)
    hi = srt[min(len(plane) - 1, int(np.floor(len(plane) * 0.9)))]
    assert (res["lo"], res["hi"]) == (srt[0], hi)
    t = np.clip((plane - srt[0]) / (hi - srt[0]), 0, 1) ** 0.5
    ref = np.floor(t * 255 + 0.5)
    got = out["bytes"].reshape(-1, 3)
    assert np.all(got[:, 0] == got[:, 1]) and np.all(got[:, 1] == got[:, 2])
    # pow is the only non-exact step: a rounding flip is at most one level
    assert np.max(np.abs(got[:, 0] - ref)) <= 1 and np.mean(got[:, 0] != ref) < 0.01


def test_plane_to_gray_known_answers():
    """ref test/png.test.js:91-115"""
    out, _ = oj.run("image", {"gray": True, "opts": {}}, {"plane": np.arange(9) / 8})
    np.testing.assert_array_equal(out["bytes"], np.floor(np.arange(9) / 8 * 255 + 0.5))
    out, _ = oj.run("image", {"gray": True, "opts": {"min": 0, "max": 4}}, {"plane": np.array([0, 0.5, 2])})
    assert abs(out["bytes"][2] - 127.5) < 1
    out, _ = oj.run("image", {"gray": True, "opts": {"invert": True}}, {"plane": np.array([0.0, 1.0, 0.25])})
    np.testing.assert_array_equal(out["bytes"], [255, 0, 191])

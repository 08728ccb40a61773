"""Target rasters, ROIs, blur and figure pixels on the device against the
oracle (makeTarget / autoRoi / resampleRaster projector.js:126-315,
blurRaster doe.js:761-794, planeToRGB / planeToGray color.js:87-160).

Gates: exact equality wherever the reference's arithmetic is exact in FP64
(disc, ring, zone, checker, lines, text and image rasters, every ROI, the
linear colour and gray mappings); 1e-14 relative where a transcendental
library call enters (dots: exp; image gamma: pow; blur: exp-built taps);
pixel levels within one step on under 1% of pixels where log10 / pow /
atan2 / hypot feed a rounding to 8 bits.
"""
import math

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
import oracle_js as oj

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]

SIZES = [(64, 64), (37, 53), (256, 192)]


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def oracle_target(kind, nx, ny, opts=None, raster=None):
    arr, res = oj.run("target", {"kind": kind, "nx": nx, "ny": ny, "opts": opts or {}}, {"raster": raster})
    return arr["irradiance"].reshape(ny, nx), arr["roi"].reshape(ny, nx), res


def text_raster(text="DOE", scale=6):
    arr, res = oj.run("render_text", {"text": text, "opts": {"scale": scale}})
    return arr["data"].reshape(res["height"], res["width"])


def device_target(ctx, specs, nx, ny, roi=(2.0, 2.0, 0.05)):
    f = ctx.target_make(specs, nx, ny)
    if roi is not None:
        ctx.target_auto_roi(f, *roi)
    re, im = f.download()
    f.free()
    return re, im


EXACT_KINDS = [("disc", {"radius": 0.35}), ("ring", {}), ("zone", {"period": 11}), ("checker", {"cells": 5}),
               ("lines", {"period": 9, "duty": 0.4})]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kind,opts", EXACT_KINDS)
def test_analytic_targets_bit_exact(ctx, nx, ny, kind, opts):
    irr, roi, _ = oracle_target(kind, nx, ny, opts)
    re, im = device_target(ctx, [ctx.target_spec(kind, **opts)], nx, ny)
    print(kind, nx, ny, flush=True)
    np.testing.assert_array_equal(re[0], irr)
    np.testing.assert_array_equal(im[0], roi)


@pytest.mark.parametrize("nx,ny", SIZES)
def test_dots_with_dark_cells(ctx, nx, ny):
    opts = {"cols": 6, "rows": 5, "sigma": 0.25, "missing": [[0, 0], [5, 4], [2, 3], [7, 1]]}
    irr, roi, _ = oracle_target("dots", nx, ny, opts)
    re, im = device_target(ctx, [ctx.target_spec("dots", **opts)], nx, ny)
    np.testing.assert_allclose(re[0], irr, rtol=1e-14, atol=0)
    assert np.all((re[0] == 0) == (irr == 0))
    thr = 0.05 * irr.max()
    near = np.abs(irr - thr) < 1e-12 * thr  # an exp ulp could move such a pixel across the threshold
    assert not near.any()
    np.testing.assert_array_equal(im[0], roi)


@pytest.mark.parametrize("nx,ny", [(256, 256), (128, 200), (61, 47)])
def test_text_target_bit_exact(ctx, nx, ny):
    r = text_raster()
    irr, roi, _ = oracle_target("text", nx, ny, {"text": "DOE"})
    re, im = device_target(ctx, [ctx.target_spec("raster", raster=r)], nx, ny)
    np.testing.assert_array_equal(re[0], irr)
    np.testing.assert_array_equal(im[0], roi)


def test_image_target_and_resample_options(ctx):
    rng = np.random.default_rng(11)
    img = rng.random((23, 31)) * 1.2 - 0.1
    irr1, roi1, _ = oracle_target("image", 96, 80, {"width": 31, "height": 23}, img)
    irr2, _, _ = oracle_target("image", 96, 80, {"width": 31, "height": 23, "gamma": 2.2}, img)
    re, im = device_target(ctx, [ctx.target_spec("raster", raster=img),
                                 ctx.target_spec("raster", raster=img, gamma=2.2)], 96, 80)
    np.testing.assert_array_equal(re[0], irr1)
    np.testing.assert_array_equal(im[0], roi1)
    np.testing.assert_allclose(re[1], irr2, rtol=1e-14, atol=0)
    # resampleRaster with fit 'cover', background and gain
    for opts in ({"fit": "cover", "gain": 2.5}, {"fit": "contain", "background": 0.3}):
        arr, _ = oj.run("resample", {"sw": 31, "sh": 23, "dw": 70, "dh": 90, "opts": opts}, {"src": img})
        spec = ctx.target_spec("raster", raster=img, cover=int(opts["fit"] == "cover"),
                               gain=opts.get("gain", 1.0), background=opts.get("background", 0.0))
        got, _ = device_target(ctx, [spec], 70, 90, roi=None)
        np.testing.assert_array_equal(got[0], arr["out"].reshape(90, 70))


@pytest.mark.parametrize("roi", [(2.0, 2.0, 0.05), (3.0, 0.0, 0.2), (12.0, 9.0, 0.05)])
def test_make_roi_equals_make_then_auto_roi(ctx, roi):
    nx, ny = 200, 136
    kinds = [("disc", {}), ("dots", {"cols": 5, "rows": 4, "missing": [[1, 1]]}),
             ("raster", {"raster": text_raster("XY")}), ("zone", {"period": 7}), ("ring", {})]
    specs = [ctx.target_spec(k, **o) for k, o in kinds]
    re1, im1 = device_target(ctx, specs, nx, ny, roi=roi)
    f = ctx.target_make_roi(specs, nx, ny, None, *roi)
    re2, im2 = f.download()
    f.free()
    np.testing.assert_array_equal(re2, re1)
    np.testing.assert_array_equal(im2, im1)


def test_mixed_batch_equals_single_calls(ctx):
    nx, ny = 128, 96
    kinds = [("disc", {}), ("dots", {"cols": 4, "rows": 3}), ("raster", {"raster": text_raster("AB")}),
             ("zone", {}), ("lines", {}), ("ring", {"inner": 0.1, "outer": 0.3})]
    specs = [ctx.target_spec(k, **o) for k, o in kinds]
    re_b, im_b = device_target(ctx, specs, nx, ny, roi=(3.0, 1.0, 0.1))
    for q, s in enumerate(specs):
        re1, im1 = device_target(ctx, [s], nx, ny, roi=(3.0, 1.0, 0.1))
        np.testing.assert_array_equal(re_b[q], re1[0])
        np.testing.assert_array_equal(im_b[q], im1[0])


def test_specs_sharing_a_raster_equal_single_calls(ctx):
    """Specs of one call that pass the same host raster share one device copy;
    each still equals its own single call (different gain, fit, gamma; a
    second raster in between)."""
    nx, ny = 120, 88
    rng = np.random.default_rng(21)
    img = np.ascontiguousarray(rng.random((23, 31)))
    txt = text_raster("Q")
    specs = [ctx.target_spec("raster", raster=img), ctx.target_spec("raster", raster=txt),
             ctx.target_spec("raster", raster=img, gain=2.0, cover=1),
             ctx.target_spec("raster", raster=img, gamma=2.2, background=0.1),
             ctx.target_spec("raster", raster=txt, cover=1)]
    re_b, im_b = device_target(ctx, specs, nx, ny)
    for q, sp in enumerate(specs):
        re1, im1 = device_target(ctx, [sp], nx, ny)
        np.testing.assert_array_equal(re_b[q], re1[0])
        np.testing.assert_array_equal(im_b[q], im1[0])


# (8, 8): the largest halo of the tiled ROI kernel; (9, 8) and (40, 7): the separable passes
@pytest.mark.parametrize("dilate,soft,threshold", [(2, 2, 0.05), (0.3, 0, 0.05), (4.5, 3, 0.2), (8, 8, 0.05),
                                                   (9, 8, 0.05), (16, 0, 0.05), (40, 7, 0.01)])
def test_auto_roi_bit_exact_on_custom_irradiance(ctx, dilate, soft, threshold):
    rng = np.random.default_rng(13)
    nx, ny = 93, 70
    irr = rng.random((ny, nx)) ** 8
    irr[3, 4] = np.nan
    irr[10, 10] = -1.0
    arr, _ = oj.run("auto_roi", {"nx": nx, "ny": ny, "opts": {"dilate": dilate, "soft": soft,
                                                               "threshold": threshold}}, {"irr": irr})
    f = ctx.upload(irr, np.zeros_like(irr), "f64")
    ctx.target_auto_roi(f, dilate, soft, threshold)
    _, im = f.download()
    f.free()
    np.testing.assert_array_equal(im[0], arr["roi"].reshape(ny, nx))


@pytest.mark.parametrize("sigma", [0.0, 0.8, 2.0, 5.5])
def test_blur_matches_oracle(ctx, sigma):
    rng = np.random.default_rng(17)
    nx, ny = 77, 64
    src = rng.random((ny, nx))
    arr, _ = oj.run("blur", {"nx": nx, "ny": ny, "sigma": sigma}, {"src": src})
    f = ctx.upload(np.stack([src, src[::-1]]), None, "f64")
    g = ctx.field(nx, ny, 2, "f64")
    ctx.raster_blur(f, g, sigma)
    re, _ = g.download()
    ref = arr["out"].reshape(ny, nx)
    np.testing.assert_allclose(re[0], ref, rtol=1e-14, atol=1e-16)
    # in place, and with the caller's taps
    if sigma > 0:
        r = max(1, math.ceil(3 * sigma))
        w = np.exp(-(np.arange(-r, r + 1) ** 2) / (2 * sigma * sigma))
        w /= w.sum()
        ctx.raster_blur(f, f, sigma, taps=w)
        re2, _ = f.download()
        np.testing.assert_allclose(re2[0], ref, rtol=1e-13, atol=1e-15)
    f.free()
    g.free()


# ------------------------------------------------------------- pixels -----
VIRIDIS = [(0, (68, 1, 84)), (0.25, (59, 82, 139)), (0.5, (33, 145, 140)), (0.75, (94, 201, 98)),
           (1, (253, 231, 37))]
TURBO = [(0, (48, 18, 59)), (0.2, (65, 122, 246)), (0.4, (26, 219, 172)), (0.6, (175, 240, 91)),
         (0.8, (251, 126, 34)), (1, (122, 4, 3))]


def device_image(ctx, plane, **kw):
    f = ctx.upload(plane.reshape(1, *plane.shape), None, "f64")
    out, rng = ctx.field_image(f, **kw)
    f.free()
    return out, rng


@pytest.mark.parametrize("shape", [(64, 64), (37, 101)])
def test_rgb_and_gray_linear_bit_exact(ctx, shape):
    rng = np.random.default_rng(19)
    plane = rng.normal(size=shape) * 2 + 0.5
    plane[1, 2] = np.nan
    for name, stops, opts in [("viridis", VIRIDIS, {}), ("turbo", TURBO, {"min": -1.0}),
                              ("viridis", VIRIDIS, {"min": 0.0, "max": 2.0})]:
        arr, res = oj.run("image", {"opts": dict(opts, map=name)}, {"plane": plane})
        got, (lo, hi) = device_image(ctx, plane, stops=stops, vmin=opts.get("min"), vmax=opts.get("max"))
        np.testing.assert_array_equal(got, arr["bytes"].astype(np.uint8))
        assert (lo, hi) == (res["lo"], res["hi"])
    for opts in ({}, {"invert": True}, {"min": 0.0, "max": 4.0}):
        arr, res = oj.run("image", {"gray": True, "opts": opts}, {"plane": plane})
        got, _ = device_image(ctx, plane, gray=True, vmin=opts.get("min"), vmax=opts.get("max"),
                              invert=opts.get("invert", False))
        np.testing.assert_array_equal(got, arr["bytes"].astype(np.uint8))


def test_percentile_upper_end_is_the_sorted_element(ctx):
    rng = np.random.default_rng(23)
    plane = rng.random((150, 171)) ** 4
    plane[0, :5] = np.nan
    plane[1, :3] = -0.0
    for p in (0.0, 0.5, 0.99, 0.999, 1.0):
        arr, res = oj.run("image", {"opts": {"map": "viridis", "percentile": p}}, {"plane": plane})
        got, (lo, hi) = device_image(ctx, plane, stops=VIRIDIS, percentile=p)
        if res["hi"] is None:  # the sorted plane ends in NaN (JSON carries it as null)
            assert p == 1.0 and np.isnan(hi)
        else:
            assert hi == res["hi"], (p, hi, res["hi"])
        assert lo == res["lo"]
        np.testing.assert_array_equal(got, arr["bytes"].astype(np.uint8))


def test_log_and_gamma_within_one_level(ctx):
    rng = np.random.default_rng(29)
    plane = rng.random((128, 96)) ** 6
    for opts in ({"log": True}, {"log": True, "logFloor": 1e-3}, {"gamma": 0.45}):
        arr, _ = oj.run("image", {"opts": dict(opts, map="viridis")}, {"plane": plane})
        got, _ = device_image(ctx, plane, stops=VIRIDIS, log=opts.get("log", False),
                              log_floor=opts.get("logFloor", 1e-4), gamma=opts.get("gamma", 1.0))
        ref = arr["bytes"].astype(np.int32)
        d = np.abs(got.astype(np.int32) - ref)
        assert d.max() <= 1 and np.mean(d > 0) < 0.01


def test_field_quantities_and_png_rows(ctx):
    rng = np.random.default_rng(31)
    ny, nx = 48, 40
    re = rng.normal(size=(ny, nx))
    im = rng.normal(size=(ny, nx))
    f = ctx.upload(re[None], im[None], "f64")
    planes = {"intensity": re * re + im * im, "amplitude": np.hypot(re, im), "phase": np.arctan2(im, re),
              "im": im}
    for q, plane in planes.items():
        arr, _ = oj.run("image", {"opts": {"map": "viridis"}}, {"plane": plane})
        got, _ = ctx.field_image(f, quantity=q, stops=VIRIDIS)
        ref = arr["bytes"].astype(np.int32)
        d = np.abs(got.astype(np.int32) - ref)
        if q in ("intensity", "im"):
            assert d.max() == 0, q
        else:
            assert d.max() <= 1 and np.mean(d > 0) < 0.01, q
    plain, _ = ctx.field_image(f, quantity="intensity", gray=True)
    rows, _ = ctx.field_image(f, quantity="intensity", gray=True, png_rows=True)
    rows = rows.reshape(ny, nx + 1)
    assert np.all(rows[:, 0] == 0)
    np.testing.assert_array_equal(rows[:, 1:].ravel(), plain)
    f.free()


def test_fp32_field_image_uses_the_stored_values(ctx):
    rng = np.random.default_rng(37)
    re = rng.normal(size=(32, 32)).astype(np.float32).astype(np.float64)
    im = rng.normal(size=(32, 32)).astype(np.float32).astype(np.float64)
    f = ctx.upload(re[None], im[None], "f32")
    arr, _ = oj.run("image", {"opts": {"map": "viridis"}}, {"plane": re * re + im * im})
    got, _ = ctx.field_image(f, quantity="intensity", stops=VIRIDIS)
    np.testing.assert_array_equal(got, arr["bytes"].astype(np.uint8))
    f.free()


def test_argument_errors(ctx):
    f = ctx.field(8, 8, 1, "f32")
    with pytest.raises(nat.DoeError, match="FP64"):
        ctx.target_make([ctx.target_spec("disc")], 8, 8, out=f)
    g = ctx.field(8, 8, 1, "f64")
    with pytest.raises(nat.DoeError, match="cols and rows"):
        ctx.target_make([ctx.target_spec("dots", cols=0)], 8, 8, out=g)
    with pytest.raises(nat.DoeError, match="raster"):
        ctx.target_make([ctx.target_spec("raster")], 8, 8, out=g)
    with pytest.raises(nat.DoeError, match="index"):
        ctx.field_image(g, index=1)
    f.free()
    g.free()

"""ComplexField#resample and #sampleImage on the device against the FP64
oracle (ref: src/field.js:350-400, 443-495): bit for bit on FP64 fields
(same operations in the same order), 1e-6 relative on FP32 fields."""
import numpy as np
import pytest

import oracle_js as oj
from bitcmp import assert_same_bits
from diffractiveopticalelement_amd import native as nat

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n,n2,dx,dx2", [(64, 128, 4e-6, 2.5e-6), (64, 32, 4e-6, 6e-6), (32, 32, 8e-6, 7e-6),
                                         (64, 128, 4e-6, 4e-6), (128, 64, 4e-6, 4e-6)])
def test_resample_vs_oracle(ctx, n, n2, dx, dx2):
    print(f"resample {n} -> {n2}, pitch {dx} -> {dx2}", flush=True)
    rng = np.random.default_rng(n + n2)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out, _ = oj.run("field_resample", {"nx": n, "ny": n, "dx": dx, "dy": dx, "nx2": n2, "ny2": n2,
                                       "dx2": dx2, "dy2": dx2}, {"re": u.real.ravel(), "im": u.imag.ravel()})
    for dtype in ("f64", "f32"):
        src = ctx.upload(u.real, u.imag, dtype)
        dst = ctx.field(n2, n2, 1, dtype)
        ctx.resample(src, dst, dx, dx, dx2, dx2)
        re, im = dst.download()
        if dtype == "f64":
            assert_same_bits(re, out["re"])
            assert_same_bits(im, out["im"])
        else:
            ref = out["re"] + 1j * out["im"]
            got = (re + 1j * im).ravel()
            assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)
        src.free()
        dst.free()


@pytest.mark.parametrize("nx,ny,w,h,scale,fill", [(64, 64, 64, 64, 1.0, 0.0), (128, 64, 40, 30, 1.0, 0.0),
                                                  (64, 64, 200, 150, 0.8, 0.3), (256, 128, 17, 9, 0.55, -0.5)])
def test_sample_image_vs_oracle(ctx, nx, ny, w, h, scale, fill):
    print(f"sampleImage {w}x{h} -> {nx}x{ny} scale {scale}", flush=True)
    img = np.random.default_rng(w * h).uniform(0, 1, (h, w))
    out, _ = oj.run("sample_image", {"nx": nx, "ny": ny, "width": w, "height": h, "scale": scale,
                                     "fill": fill}, {"img": img.ravel()})
    for dtype in ("f64", "f32"):
        dst = ctx.field(nx, ny, 1, dtype)
        ctx.sample_image(img, dst, scale, fill)
        re, im = dst.download()
        if dtype == "f64":
            assert_same_bits(re, out["re"])
        else:
            assert np.max(np.abs(re.ravel() - out["re"])) <= 1e-6 * np.max(np.abs(out["re"]))
        assert np.all(im == 0)
        dst.free()


def test_sampling_argument_errors(ctx):
    a = ctx.field(8, 8, 1, "f64")
    b = ctx.field(16, 16, 1, "f32")
    with pytest.raises(nat.DoeError, match="dtypes differ"):
        ctx.resample(a, b, 1e-6, 1e-6, 2e-6, 2e-6)
    c = ctx.field(16, 16, 1, "f64")
    with pytest.raises(nat.DoeError, match="positive"):
        ctx.resample(a, c, 0.0, 1e-6, 2e-6, 2e-6)
    for f in (a, b, c):
        f.free()

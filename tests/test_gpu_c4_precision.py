"""Precision sweep at 4096 x 4096 (ref: src/fft.js:148-154, 208-237;
src/propagate.js:57-133): fft2d and the band-limited angular spectrum
(z = 50 mm, 532 nm, pitch 2 um) in FP64, FP32 and FP16/BF16 storage with
FP32 arithmetic, on a field of complex Gaussian N(0, 1) entries (unit
variance) and on a smooth Gaussian beam (waist 0.2 N dx, a gentle
defocus phase).

References: the device FP64 check mode (itself checked here against numpy's
FFT, against a naive FP64 DFT on 256^2 crops of the spectrum, and, for the
angular spectrum, against the FP64 oracle at full size) and the naive DFT
crops directly.  Gates: FP64 <= 1e-12, FP32 <= 1e-5; half-width storage
rounds every pass's output to 11 (FP16) or 8 (BF16) significant bits:
<= 2e-3 and <= 1.5e-2 relative L2.  An unscaled 4096^2 spectrum of the
beam peaks at ~2.6e5, beyond FP16's largest finite value (65504): its FP16
fft2d overflows, which the test asserts, and the same beam scaled by 2^-6
stays in range.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_js as oj
from diffractiveopticalelement_amd import native as nat

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import dft_crop  # noqa: E402  (the naive-DFT oracle)

pytestmark = pytest.mark.gpu

N, DX, LAM, Z = 4096, 2e-6, 532e-9, 0.05
TOL = {"f64": 1e-12, "f32": 1e-5, "f16": 2e-3, "bf16": 1.5e-2}
# 256^2 crops of the spectrum (start row, start column; indices wrap): the
# N(0, 1) field's energy is spread over all frequencies; the beam's sits
# around DC (natural order: the corners), so its crop wraps around DC (a
# crop far from DC holds ~1e-10 of the energy, where a relative error only
# measures the rounding of the large DC values)
CROPS = {"gauss": ((1920, 1920), (0, 0), (3000, 700)),
         "beam": ((N - 128, N - 128),)}


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def field(kind):
    if kind == "gauss":
        rng = np.random.default_rng(44)
        return (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / np.sqrt(2)
    c = (np.arange(N) - N / 2) * DX
    r2 = c[None, :] ** 2 + c[:, None] ** 2
    w = 0.2 * N * DX
    return np.exp(-r2 / w ** 2) * np.exp(1j * np.pi * r2 / (LAM * 2.0))  # 2 m radius of curvature


def dev(ctx, x, dtype, op):
    f = ctx.upload(x.real, x.imag, dtype)
    op(f)
    re, im = f.download()
    f.free()
    return re[0] + 1j * im[0]


@pytest.mark.parametrize("kind", ["gauss", "beam"])
def test_c4_fft2d_precision_sweep(ctx, kind):
    x = field(kind)
    print(f"C4 fft2d {kind} 4096^2", flush=True)
    y = {dt: dev(ctx, x, dt, ctx.fft2d) for dt in ("f64", "f32", "bf16")}
    ref = np.fft.fft2(x)
    crops = [dft_crop.dft2_crop(x, ky, kx, 256) for ky, kx in CROPS[kind]]

    def crop_err(yy):
        return max(rel(yy[np.ix_((ky + np.arange(256)) % N, (kx + np.arange(256)) % N)], c)
                   for (ky, kx), c in zip(CROPS[kind], crops))

    rep = {"f64_vs_numpy": rel(y["f64"], ref), "f64_vs_dft_crops": crop_err(y["f64"])}
    for dt in ("f32", "bf16"):
        rep[f"{dt}_vs_f64"] = rel(y[dt], y["f64"])
        rep[f"{dt}_vs_dft_crops"] = crop_err(y[dt])
    if kind == "gauss":
        y16 = dev(ctx, x, "f16", ctx.fft2d)
        rep["f16_vs_f64"] = rel(y16, y["f64"])
        rep["f16_vs_dft_crops"] = crop_err(y16)
    else:
        y16 = dev(ctx, x, "f16", ctx.fft2d)
        rep["f16_peak_over_range"] = float(np.abs(ref).max())
        assert not np.all(np.isfinite(y16))  # beyond FP16's range, as stated
        s = 2.0 ** -6
        y16 = dev(ctx, x * s, "f16", ctx.fft2d) / s
        rep["f16_scaled_vs_f64"] = rel(y16, y["f64"])
    print(f"C4 fft2d {kind}:", {k: f"{v:.2e}" for k, v in rep.items()}, flush=True)
    assert rep["f64_vs_numpy"] < TOL["f64"] and rep["f64_vs_dft_crops"] < TOL["f64"]
    assert rep["f32_vs_f64"] < TOL["f32"] and rep["f32_vs_dft_crops"] < TOL["f32"]
    assert rep["bf16_vs_f64"] < TOL["bf16"] and rep["bf16_vs_dft_crops"] < TOL["bf16"]
    f16 = rep.get("f16_vs_f64", rep.get("f16_scaled_vs_f64"))
    assert f16 < TOL["f16"]


@pytest.mark.parametrize("kind", ["gauss", "beam"])
def test_c4_asm_precision_sweep(ctx, kind):
    x = field(kind)
    print(f"C4 asm {kind} 4096^2 z = 50 mm", flush=True)

    def op(f):
        ctx.asm(f, f, DX, DX, LAM, Z)  # band-limited (the reference's default)

    y = {dt: dev(ctx, x, dt, op) for dt in ("f64", "f32", "f16", "bf16")}
    rep = {f"{dt}_vs_f64": rel(y[dt], y["f64"]) for dt in ("f32", "f16", "bf16")}
    if kind == "beam":
        out, _ = oj.run("asm", {"nx": N, "ny": N, "dx": DX, "dy": DX, "lambda": LAM, "z": Z,
                                "bandLimit": True}, {"re": x.real.ravel(), "im": x.imag.ravel()},
                        timeout=1200)
        ref = (out["re"] + 1j * out["im"]).reshape(N, N)
        rep["f64_vs_oracle"] = rel(y["f64"], ref)
        rep["f32_vs_oracle"] = rel(y["f32"], ref)
    print(f"C4 asm {kind}:", {k: f"{v:.2e}" for k, v in rep.items()}, flush=True)
    for dt in ("f32", "f16", "bf16"):
        assert rep[f"{dt}_vs_f64"] < TOL[dt], dt
    if kind == "beam":
        assert rep["f64_vs_oracle"] < TOL["f64"] and rep["f32_vs_oracle"] < TOL["f32"]

"""The batched-propagation configuration at its full size: 16 complex fields
of 2048 x 2048, Gaussian amplitude (waist 0.3 N dx) times a seeded uniform
phase, pitch 4 um, 633 nm, z_b ~ U[10, 100] mm per field.  Band-limited
angular spectrum forward, then inverse (the round trip), and single-FFT
Fresnel forward (ref: src/propagate.js:57-133, 155-239).

The FP32 results are compared with the device's FP64 check mode on the same
inputs (all 16 fields, in chunks), and the FP64 check mode with the FP64
oracle itself on one field at full size.  Gates: FP32 <= 1e-5 relative L2;
FP64 vs oracle <= 1e-12 (ASM), 1e-10 (Fresnel: the reference forms
k z - pi (nx + ny) / 2 and the chirp from products near 1e5 rad, whose
last-ulp differences reach ~1e-11); Parseval of the Fresnel transform
<= 1e-6 (FP32) and 1e-12 (FP64).  The band-limited round trip is lossy by
design (the band edge removes power), so it is compared between precisions,
not with the input.
"""
import numpy as np
import pytest

import oracle_js as oj
from diffractiveopticalelement_amd import native as nat

pytestmark = pytest.mark.gpu

B, N, DX, LAM = 16, 2048, 4e-6, 633e-9
CHUNK = 4


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def c2_fields(first, count):
    """fields first..first+count-1 of the batch (field b from seed b)"""
    c = (np.arange(N) - N / 2) * DX
    w = 0.3 * N * DX
    amp = np.exp(-(c[None, :] ** 2 + c[:, None] ** 2) / w ** 2)
    out = np.empty((count, N, N), np.complex128)
    for q in range(count):
        ph = np.random.default_rng(1000 + first + q).uniform(0, 2 * np.pi, (N, N))
        out[q] = amp * np.exp(1j * ph)
    return out


Z = np.random.default_rng(0).uniform(0.010, 0.100, B)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def run(ctx, x, dtype, op):
    f = ctx.upload(x.real, x.imag, dtype)
    g = ctx.field(N, N, x.shape[0], dtype)
    op(f, g)
    re, im = g.download()
    f.free()
    g.free()
    return re + 1j * im


def test_c2_full_batch_fp32_vs_fp64(ctx):
    worst = {"asm": 0.0, "round_trip": 0.0, "fresnel": 0.0, "parseval32": 0.0, "parseval64": 0.0}
    for b0 in range(0, B, CHUNK):
        print(f"C2 chunk fields {b0}..{b0 + CHUNK - 1}", flush=True)
        x = c2_fields(b0, CHUNK)
        z = Z[b0:b0 + CHUNK]

        def asm_fwd(f, g):
            ctx.asm(f, g, DX, DX, LAM, z)

        def asm_rt(f, g):
            ctx.asm(f, g, DX, DX, LAM, z)
            ctx.asm(g, g, DX, DX, LAM, -z)

        def fres(f, g):
            ctx.fresnel(f, g, DX, DX, LAM, z)

        for key, op in (("asm", asm_fwd), ("round_trip", asm_rt), ("fresnel", fres)):
            y32 = run(ctx, x, "f32", op)
            y64 = run(ctx, x, "f64", op)
            for q in range(CHUNK):
                worst[key] = max(worst[key], rel(y32[q], y64[q]))
            if key == "fresnel":
                odx = LAM * np.abs(z) / (N * DX)
                p_in = np.sum(np.abs(x) ** 2, axis=(1, 2)) * DX * DX
                for y, k in ((y32, "parseval32"), (y64, "parseval64")):
                    p_out = np.sum(np.abs(y) ** 2, axis=(1, 2)) * odx * odx
                    worst[k] = max(worst[k], float(np.max(np.abs(p_out / p_in - 1))))
            del y32, y64
    print("C2 full batch:", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    assert worst["asm"] < 1e-5 and worst["round_trip"] < 1e-5 and worst["fresnel"] < 1e-5
    assert worst["parseval32"] < 1e-6 and worst["parseval64"] < 1e-12


def test_c2_fp64_against_oracle_one_field(ctx):
    b = int(np.argmax(Z))  # the largest distance: the narrowest band limit
    x = c2_fields(b, 1)
    z = float(Z[b])
    print(f"C2 field {b} (z = {z * 1e3:.2f} mm) vs the oracle, 2048^2", flush=True)
    args = {"nx": N, "ny": N, "dx": DX, "dy": DX, "lambda": LAM, "z": z, "bandLimit": True}
    out, _ = oj.run("asm", args, {"re": x[0].real.ravel(), "im": x[0].imag.ravel()}, timeout=900)
    ref = (out["re"] + 1j * out["im"]).reshape(N, N)
    got = run(ctx, x, "f64", lambda f, g: ctx.asm(f, g, DX, DX, LAM, z))[0]
    e_asm = rel(got, ref)
    got32 = run(ctx, x, "f32", lambda f, g: ctx.asm(f, g, DX, DX, LAM, z))[0]
    e_asm32 = rel(got32, ref)
    out, res = oj.run("fresnel", {"nx": N, "ny": N, "dx": DX, "dy": DX, "lambda": LAM, "z": z},
                      {"re": x[0].real.ravel(), "im": x[0].imag.ravel()}, timeout=900)
    ref = (out["re"] + 1j * out["im"]).reshape(N, N)
    got = run(ctx, x, "f64", lambda f, g: ctx.fresnel(f, g, DX, DX, LAM, z))[0]
    e_fr = rel(got, ref)
    print(f"ASM f64 {e_asm:.2e}, f32 {e_asm32:.2e}; Fresnel f64 {e_fr:.2e}", flush=True)
    assert e_asm < 1e-12 and e_asm32 < 1e-5
    assert e_fr < 1e-10

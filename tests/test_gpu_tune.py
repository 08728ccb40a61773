"""fft2d tuning: every configuration the tuner can pick (column split, column
tile variant, row-pass LDS planes, chunking) computes the same transform.

Each variant is checked against numpy's FFT (FP32 gate 1e-5 relative L2) and
by its inverse (round trip 1e-6), and the autotuner's pick is checked the
same way.
"""
import itertools

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def valid(nx, ny, n1, colvar, rowplanes):
    """Which configurations exist (FP32)."""
    if n1 > 1 and (ny % n1 or nx < 256 or nx * n1 > 16384):
        return False
    variant = colvar != 0 or rowplanes != 2
    if variant and (nx < 256 or ny // n1 < 256):
        return False
    if rowplanes == 1 and n1 > 1 and nx < 512:
        return False
    return True


@pytest.mark.parametrize("nx,ny,batch", [(2048, 2048, 2), (1024, 1024, 3), (512, 4096, 1),
                                         (4096, 1024, 1), (256, 2048, 2)])
def test_every_variant_computes_fft2d(ctx, nx, ny, batch):
    print(f"tune variants {nx}x{ny}x{batch}", flush=True)
    rng = np.random.default_rng(nx + ny + batch)
    u = rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))
    ref = np.fft.fft2(u)
    f = ctx.field(nx, ny, batch, "f32")
    ran = 0
    for n1, cv, rp, ch in itertools.product((1, 2, 4), range(6), (2, 1), (0, 8 << 20)):
        if not valid(nx, ny, n1, cv, rp):
            continue
        ctx.tune_set(nx, ny, batch, "f32", n1, cv, rp, ch)
        f.upload(u.real, u.imag)
        ctx.fft2d(f)
        re, im = f.download()
        assert rel(re + 1j * im, ref) < 1e-5, (n1, cv, rp, ch)
        ctx.fft2d(f, inverse=True)
        re, im = f.download()
        assert rel(re + 1j * im, u) < 1e-6, (n1, cv, rp, ch)
        ran += 1
    assert ran >= 8
    ctx.tune_set(nx, ny, batch, "f32")  # back to the built-in configuration
    f.free()


# half-width storage: FP32 arithmetic, rounding to the storage type after
# each pass (the storage gates of the fft2d parity tests)
HALF_TOL = {"f16": 2e-3, "bf16": 1.5e-2}


@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_half_width_variants_compute_fft2d(ctx, dt):
    nx, ny, batch = 1024, 1024, 2
    print(f"tune variants {dt} {nx}x{ny}x{batch}", flush=True)
    rng = np.random.default_rng(17)
    u = (rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))) * 0.01
    ref = np.fft.fft2(u)
    f = ctx.field(nx, ny, batch, dt)
    ran = 0
    for n1, cv, rp in itertools.product((1, 2, 4), range(6), (2, 1)):
        if not valid(nx, ny, n1, cv, rp):
            continue
        ctx.tune_set(nx, ny, batch, dt, n1, cv, rp, 0)
        f.upload(u.real, u.imag)
        ctx.fft2d(f)
        re, im = f.download()
        assert rel(re + 1j * im, ref) < HALF_TOL[dt], (dt, n1, cv, rp)
        ran += 1
    assert ran >= 30
    ctx.tune_set(nx, ny, batch, dt)
    f.free()


@pytest.mark.parametrize("nx,ny,batch", [(1024, 1024, 2), (2048, 2048, 1), (4096, 512, 1)])
def test_fp64_variants_compute_fft2d(ctx, nx, ny, batch):
    # FP64: the split (N1 rows of at most 8192 points per row group), the
    # two column tiles and both row passes
    print(f"tune variants f64 {nx}x{ny}x{batch}", flush=True)
    rng = np.random.default_rng(23)
    u = rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))
    ref = np.fft.fft2(u)
    f = ctx.field(nx, ny, batch, "f64")
    ran = 0
    for n1, cv, rp, ch in itertools.product((1, 2, 4), (0, 1), (2, 1), (0, 8 << 20)):
        if n1 > 1 and nx * n1 > 8192:
            continue
        if not valid(nx, ny, n1, cv, rp):
            continue
        ctx.tune_set(nx, ny, batch, "f64", n1, cv, rp, ch)
        f.upload(u.real, u.imag)
        ctx.fft2d(f)
        re, im = f.download()
        assert rel(re + 1j * im, ref) < 1e-13, (n1, cv, rp, ch)
        ctx.fft2d(f, inverse=True)
        re, im = f.download()
        assert rel(re + 1j * im, u) < 1e-14, (n1, cv, rp, ch)
        ran += 1
    assert ran >= 6
    ctx.tune_set(nx, ny, batch, "f64")
    f.free()


def test_autotune_picks_a_working_configuration(ctx):
    nx, ny, batch = 1024, 1024, 2
    rep = ctx.autotune_fft2d(nx, ny, batch, "f32", reps=2)
    lines = rep.strip().splitlines()
    assert lines[-1].startswith("best ")
    assert len(lines) >= 9
    t = ctx.tune_get(nx, ny, batch, "f32")
    best = dict(kv.split("=") for kv in lines[-1].split()[1:])
    assert int(best["n1"]) == t["n1"] and int(best["colvar"]) == t["colvar"]
    assert int(best["rowplanes"]) == t["rowplanes"] and int(best["chunk"]) == t["chunk"]
    rng = np.random.default_rng(5)
    u = rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))
    f = ctx.upload(u.real, u.imag, "f32")
    ctx.fft2d(f)
    re, im = f.download()
    assert rel(re + 1j * im, np.fft.fft2(u)) < 1e-5
    f.free()


@pytest.mark.parametrize("dt,tol", [("f16", 2e-3), ("f64", 1e-13)])
def test_autotune_other_storage(ctx, dt, tol):
    # the tuner over the FP16 / FP64 variants: every candidate line parses,
    # the pick is the fastest line and computes the transform
    nx, ny, batch = 1024, 1024, 2
    rep = ctx.autotune_fft2d(nx, ny, batch, dt, reps=2)
    lines = rep.strip().splitlines()
    cand = [dict(kv.split("=") for kv in ln.split()) for ln in lines[:-1]]
    best = dict(kv.split("=") for kv in lines[-1].split()[1:])
    assert len(cand) >= (8 if dt == "f64" else 30)
    fastest = min(cand, key=lambda c: float(c["ms"]))
    assert float(best["ms"]) == float(fastest["ms"])
    t = ctx.tune_get(nx, ny, batch, dt)
    assert int(best["n1"]) == t["n1"] and int(best["colvar"]) == t["colvar"]
    assert int(best["rowplanes"]) == t["rowplanes"]
    rng = np.random.default_rng(8)
    u = (rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))) * 0.01
    f = ctx.upload(u.real, u.imag, dt)
    ctx.fft2d(f)
    re, im = f.download()
    assert rel(re + 1j * im, np.fft.fft2(u)) < tol
    f.free()
    ctx.tune_set(nx, ny, batch, dt)


def test_tune_argument_errors(ctx):
    with pytest.raises(nat.DoeError, match="n1"):
        ctx.tune_set(512, 512, 1, "f32", n1=3)
    with pytest.raises(nat.DoeError, match="colvar"):
        ctx.tune_set(512, 512, 1, "f32", colvar=7)
    with pytest.raises(nat.DoeError, match="colvar of FP64"):
        ctx.tune_set(512, 512, 1, "f64", colvar=2)
    # a split the shape cannot take is refused when the transform runs
    f = ctx.field(128, 256, 1, "f32")
    ctx.tune_set(128, 256, 1, "f32", n1=4)
    with pytest.raises(nat.DoeError, match="column split 4 does not fit 128x256"):
        ctx.fft2d(f)
    ctx.tune_set(128, 256, 1, "f32")
    f.free()
    # an untuned shape reports the built-in choice
    t = ctx.tune_get(2048, 2048, 16, "f32")
    assert t["n1"] == 4 and t["colvar"] == 0 and t["rowplanes"] == 2
    # half-width storage of 2048+ lines: plain passes, one-plane rows
    for dt in ("f16", "bf16"):
        t = ctx.tune_get(2048, 2048, 16, dt)
        assert t["n1"] == 1 and t["colvar"] == 0 and t["rowplanes"] == 1
    t = ctx.tune_get(1024, 1024, 16, "f16")
    assert t["n1"] == 2 and t["rowplanes"] == 2
    # FP64: columns of 1024+ points split in two, with one-plane rows
    t = ctx.tune_get(2048, 2048, 16, "f64")
    assert t["n1"] == 2 and t["rowplanes"] == 1
    assert ctx.tune_get(4096, 4096, 1, "f64")["n1"] == 2
    t = ctx.tune_get(8192, 4096, 1, "f64")
    assert t["n1"] == 1 and t["rowplanes"] == 2

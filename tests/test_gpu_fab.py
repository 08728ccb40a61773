"""Fabrication output on the device against the oracle: error diffusion
(both scan orders, one lane per image for serpentine rows, a wavefront for
plain rows) and the etch-depth map (ref: src/doe.js:185-226, 967-976).
Gate: equal bit for bit (every step is exact IEEE arithmetic in the
reference's order)."""
import math
import time

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
import oracle_js as oj
from bitcmp import assert_same_bits

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def device_fs(ctx, phases, n, levels, snake, wrap_first=True):
    f = ctx.upload(np.stack([p.reshape(n, n) for p in phases]), None, "f64")
    o = ctx.field(n, n, len(phases), "f64")
    ctx.error_diffusion(f, o, levels, snake, wrap_first)
    re, _ = o.download()
    f.free()
    o.free()
    return re.reshape(len(phases), -1)


@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("snake", [True, False])
@pytest.mark.parametrize("levels", [2, 4, 3, 1, -3])
def test_error_diffusion_bit_exact(ctx, n, snake, levels):
    rng = np.random.default_rng(n * 10 + levels)
    phi = rng.uniform(-6, 6, n * n)
    print("error diffusion", n, snake, levels, flush=True)
    out, _ = oj.run("error_diffusion", {"levels": levels, "opts": {"snake": snake}}, {"phase": phi})
    got = device_fs(ctx, [phi], n, levels, snake)[0]
    assert_same_bits(got, out["out"])


@pytest.mark.parametrize("snake", [True, False])
def test_error_diffusion_batch_and_no_wrap(ctx, snake):
    rng = np.random.default_rng(3)
    n = 48
    phis = [rng.uniform(-3, 3, n * n) for _ in range(5)]
    got = device_fs(ctx, phis, n, 8, snake, wrap_first=False)
    for q, phi in enumerate(phis):
        out, _ = oj.run("error_diffusion", {"levels": 8, "opts": {"snake": snake, "wrapFirst": False}},
                        {"phase": phi})
        assert_same_bits(got[q], out["out"])


@pytest.mark.parametrize("snake", [True, False])
@pytest.mark.parametrize("levels", [4, 3])
def test_error_diffusion_half_level_inputs(ctx, snake, levels):
    """Quotients on and next to half-integers: the level decision there goes
    through the IEEE division (the fast reciprocal quotient is only trusted
    away from them)."""
    n = 64
    st = 2 * np.pi / levels
    rng = np.random.default_rng(levels)
    k = rng.integers(-levels, levels, n * n)
    for phi in ((k + 0.5) * st, np.full(n * n, 0.5 * st),
                np.nextafter((k + 0.5) * st, np.where(k % 2 == 0, np.inf, -np.inf))):
        for wrap_first in (True, False):
            out, _ = oj.run("error_diffusion", {"levels": levels, "opts": {"snake": snake, "wrapFirst": wrap_first}},
                            {"phase": phi})
            got = device_fs(ctx, [phi], n, levels, snake, wrap_first)[0]
            assert_same_bits(got, out["out"])


def _wrap(p):
    p = math.fmod(p, 2 * math.pi)
    if p > math.pi:
        p -= 2 * math.pi
    elif p <= -math.pi:
        p += 2 * math.pi
    return p


def _gather_fs(x, levels, wrap_first=True, snake=True):
    """Floyd–Steinberg of an ny x nx grid in gather form (each sample pulls
    its 1/16, 5/16, 3/16 shares from the row above in that row's scan order,
    then 7/16 of its predecessor); equal to the oracle's push form on square
    grids (tests/test_fab_cpu.py)."""
    ny, nx = x.shape
    step = 2 * math.pi / levels
    err = np.zeros((ny, nx))
    out = np.zeros((ny, nx))
    for j in range(ny):
        d = -1 if (snake and j % 2 == 1) else 1
        du = -1 if (snake and (j - 1) % 2 == 1) else 1
        prev = None
        for c in range(nx):
            i = c if d == 1 else nx - 1 - c
            v = _wrap(x[j, i]) if wrap_first else x[j, i]
            if j > 0:
                for src, w in ((i - du, 1 / 16), (i, 5 / 16), (i + du, 3 / 16)):
                    if 0 <= src < nx:
                        v += err[j - 1, src] * w
            if prev is not None:
                v += prev * 7 / 16
            f = math.floor(v / step)
            r = float(f + 1 if v / step - f >= 0.5 else f)
            q = (math.copysign(0.0, v / step) if r == 0 else r) * step  # Math.round's -0
            e = v - q
            if e > math.pi:
                e -= 2 * math.pi
            elif e <= -math.pi:
                e += 2 * math.pi
            out[j, i], err[j, i], prev = q, e, e
    return np.vectorize(_wrap)(out)


@pytest.mark.parametrize("nx,ny", [(40, 1500), (8, 2100), (4, 5000)])
def test_error_diffusion_plain_rows_tall_images(ctx, nx, ny):
    """More rows than the wavefront's 1024 threads: 2, 4 and 8 rows per thread."""
    rng = np.random.default_rng(ny)
    x = rng.uniform(-5, 5, (1, ny, nx))
    f = ctx.upload(x, None, "f64")
    o = ctx.field(nx, ny, 1, "f64")
    ctx.error_diffusion(f, o, 4, False, True)
    re, _ = o.download()
    assert_same_bits(re[0], _gather_fs(x[0], 4, snake=False))
    f.free()
    o.free()


def test_error_diffusion_widest_lds_rows(ctx):
    """Rows of 4096 samples: the widest the workgroup schedule stages (four
    LDS rows of 4112 doubles)."""
    nx, ny, B = 4096, 3, 2
    rng = np.random.default_rng(13)
    x = rng.uniform(-5, 5, (B, ny, nx))
    f = ctx.upload(x, None, "f64")
    o = ctx.field(nx, ny, B, "f64")
    ctx.error_diffusion(f, o, 3, True, True)
    re, _ = o.download()
    for b in range(B):
        assert_same_bits(re[b], _gather_fs(x[b], 3))
    f.free()
    o.free()


@pytest.mark.parametrize("levels", [4, 1])
def test_error_diffusion_long_rows_one_lane_per_image(ctx, levels):
    """Rows longer than 4096 samples take the one-lane-per-image schedule."""
    nx, ny, B = 4100, 3, 2
    rng = np.random.default_rng(12)
    x = rng.uniform(-5, 5, (B, ny, nx))
    f = ctx.upload(x, None, "f64")
    o = ctx.field(nx, ny, B, "f64")
    ctx.error_diffusion(f, o, levels, True, True)
    re, _ = o.download()
    for b in range(B):
        assert_same_bits(re[b], _gather_fs(x[b], levels))
    f.free()
    o.free()


def test_error_diffusion_smooth_phase_and_timing(ctx):
    n = 1024
    j, i = np.mgrid[0:n, 0:n].astype(float)
    phi = (np.sin(i / 37) * 3 + np.cos(j / 23) * 2).ravel()
    for snake in (False, True):
        f = ctx.upload(phi.reshape(1, n, n), None, "f64")
        o = ctx.field(n, n, 1, "f64")
        ctx.error_diffusion(f, o, 4, snake)  # warm
        t0 = time.perf_counter()
        ctx.error_diffusion(f, o, 4, snake)
        ms = (time.perf_counter() - t0) * 1e3
        re, _ = o.download()
        print(f"errorDiffusionQuantize 1024^2 snake={snake}: {ms:.2f} ms", flush=True)
        st = 2 * np.pi / 4
        m = np.mod(re.ravel() + np.pi, st)
        assert np.all(np.minimum(m, st - m) < 1e-9)
        f.free()
        o.free()
    out, _ = oj.run("error_diffusion", {"levels": 4, "opts": {"snake": False}}, {"phase": phi})
    f = ctx.upload(phi.reshape(1, n, n), None, "f64")
    o = ctx.field(n, n, 1, "f64")
    ctx.error_diffusion(f, o, 4, False)
    re, _ = o.download()
    assert_same_bits(re.ravel(), out["out"])


def test_phase_to_height_bit_exact(ctx):
    rng = np.random.default_rng(8)
    phase = rng.uniform(-50, 50, (3, 40, 24))
    for lam, nidx in ((633e-9, 1.52), (532e-9, 1.46)):
        out, _ = oj.run("phase_to_height", {"lambda": lam, "n": nidx}, {"phase": phase})
        f = ctx.upload(phase, None, "f64")
        o = ctx.field(24, 40, 3, "f64")
        ctx.phase_to_height(f, o, lam, nidx)
        re, _ = o.download()
        assert_same_bits(re.ravel(), out["out"])
        f.free()
        o.free()


def test_argument_errors(ctx):
    f = ctx.field(8, 8, 1, "f64")
    g = ctx.field(8, 8, 1, "f32")
    with pytest.raises(nat.DoeError, match="FP64"):
        ctx.error_diffusion(f, g)
    with pytest.raises(nat.DoeError, match="overlaps"):
        ctx.error_diffusion(f, f)
    # an output that starts one row into the phase overlaps it too
    shifted = ctx.wrap(f.ptrs()[0] + 8 * 8, f.ptrs()[1] + 8 * 8, 8, 7, 1, "f64")
    head = ctx.wrap(f.ptrs()[0], f.ptrs()[1], 8, 7, 1, "f64")
    with pytest.raises(nat.DoeError, match="overlaps"):
        ctx.error_diffusion(head, shifted)
    shifted.free()
    head.free()
    with pytest.raises(nat.DoeError, match="non-zero"):
        ctx.error_diffusion(f, ctx.field(8, 8, 1, "f64"), levels=0)
    f.free()
    g.free()

"""doe_screen_metrics_noalloc (no device allocation; fixed per-GPU scratch)
against doe_screen_metrics on the same screens: every figure equal bit for
bit (the same kernels reduce in the same order), on two calls in a row, the
second on a smaller screen, with and without spot cells; sizes beyond the
scratch are refused with DOE_ERR_UNSUPPORTED before anything runs."""
import math

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def case(ctx, n, seed, dtype="f32"):
    rng = np.random.default_rng(seed)
    scr = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    j, i = np.mgrid[0:n, 0:n]
    t = (np.hypot(i - n / 2, j - n / 2) < 0.3 * n).astype(float)
    roi = (np.hypot(i - n / 2, j - n / 2) < 0.35 * n).astype(float) * 0.8
    el = np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n)))
    return (ctx.upload(scr.real, scr.imag, dtype), ctx.upload(t, roi, "f64"),
            ctx.upload(el.real, el.imag, "f32"))


def same(a, b):
    for k in a:
        x, y = a[k], b[k]
        if x is None or y is None:
            assert x is None and y is None, k
        elif math.isnan(x) or math.isnan(y):
            assert math.isnan(x) and math.isnan(y), k
        else:
            assert x == y, (k, x, y)


@pytest.mark.parametrize("grid", [None, (8, 6)])
def test_noalloc_equals_allocating_entry_point(ctx, grid):
    for n, seed, dtype in ((1024, 1, "f32"), (200, 2, "f64")):  # second call: a smaller screen
        print(f"noalloc screen metrics n={n} grid={grid}", flush=True)
        scr, tgt, el = case(ctx, n, seed, dtype)
        args = (scr, tgt, el, 4e-6, 4e-6, 8e-6, 8e-6, grid)
        ref = ctx.screen_metrics(*args)
        got = ctx.screen_metrics(*args, noalloc=True)
        same(got, ref)
        again = ctx.screen_metrics(*args, noalloc=True)  # the scratch keeps the last call: no effect
        same(again, ref)
        for f in (scr, tgt, el):
            f.free()


def test_noalloc_refuses_beyond_the_scratch(ctx):
    scr, tgt, el = case(ctx, 64, 3)
    with pytest.raises(nat.DoeError, match="4096 spot cells"):
        ctx.screen_metrics(scr, tgt, el, 1, 1, 1, 1, (65, 64), noalloc=True)
    for f in (scr, tgt, el):
        f.free()
    # 16384 x 8192 points need ~6.4 MB of partials: refused before any kernel runs
    big = ctx.field(16384, 8192, 1, "f32")
    bt = ctx.field(16384, 8192, 1, "f64")
    small = ctx.field(4, 4, 1, "f32")
    with pytest.raises(nat.DoeError, match="bytes of scratch"):
        ctx.screen_metrics(big, bt, small, 1, 1, 1, 1, None, noalloc=True)
    for f in (big, bt, small):
        f.free()

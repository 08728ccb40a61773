"""Pins for the oracle's fabrication output (CPU): error diffusion, phase
rounding, etch depth (ref: src/doe.js:167-226, 967-976).

* error diffusion against an independent gather-form restatement (each
  sample pulls its shares from the already-quantised neighbours instead of
  the scan pushing them), which catches a mirrored or misplaced tap;
* the reference's own behaviour tests (ref test/doe.test.js:236-275, 341-348):
  levels on the level grid, error diffusion beats rounding on a low-passed
  field, heights inside [0, lambda / (n - 1)];
* closed forms: one-row images are 1-D error feedback, a phase already on
  the levels passes unchanged, heights equal numpy's fmod form.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle_js as oj
from bitcmp import assert_same_bits

pytestmark = pytest.mark.skipif(not oj.node_available(), reason="node is required for the oracle")


def js_round(x):
    """ECMAScript Math.round: ties toward +Infinity, and -0 for x in [-0.5, -0]."""
    f = math.floor(x)
    r = float(f + 1 if x - f >= 0.5 else f)
    return math.copysign(0.0, x) if r == 0 else r


def wrap(p):
    p = math.fmod(p, 2 * math.pi)
    if p > math.pi:
        p -= 2 * math.pi
    elif p <= -math.pi:
        p += 2 * math.pi
    return p


def gather_fs(phi, n, levels, snake=True, wrap_first=True):
    """Floyd–Steinberg in gather form: sample (i, j) = its value + 1/16, 5/16,
    3/16 of the errors above (in the row above's scan order) + 7/16 of the
    previous error in its own row."""
    x = phi.reshape(n, n)
    step = 2 * math.pi / levels
    err = np.zeros((n, n))
    out = np.zeros((n, n))
    for j in range(n):
        d = -1 if (snake and j % 2 == 1) else 1
        du = -1 if (snake and (j - 1) % 2 == 1) else 1  # direction of the row above
        prev = None
        for c in range(n):
            i = c if d == 1 else n - 1 - c
            v = wrap(x[j, i]) if wrap_first else x[j, i]
            if j > 0:
                # the row above scatters 1/16 ahead, 5/16 below, 3/16 behind (in its direction)
                for src, w in ((i - du, 1 / 16), (i, 5 / 16), (i + du, 3 / 16)):
                    if 0 <= src < n:
                        v += err[j - 1, src] * w
            if prev is not None:
                v += prev * 7 / 16
            q = js_round(v / step) * step
            e = v - q
            if e > math.pi:
                e -= 2 * math.pi
            elif e <= -math.pi:
                e += 2 * math.pi
            out[j, i], err[j, i], prev = q, e, e
    return np.vectorize(wrap)(out).ravel()


@pytest.mark.parametrize("n,levels,snake,wrap_first", [(12, 4, True, True), (16, 3, False, True),
                                                       (9, 8, True, False), (16, 2, False, False)])
def test_error_diffusion_matches_gather_form(n, levels, snake, wrap_first):
    rng = np.random.default_rng(n + levels)
    phi = rng.uniform(-7, 7, n * n)
    out, _ = oj.run("error_diffusion", {"levels": levels, "opts": {"snake": snake, "wrapFirst": wrap_first}},
                    {"phase": phi})
    assert_same_bits(out["out"], gather_fs(phi, n, levels, snake, wrap_first))


def test_error_diffusion_keeps_negative_zero_levels():
    """Math.round gives -0 for quotients in [-0.5, 0), so such levels are -0
    in the reference's output (and the wrap keeps the sign)."""
    rng = np.random.default_rng(5)
    n = 16
    phi = rng.uniform(-0.6, 0.6, n * n)
    out, _ = oj.run("error_diffusion", {"levels": 4, "opts": {"snake": True}}, {"phase": phi})
    neg0 = (out["out"] == 0) & np.signbit(out["out"])
    assert neg0.sum() > 10
    assert_same_bits(out["out"], gather_fs(phi, n, 4))


def test_error_diffusion_reference_behaviour():
    """ref test/doe.test.js:236-275"""
    n = 64
    j, i = np.mgrid[0:n, 0:n].astype(float)
    phase = (np.sin(i / 7) + np.cos(j / 5.5) + 0.3 * np.sin((i + j) / 3)).ravel()
    d, _ = oj.run("error_diffusion", {"levels": 4}, {"phase": phase})
    r, _ = oj.run("quantize", {"levels": 4}, {"phase": phase})

    def smooth_err(q):
        f = np.exp(1j * q.reshape(n, n))
        g = np.exp(1j * phase.reshape(n, n))
        lp = lambda z: ndimage.gaussian_filter(z.real, 2, mode="nearest") + 1j * ndimage.gaussian_filter(
            z.imag, 2, mode="nearest")
        return np.linalg.norm(lp(f) - lp(g)) / np.linalg.norm(lp(g))

    assert smooth_err(d["out"]) < smooth_err(r["out"])
    st = 2 * np.pi / 4
    m = np.mod(d["out"] + np.pi, st)
    assert np.all(np.minimum(m, st - m) < 1e-9)


def test_error_diffusion_closed_forms():
    # one row: plain error feedback
    rng = np.random.default_rng(4)
    phi = rng.uniform(-3, 3, 1)  # a 1 x 1 grid
    out, _ = oj.run("error_diffusion", {"levels": 5}, {"phase": phi})
    st = 2 * np.pi / 5
    assert out["out"][0] == wrap(js_round(phi[0] / st) * st)
    # a phase already on the levels passes unchanged (zero error everywhere)
    n, levels = 16, 8
    st = 2 * np.pi / levels
    k = rng.integers(-3, 4, n * n)
    lv = np.vectorize(wrap)(k * st)
    out, _ = oj.run("error_diffusion", {"levels": levels}, {"phase": lv})
    np.testing.assert_allclose(out["out"], lv, atol=1e-12)


def test_phase_to_height_range_and_closed_form():
    """ref test/doe.test.js:341-348"""
    lam, nidx = 633e-9, 1.52
    rng = np.random.default_rng(6)
    phase = rng.uniform(-40, 40, 4096)
    out, _ = oj.run("phase_to_height", {"lambda": lam, "n": nidx}, {"phase": phase})
    h = out["out"]
    s, step = lam / (2 * np.pi * (nidx - 1)), lam / (nidx - 1)
    np.testing.assert_array_equal(h, np.fmod(np.fmod(phase * s, step) + step, step))
    assert np.all(h >= 0) and np.all(h <= lam / 0.52 + 1e-18)
    # a 2 pi step of phase is one full wave of etch depth
    h2, _ = oj.run("phase_to_height", {"lambda": lam, "n": nidx}, {"phase": np.array([0.5, 0.5 + 2 * np.pi])})
    assert abs(h2["out"][1] - h2["out"][0]) < 1e-15

"""The warm-start network's CPU oracle (oracle/nn_ref.py) pinned by finite
differences: its analytic gradient of the phase-aligned loss equals central
differences of its forward pass (FP64, tiny network)."""
import numpy as np
import pytest

import sys
from pathlib import Path
sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import nn_ref  # noqa: E402


def tiny_problem(seed=0, res=16, base=2, B=2):
    rng = np.random.default_rng(seed)
    p = nn_ref.init_params(res, base, seed)
    x = rng.uniform(0, 1, (B, res, res, 1))
    ph = rng.uniform(-np.pi, np.pi, (B, res, res))
    w = rng.uniform(0.1, 1, (B, res, res))
    return p, x, np.cos(ph), np.sin(ph), w


def test_layer_count_and_params():
    L, slots = nn_ref.layers(256, 32)
    _, n = nn_ref.param_offsets(L)
    assert sum(1 for l in L if l[0] == "conv") == 15
    assert n == 1753314  # 1.75 M parameters at base width 32


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 5, 6, 3))
    W = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = nn_ref.conv3x3(x, W, b)
    ref = np.zeros((1, 5, 6, 4))
    for h in range(5):
        for w in range(6):
            for co in range(4):
                s = b[co]
                for kh in range(3):
                    for kw in range(3):
                        hh, ww = h + kh - 1, w + kw - 1
                        if 0 <= hh < 5 and 0 <= ww < 6:
                            s += np.dot(x[0, hh, ww], W[co, kh, kw])
                ref[0, h, w, co] = s
    assert np.max(np.abs(y - ref)) < 1e-12


def test_loss_is_phase_and_conjugate_invariant():
    rng = np.random.default_rng(2)
    ph = rng.uniform(-np.pi, np.pi, (1, 8, 8))
    w = rng.uniform(0.1, 1, (1, 8, 8))
    yr, yi = np.cos(ph), np.sin(ph)
    for z in (np.exp(1j * (ph + 0.7)), np.exp(-1j * ph + 0.3j)):
        out = np.stack([z.real, z.imag], -1)
        assert nn_ref.loss(out, yr, yi, w) < 1e-12
    out = np.stack([np.cos(ph + 1.0 * np.arange(8)), np.sin(ph + 1.0 * np.arange(8))], -1)
    assert nn_ref.loss(out, yr, yi, w) > 0.1


def test_backward_matches_finite_differences():
    p, x, yr, yi, w = tiny_problem()
    l0, g = nn_ref.backward(p, x, yr, yi, w, 16, 2)
    assert abs(l0 - nn_ref.loss(nn_ref.forward(p, x, 16, 2)[0], yr, yi, w)) < 1e-14
    rng = np.random.default_rng(3)
    idx = rng.choice(p.size, 60, replace=False)
    eps = 1e-6
    fd = np.zeros(idx.size)
    for j, i in enumerate(idx):
        q = p.copy()
        q[i] += eps
        lp = nn_ref.loss(nn_ref.forward(q, x, 16, 2)[0], yr, yi, w)
        q[i] -= 2 * eps
        lm = nn_ref.loss(nn_ref.forward(q, x, 16, 2)[0], yr, yi, w)
        fd[j] = (lp - lm) / (2 * eps)
    assert np.linalg.norm(fd - g[idx]) <= 1e-6 * np.linalg.norm(fd) + 1e-10

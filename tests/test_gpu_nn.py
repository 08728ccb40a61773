"""Warm-start network on the GPU (BF16 MFMA kernels) against the CPU oracle
(oracle/nn_ref.py, pinned by finite differences in test_nn_cpu.py).

Gates: forward output within 1e-2 relative L2 of the oracle run with the
GPU's BF16 storage rounding; loss within 1e-3; gradients within 5e-2
relative L2 of the oracle's (BF16 operands of the backward products), and
within 5e-2 of central finite differences of the oracle's FP64 loss; a few
Adam steps on a fixed batch lower the loss.  The warm-start hand-off
(inputs, labels, initial phase) is checked against numpy restatements.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
from test_gpu_gs import problem

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import nn_ref  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


class Bufs:
    """device buffers for a batch of res x res network inputs/outputs"""

    def __init__(self, ctx, res, B):
        self.inp = ctx.field(res, res, B, "f32")
        self.lab = ctx.field(res, res, B, "f32")
        self.w = ctx.field(res, res, B, "f32")
        self.out = ctx.field(2 * res, res, B, "f32")
        self.res, self.B = res, B

    def set(self, x, yr, yi, w):
        z = np.zeros_like(x)
        self.inp.upload(x, z)
        self.lab.upload(yr, yi)
        self.w.upload(w, z)

    def output(self):
        re, _ = self.out.download()
        return re.reshape(self.B, self.res, self.res, 2)

    def free(self):
        for f in (self.inp, self.lab, self.w, self.out):
            f.free()


def problem_nn(res, B, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (B, res, res))
    ph = rng.uniform(-np.pi, np.pi, (B, res, res))
    w = rng.uniform(0.1, 1, (B, res, res))
    return x, np.cos(ph), np.sin(ph), w


@pytest.mark.parametrize("res,base", [(16, 4), (64, 8)])
def test_forward_matches_oracle(ctx, res, base):
    B = 2
    net = nat.WsNet(ctx, res, base, B, seed=1)
    p = nn_ref.bf16_round(nn_ref.init_params(res, base, 5)).astype(np.float32)
    net.set_params(p)
    x, yr, yi, w = problem_nn(res, B, 0)
    buf = Bufs(ctx, res, B)
    buf.set(x, yr, yi, w)
    net.forward(B, buf.inp.ptrs()[0], buf.out.ptrs()[0])
    got = buf.output()
    ref, _ = nn_ref.forward(p.astype(np.float64), x[..., None], res, base, bf16=True)
    assert rel(got, ref) < 1e-2
    buf.free()
    net.close()


def test_gradients_match_oracle_and_finite_differences(ctx):
    res, base, B = 16, 4, 2
    net = nat.WsNet(ctx, res, base, B, seed=1)
    p = nn_ref.bf16_round(nn_ref.init_params(res, base, 6)).astype(np.float32)
    net.set_params(p)
    x, yr, yi, w = problem_nn(res, B, 1)
    buf = Bufs(ctx, res, B)
    buf.set(x, yr, yi, w)
    loss, g = net.loss_grad(B, buf.inp.ptrs()[0], buf.lab.ptrs()[0], buf.lab.ptrs()[1],
                            buf.w.ptrs()[0], want_grad=True)
    p64 = p.astype(np.float64)
    lref, gref = nn_ref.backward(p64, x[..., None], yr, yi, w, res, base, bf16=True)
    assert abs(loss - lref) < 1e-3 * max(1.0, abs(lref))
    assert rel(g, gref) < 5e-2
    # per layer, so no layer hides behind a large one
    L, _ = nn_ref.layers(res, base)
    offs, _ = nn_ref.param_offsets(L)
    for o in offs:
        if o is None:
            continue
        sl = slice(o[0], o[1])
        if np.linalg.norm(gref[sl]) > 1e-3 * np.linalg.norm(gref):
            assert rel(g[sl], gref[sl]) < 1e-1, o
    # against central differences of the oracle's FP64 loss (no rounding)
    rng = np.random.default_rng(9)
    idx = rng.choice(p.size, 40, replace=False)
    fd = np.zeros(idx.size)
    eps = 1e-4
    for j, i in enumerate(idx):
        q = p64.copy()
        q[i] += eps
        lp = nn_ref.loss(nn_ref.forward(q, x[..., None], res, base)[0], yr, yi, w)
        q[i] -= 2 * eps
        lm = nn_ref.loss(nn_ref.forward(q, x[..., None], res, base)[0], yr, yi, w)
        fd[j] = (lp - lm) / (2 * eps)
    assert rel(g[idx], fd) < 5e-2
    buf.free()
    net.close()


def test_adam_steps_lower_the_loss(ctx):
    res, base, B = 32, 4, 4
    net = nat.WsNet(ctx, res, base, B, seed=3)
    x, yr, yi, w = problem_nn(res, B, 2)
    # a learnable target: the label phase is a smooth function of the input
    ph = 2.0 * x + 0.5
    yr, yi = np.cos(ph), np.sin(ph)
    buf = Bufs(ctx, res, B)
    buf.set(x, yr, yi, w)
    l0, _ = net.loss_grad(B, buf.inp.ptrs()[0], buf.lab.ptrs()[0], buf.lab.ptrs()[1],
                          buf.w.ptrs()[0])
    for _ in range(40):
        net.adam_step(2e-3)
        l, _ = net.loss_grad(B, buf.inp.ptrs()[0], buf.lab.ptrs()[0], buf.lab.ptrs()[1],
                             buf.w.ptrs()[0])
    assert np.isfinite(l) and l < 0.5 * l0, (l0, l)
    buf.free()
    net.close()


def centred_far_field(ill, phi):
    n = ill.shape[0]
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(ill * np.exp(1j * phi)))) / n


def test_warmstart_handoff_matches_numpy(ctx):
    sides, res = [64, 128], 16
    b = nat.GsBatch(ctx, sides, nat.ALG_WGS, 0.9)
    probs = []
    for i, n in enumerate(sides):
        ill, t, _, phi0 = problem(n, 8e-6, 60 + i)
        b.set(i, ill, t, None, phi0)
        probs.append((ill, t, phi0))
    order = b.design_order()
    assert sorted(order) == [0, 1] and order[0] == 1  # larger designs first
    per = res * res
    scratch = ctx.field(per, 4, len(sides), "f32")  # >= count * res^2 * 3 floats
    sp = scratch.ptrs()[0]
    # inputs: pooled irradiance, max 1
    b.warm_inputs(res, sp)
    re, _ = scratch.download()
    flat = re.ravel()
    for j, u in enumerate(order):
        ill, t, _ = probs[u]
        n = sides[u]
        f = n // res
        pooled = (t ** 2).reshape(res, f, res, f).mean(axis=(1, 3))
        assert rel(flat[j * per:(j + 1) * per], (pooled / pooled.max()).ravel()) < 1e-5
    # labels from the current best phase: run 3 iterations, compare with numpy
    b.run(3, graph=False)
    lab = ctx.field(res, res, len(sides), "f32")
    wts = ctx.field(res, res, len(sides), "f32")
    b.warm_labels(res, lab.ptrs()[0], lab.ptrs()[1], wts.ptrs()[0])
    lr_, li_ = lab.download()
    wr, _ = wts.download()
    for j, u in enumerate(order):
        ill, t, _ = probs[u]
        n = sides[u]
        f = n // res
        Y = centred_far_field(ill, b.get(u)["phase"].reshape(n, n))
        S = Y.reshape(res, f, res, f).sum(axis=(1, 3))
        ref = S / np.abs(S)
        got = lr_[j] + 1j * li_[j]
        assert np.max(np.abs(got - ref)) < 1e-3
        assert rel(wr[j], t.reshape(res, f, res, f).mean(axis=(1, 3))) < 1e-5
    # initial phase from a prediction: one WGS iteration from the warm start
    # equals the single-design loop from numpy's arg(F^-1(t * phasor))
    net = nat.WsNet(ctx, res, 4, 2, seed=11)
    b.warm_apply(net, sp)
    b.run(1, graph=False)
    out = ctx.field(2 * res, res, len(sides), "f32")
    b.warm_inputs(res, sp)
    net.forward(len(sides), sp, out.ptrs()[0])
    z, _ = out.download()
    z = z.reshape(len(sides), res, res, 2)
    for j, u in enumerate(order):
        ill, t, _ = probs[u]
        n = sides[u]
        f = n / res
        c = (np.arange(n) + 0.5) / f - 0.5
        i0 = np.clip(np.floor(c).astype(int), 0, res - 1)
        i1 = np.minimum(i0 + 1, res - 1)
        a = np.clip(c - i0, 0, 1)
        zc = z[j, ..., 0] + 1j * z[j, ..., 1]
        rows = (1 - a)[:, None] * zc[i0] + a[:, None] * zc[i1]
        zz = (1 - a)[None, :] * rows[:, i0] + a[None, :] * rows[:, i1]
        Y = t * zz / np.abs(zz)
        x = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(Y)))
        phi0 = np.angle(x)
        plan = nat.GsPlan(mode=nat.MODE_FAR_FIELD, algorithm=nat.ALG_WGS, iterations=1, nx=n, ny=n,
                          dx=8e-6, dy=8e-6, lambda_=532e-9, z=0.03, beta=0.9, feedback=0,
                          escape=0, dtype=nat.DTYPES["f32"])
        r = ctx.gs_run(plan, ill, t, None, phi0)
        g = b.get(u)
        assert abs(g["error"][0] - r["error"][0]) < 1e-5, (u, g["error"], r["error"])
    for f_ in (scratch, lab, wts, out):
        f_.free()
    net.close()
    b.close()

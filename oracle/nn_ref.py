"""ORACLE — test infrastructure only (imported by tests/, never by the product).

CPU restatement (numpy, FP64) of the warm-start network of
diffractiveopticalelement_amd/csrc/nn.hip: the same layer list, parameter
layout ([Cout][3][3][Cin] weights then Cout biases, per convolution), NHWC
activations, additive-skip U-Net, ReLU, 2x2 average pooling, nearest
upsampling, and the phase-aligned, conjugate-aware weighted loss.  The
network has no counterpart in the reference (no file:line to follow); this
module is pinned by finite differences of its own loss (analytic backward vs
central differences), and the GPU is checked against it.  `bf16=True`
emulates the GPU's storage rounding (weights, activations and the matrix
cores' operands to bfloat16).
"""
from __future__ import annotations

import numpy as np


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(res, base):
    """(kind, cin, cout, relu, res_out, in_slot, out_slot, skip_slot) list and slot shapes."""
    C, R = base, res
    slots = []
    L = []

    def slot(c, r):
        slots.append((c, r))
        return len(slots) - 1

    cur = [slot(1, R)]

    def conv(cin, cout, r, relu):
        out = slot(cout, r) if (relu or cout != 2) else -1
        L.append(("conv", cin, cout, relu, r, cur[0], out, -1))
        cur[0] = out
        return out

    def pool(c, r):
        out = slot(c, r // 2)
        L.append(("pool", c, c, False, r // 2, cur[0], out, -1))
        cur[0] = out

    def up(c, r, skip):
        out = slot(c, r)
        L.append(("up", c, c, False, r, cur[0], out, skip))
        cur[0] = out

    conv(1, C, R, True)
    s1 = conv(C, C, R, True)
    pool(C, R)
    conv(C, 2 * C, R // 2, True)
    s2 = conv(2 * C, 2 * C, R // 2, True)
    pool(2 * C, R // 2)
    conv(2 * C, 4 * C, R // 4, True)
    s3 = conv(4 * C, 4 * C, R // 4, True)
    pool(4 * C, R // 4)
    conv(4 * C, 8 * C, R // 8, True)
    conv(8 * C, 8 * C, R // 8, True)
    conv(8 * C, 4 * C, R // 8, True)
    up(4 * C, R // 4, s3)
    conv(4 * C, 4 * C, R // 4, True)
    conv(4 * C, 2 * C, R // 4, True)
    up(2 * C, R // 2, s2)
    conv(2 * C, 2 * C, R // 2, True)
    conv(2 * C, C, R // 2, True)
    up(C, R, s1)
    conv(C, C, R, True)
    conv(C, 2, R, False)
    return L, slots


def param_offsets(L):
    offs = []
    o = 0
    for (kind, cin, cout, *_rest) in L:
        if kind == "conv":
            offs.append((o, o + cout * 9 * cin))
            o += cout * 9 * cin + cout
        else:
            offs.append(None)
    return offs, o


def conv3x3(x, W, b):
    """x [B,H,W,Cin], W [Cout,3,3,Cin] -> [B,H,W,Cout] (zero padding)."""
    B, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, Wd, W.shape[0]))
    for kh in range(3):
        for kw in range(3):
            y += xp[:, kh:kh + H, kw:kw + Wd, :] @ W[:, kh, kw, :].T
    return y + b


def conv3x3_bwd(x, W, dz):
    """gradients of conv3x3 wrt x, W, b given dz = dL/dy."""
    B, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for kh in range(3):
        for kw in range(3):
            xs = xp[:, kh:kh + H, kw:kw + Wd, :]
            dW[:, kh, kw, :] = np.einsum("bhwo,bhwi->oi", dz, xs)
            dxp[:, kh:kh + H, kw:kw + Wd, :] += dz @ W[:, kh, kw, :]
    return dxp[:, 1:-1, 1:-1, :], dW, dz.sum(axis=(0, 1, 2))


def pool2(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def forward(params, x, res, base, bf16=False):
    """x [B,res,res,1] -> (out [B,res,res,2], activations per slot)."""
    L, slots = layers(res, base)
    offs, _ = param_offsets(L)
    rnd = bf16_round if bf16 else (lambda a: a)
    act = [None] * len(slots)
    act[0] = rnd(x)
    out = None
    for li, (kind, cin, cout, relu, r, i, o, s) in enumerate(L):
        if kind == "conv":
            w0, b0 = offs[li]
            W = rnd(params[w0:w0 + cout * 9 * cin].reshape(cout, 3, 3, cin))
            b = params[b0:b0 + cout]
            y = conv3x3(act[i], W, b)
            if relu:
                y = np.maximum(y, 0)
            if o >= 0:
                act[o] = rnd(y)
            else:
                out = y
        elif kind == "pool":
            act[o] = rnd(pool2(act[i]))
        else:
            act[o] = rnd(up2(act[i]) + act[s])
    return out, act


def loss(out, yr, yi, w):
    """batch-mean phase-aligned, conjugate-aware weighted loss."""
    zr, zi = out[..., 0], out[..., 1]
    tot = 0.0
    for b in range(out.shape[0]):
        wc = w[b]
        W = wc.sum()
        Sr = np.sum(wc * (zr[b] * yr[b] + zi[b] * yi[b]))
        Si = np.sum(wc * (zi[b] * yr[b] - zr[b] * yi[b]))
        Tr = np.sum(wc * (zr[b] * yr[b] - zi[b] * yi[b]))
        Ti = np.sum(wc * (zi[b] * yr[b] + zr[b] * yi[b]))
        m = max(np.hypot(Sr, Si), np.hypot(Tr, Ti))
        tot += (np.sum(wc * (zr[b] ** 2 + zi[b] ** 2)) + W - 2 * m) / W
    return tot / out.shape[0]


def loss_grad(out, yr, yi, w):
    zr, zi = out[..., 0], out[..., 1]
    g = np.zeros_like(out)
    B = out.shape[0]
    for b in range(B):
        wc = w[b]
        W = wc.sum()
        Sr = np.sum(wc * (zr[b] * yr[b] + zi[b] * yi[b]))
        Si = np.sum(wc * (zi[b] * yr[b] - zr[b] * yi[b]))
        Tr = np.sum(wc * (zr[b] * yr[b] - zi[b] * yi[b]))
        Ti = np.sum(wc * (zi[b] * yr[b] + zr[b] * yi[b]))
        aS, aT = np.hypot(Sr, Si), np.hypot(Tr, Ti)
        conj = aT > aS
        m = aT if conj else aS
        dr, di = ((Tr, Ti) if conj else (Sr, Si))
        dr, di = (dr / m, di / m) if m > 0 else (0.0, 0.0)
        lr, li = yr[b], (-yi[b] if conj else yi[b])
        # |S| with S = sum w z conj(l): d|S|/dzr = w (dr lr - di li), d|S|/dzi = w (dr li + di lr)
        g[b, ..., 0] = (2 * wc * zr[b] - 2 * wc * (dr * lr - di * li)) / (W * B)
        g[b, ..., 1] = (2 * wc * zi[b] - 2 * wc * (dr * li + di * lr)) / (W * B)
    return g


def backward(params, x, yr, yi, w, res, base, bf16=False):
    """(loss, dL/dparams) by reverse-mode differentiation of forward()."""
    L, slots = layers(res, base)
    offs, npar = param_offsets(L)
    rnd = bf16_round if bf16 else (lambda a: a)
    out, act = forward(params, x, res, base, bf16)
    grad = np.zeros(npar)
    gact = [np.zeros_like(a) if a is not None else None for a in act]
    dout = loss_grad(out, yr, yi, w)
    for li in range(len(L) - 1, -1, -1):
        kind, cin, cout, relu, r, i, o, s = L[li]
        if kind == "conv":
            dz = dout if o < 0 else gact[o]
            if relu:
                dz = dz * (act[o] > 0)
            w0, b0 = offs[li]
            W = rnd(params[w0:w0 + cout * 9 * cin].reshape(cout, 3, 3, cin))
            dx, dW, db = conv3x3_bwd(act[i], W, dz)
            grad[w0:w0 + cout * 9 * cin] = dW.ravel()
            grad[b0:b0 + cout] = db
            if i > 0:
                gact[i] += dx
        elif kind == "pool":
            gact[i] += 0.25 * up2(gact[o])
        else:
            gact[s] += gact[o]
            g = gact[o]
            gact[i] += g[:, 0::2, 0::2] + g[:, 1::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 1::2]
    return loss(out, yr, yi, w), grad


def init_params(res, base, seed):
    L, _ = layers(res, base)
    offs, npar = param_offsets(L)
    rng = np.random.default_rng(seed)
    p = np.zeros(npar)
    for li, (kind, cin, cout, *_r) in enumerate(L):
        if kind == "conv":
            w0, b0 = offs[li]
            p[w0:w0 + cout * 9 * cin] = rng.normal(0, np.sqrt(2 / (9 * cin)), cout * 9 * cin)
            p[b0:b0 + cout] = rng.normal(0, 0.05, cout)
    return p

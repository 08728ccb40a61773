"""Independent pins for the oracle functions that no golden vector or
published figure reaches (CPU only):

* HIO relaxation t + fb (a - t) inside the ROI, field kept outside
  (ref: src/doe.js:379-397), with the escape period;
* far-field WGS under fourierCgh: weight update clamp(t/a, 1/4, 4)^beta,
  clamp to [1e-4, 1e4], renormalisation by sqrt(sum t^2 / sum w^2 t^2) over
  the ROI (doe.js:506-524), and the best-iterate bookkeeping that returns the
  phase after the best iteration's projection (doe.js:490-504);
* the a = 0 branch of the far-field projection, spectrum := target
  (doe.js:708-711);
* the raised-cosine band edge of the angular-spectrum transfer function
  (ref: src/propagate.js:38-42, 74-80);
* SSIM (metrics.js:198-259), zeroOrderFraction (145-157), rmsRadius
  (164-185), spotUniformity (284-321): closed forms;
* illumination profiles (doe.js:979-1000, field.js discMask): closed forms.

Each expectation is computed here with numpy's FFT and array code, from the
mathematics of the step (not from the oracle's loops), so a dropped term, a
swapped operand or a shifted index in the oracle fails it.
tests/test_oracle_mutants_cpu.py plants such mistakes in a copy of the
oracle and checks that these tests catch them.
"""
import numpy as np
import pytest

import oracle_js as oj

pytestmark = pytest.mark.skipif(not oj.node_available(), reason="node is required for the oracle")

LAM = 532e-9


def disc(n, r):
    i = np.arange(n)
    X, Y = np.meshgrid(i - n / 2, i - n / 2)
    return (np.hypot(X, Y) <= r).astype(float)


def soft_disc(n, dx, radius, soft):
    c = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(c, c)
    r = np.hypot(X, Y)
    edge = 0.5 * (1 + np.cos(np.pi * (r - radius + soft) / soft))
    return np.where(r <= radius - soft, 1.0, np.where(r >= radius, 0.0, edge))


def asm_np(u, dx, lam, z):
    """Angular spectrum without band limit, evanescent waves removed:
    centred spectrum (numpy shifts) times exp(i 2 pi z sqrt(1/lam^2 - f^2))."""
    n = u.shape[0]
    f = (np.arange(n) - n / 2) / (n * dx)
    FX, FY = np.meshgrid(f, f)
    arg = 1 / lam ** 2 - FX ** 2 - FY ** 2
    H = np.where(arg >= 0, np.exp(2j * np.pi * z * np.sqrt(np.maximum(arg, 0))), 0)
    U = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(u)))
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(U * H)))


def cdft(x):
    """Unitary DFT in centred index space."""
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x))) / np.sqrt(x.size)


def icdft(y):
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(y))) * np.sqrt(y.size)


def phasor_dist(a, b, w):
    """amplitude-weighted RMS distance between exp(i a) and exp(i b)"""
    return float(np.sqrt(np.sum(w * np.abs(np.exp(1j * a) - np.exp(1j * b)) ** 2) / np.sum(w)))


def design_case(n, dx, seed):
    ill = soft_disc(n, dx, 0.45 * n * dx, 2 * dx)
    t = np.sqrt(disc(n, 0.26 * n) + 0.6 * disc(n, 0.1 * n))
    t *= np.sqrt(np.sum(ill ** 2) / np.sum(t ** 2))
    roi = 0.4 * disc(n, 0.33 * n) + 0.6 * disc(n, 0.29 * n)
    rng = np.random.default_rng(seed)
    phi0 = rng.uniform(-np.pi, np.pi, (n, n))
    return ill, t, roi, phi0


# ------------------------------------------------------------ HIO ---------
def hio_np(ill, t, roi, phi, dx, z, fb):
    """One image-plane GS step with HIO feedback fb: inside the ROI the
    amplitude becomes t + fb (a - t) (t where a = 0), outside it stays a."""
    U = asm_np(ill * np.exp(1j * phi), dx, LAM, z)
    a = np.abs(U)
    inside = roi > 0
    na = np.where(inside, np.where(a > 0, t + fb * (a - t), t), a)
    Up = np.where(a > 0, U * na / np.where(a > 0, a, 1), U)
    return np.angle(asm_np(Up, dx, LAM, -z)), a


@pytest.mark.parametrize("fb", [0.7, 0.35])
def test_hio_step_closed_form(fb):
    n, dx, z = 64, 8e-6, 0.03
    ill, t, roi, phi0 = design_case(n, dx, 3)
    args = {"mode": "image-plane", "algorithm": "gs", "nx": n, "ny": n, "dx": dx, "dy": dx,
            "lambda": LAM, "z": z, "iterations": 1, "feedback": fb, "escape": 0}
    out, r = oj.run("gs", args, {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
    phi1, a0 = hio_np(ill, t, roi, phi0, dx, z, fb)
    assert phasor_dist(out["phase"].reshape(n, n), phi1, ill) < 1e-10
    # plain GS from the same start lands elsewhere: the feedback term matters
    phi_gs, _ = hio_np(ill, t, roi, phi0, dx, z, 0.0)
    assert phasor_dist(phi_gs, phi1, ill) > 1e-3
    m = roi > 0
    e0 = np.sqrt(np.sum(((a0 - t) * roi)[m] ** 2) / np.sum(t[m] ** 2))
    assert abs(r["error"][0] - e0) < 1e-12


def test_hio_escape_period():
    """escape = 1: iteration 0 is plain GS, iteration 1 uses the feedback."""
    n, dx, z, fb = 64, 8e-6, 0.03, 0.8
    ill, t, roi, phi0 = design_case(n, dx, 4)
    args = {"mode": "image-plane", "algorithm": "gs", "nx": n, "ny": n, "dx": dx, "dy": dx,
            "lambda": LAM, "z": z, "iterations": 2, "feedback": fb, "escape": 1}
    out, _ = oj.run("gs", args, {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
    phi1, _ = hio_np(ill, t, roi, phi0, dx, z, 0.0)
    phi2, _ = hio_np(ill, t, roi, phi1, dx, z, fb)
    assert phasor_dist(out["phase"].reshape(n, n), phi2, ill) < 1e-10


# ------------------------------------------------- far-field WGS ----------
def far_wgs_np(ill, t, roi, phi0, iters, beta=0.9):
    """Far-field WGS written from its definition: per iteration measure
    Y = F(ill e^{i phi}), project (|Y| -> t w inside the ROI, Y -> t w where
    |Y| = 0), phi' = arg F^-1; then update the weights from |Y|.  Returns per
    iteration (rmse of |Y| against t over the ROI, phi', weights after the
    update)."""
    m = roi > 0
    w = m.astype(float)
    phi = phi0.copy()
    hist = []
    t2 = np.sum(t[m] ** 2)
    for _ in range(iters):
        Y = cdft(ill * np.exp(1j * phi))
        a = np.abs(Y)
        rmse = np.sqrt(np.sum((a - t)[m] ** 2) / t2)
        tt = np.where(m, t * w, a)
        Yp = np.where(a > 0, Y * tt / np.where(a > 0, a, 1), tt)
        phi = np.angle(icdft(Yp))
        ratio = np.where(a > 1e-30, t / np.where(a > 1e-30, a, 1), 2.0)
        wn = np.clip(w * np.clip(ratio, 0.25, 4) ** beta, 1e-4, 1e4)
        wn = np.where(m, wn, w)
        w = wn * np.sqrt(t2 / np.sum(wn[m] ** 2 * t[m] ** 2))
        hist.append((rmse, phi.copy(), w.copy()))
    return hist


def ff_args(n, dx, iters):
    return {"mode": "far-field", "algorithm": "wgs", "nx": n, "ny": n, "dx": dx, "dy": dx,
            "lambda": LAM, "z": 0.03, "iterations": iters, "beta": 0.9}


def test_far_field_wgs_weights():
    n, dx = 32, 8e-6
    ill, t, roi, phi0 = design_case(n, dx, 5)
    hist = far_wgs_np(ill, t, roi, phi0, 2)
    # the first measurement has amplitude ratios beyond both clamp bounds, so
    # the order of power and clamp shows in the weights
    a0 = np.abs(cdft(ill * np.exp(1j * phi0)))
    ratio = t[roi > 0] / a0[roi > 0]
    assert np.sum(ratio > 4) > 5 and np.sum(ratio < 0.25) > 5
    for it in (1, 2):
        out, r = oj.run("gs", ff_args(n, dx, it), {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
        w_ref = hist[it - 1][2].ravel()
        assert np.max(np.abs(out["weights"] - w_ref)) < 1e-9 * np.max(w_ref), it
        assert abs(r["rmse"] - min(h[0] for h in hist[:it])) < 1e-12
    # renormalised: sum w^2 t^2 = sum t^2 over the ROI
    m = (roi > 0).ravel()
    tv = t.ravel()[m]
    assert abs(np.sum(out["weights"][m] ** 2 * tv ** 2) / np.sum(tv ** 2) - 1) < 1e-12


def test_best_iterate_off_by_one():
    """The phase returned with the best rmse is the one its iteration's
    projection produced (the start of the next iteration), not the phase the
    iteration was measured with."""
    n, dx, K = 32, 8e-6, 9
    ill, t, roi, phi0 = design_case(n, dx, 6)
    hist = far_wgs_np(ill, t, roi, phi0, K)
    rm = np.array([h[0] for h in hist])
    kb = int(np.argmin(rm))
    out, r = oj.run("gs", ff_args(n, dx, K), {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
    assert np.allclose(r["error"], np.minimum.accumulate(rm), rtol=0, atol=1e-10)
    got = out["phase"].reshape(n, n)
    measured_with = phi0 if kb == 0 else hist[kb - 1][1]
    assert phasor_dist(got, hist[kb][1], ill) < 1e-8
    assert phasor_dist(got, measured_with, ill) > 1e-3


def test_far_field_zero_amplitude_branch():
    """Zero illumination: |Y| = 0 everywhere, so the projection sets the
    spectrum to the target amplitude (real) inside the ROI and keeps 0
    outside; the phase is arg F^-1 of that spectrum."""
    n, dx = 32, 8e-6
    _, t, roi, phi0 = design_case(n, dx, 7)
    ill = np.zeros((n, n))
    args = dict(ff_args(n, dx, 1), algorithm="gs")
    out, r = oj.run("gs", args, {"illum": ill, "target": t, "roi": roi, "phase0": phi0})
    x = icdft(np.where(roi > 0, t, 0.0))
    assert phasor_dist(out["phase"].reshape(n, n), np.angle(x), np.abs(x)) < 1e-12
    assert r["error"][0] == 1.0 and r["efficiency"][0] == 0.0


# ---------------------------------------------- band-limit soft edge ------
def test_band_limit_soft_window():
    n, dx, z, soft = 256, 8e-6, 0.2, 0.3
    out, _ = oj.run("transfer", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z,
                                 "bandLimit": True, "softEdge": soft})
    H = np.abs(out["re"] + 1j * out["im"]).reshape(n, n)
    df = 1 / (n * dx)
    lim = min(1 / (LAM * np.sqrt((2 * df * z) ** 2 + 1)), 1 / (2 * dx))
    inner = lim * (1 - soft)
    v = np.abs((np.arange(n) - n / 2) * df)
    row = H[n // 2]  # fy = 0: the y window is 1
    assert np.all(np.abs(row[v <= inner] - 1) < 1e-14) and np.sum(v <= inner) > 20
    assert np.all(row[v >= lim] == 0)
    tr = (v > inner) & (v < lim)
    assert np.sum(tr) >= 8
    pos = tr & (np.arange(n) > n // 2)  # positive frequencies, increasing
    assert np.all(np.diff(row[pos]) < 0)  # strictly decreasing across the edge
    # raised cosine from 1 at inner to 0 at lim (1/2 at the midpoint)
    edge = 0.5 * (1 + np.cos(np.pi * (v - inner) / (lim - inner)))
    assert np.max(np.abs(row[tr] - edge[tr])) < 1e-13
    # separable: |H| = w(fx) w(fy) inside the propagating disc
    wv = np.where(v <= inner, 1.0, np.where(v >= lim, 0.0, edge))
    assert np.max(np.abs(H - np.outer(wv, wv))) < 1e-13


# --------------------------------------------------------- SSIM -----------
def ssim_case(nx, ny, seed, affine):
    rng = np.random.default_rng(seed)
    wy, wx = ny // 8, nx // 8
    if affine:
        t = rng.uniform(0.5, 1.5, (ny, nx))
        al = rng.uniform(-1.5, 2.0, (wy, wx))
        be = rng.uniform(0.0, 1.0, (wy, wx))
        p = t * np.kron(al, np.ones((8, 8)))[:ny, :nx] + np.kron(be, np.ones((8, 8)))[:ny, :nx]
    else:
        t = np.kron(rng.uniform(0.2, 2.0, (wy, wx)), np.ones((8, 8)))
        p = np.kron(rng.uniform(0.2, 2.0, (wy, wx)), np.ones((8, 8)))
    return p, t


def ssim_windows(p, t, L):
    """per-window SSIM from the definition (population moments), and the
    target-mean weights, for the complete 8x8 windows"""
    C1, C2 = (0.01 * L) ** 2, (0.03 * L) ** 2
    vals, wts = [], []
    for y0 in range(0, p.shape[0] - 7, 8):
        for x0 in range(0, p.shape[1] - 7, 8):
            a = p[y0:y0 + 8, x0:x0 + 8]
            b = t[y0:y0 + 8, x0:x0 + 8]
            ma, mb = a.mean(), b.mean()
            va, vb = a.var(), b.var()
            cab = np.mean((a - ma) * (b - mb))
            vals.append((2 * ma * mb + C1) * (2 * cab + C2) / ((ma ** 2 + mb ** 2 + C1) * (va + vb + C2)))
            wts.append(mb)
    return np.array(vals), np.array(wts)


def oracle_ssim(p, t, weight):
    ny, nx = p.shape
    _, r = oj.run("metrics", {"nx": nx, "ny": ny, "weight": weight},
                  {"plane": p.ravel(), "target": t.ravel(), "mask": np.ones(p.size)})
    return r["ssim"]


@pytest.mark.parametrize("weight", [None, "target"])
def test_ssim_piecewise_constant(weight):
    p, t = ssim_case(32, 32, 1, affine=False)
    L = max(p.max(), t.max())
    C1 = (0.01 * L) ** 2
    pw, tw = p[::8, ::8], t[::8, ::8]
    s = (2 * pw * tw + C1) / (pw ** 2 + tw ** 2 + C1)  # the variance factor is C2 / C2
    want = np.sum(s * tw) / np.sum(tw) if weight == "target" else s.mean()
    # (the oracle's window variances are differences of sums: ~1e-16 / C2)
    assert abs(oracle_ssim(p, t, weight) - want) < 1e-11


@pytest.mark.parametrize("weight", [None, "target"])
def test_ssim_affine_windows(weight):
    p, t = ssim_case(32, 40, 2, affine=True)
    L = max(p.max(), t.max())
    vals, wts = ssim_windows(p, t, L)
    want = np.sum(vals * wts) / np.sum(wts) if weight == "target" else vals.mean()
    assert abs(oracle_ssim(p, t, weight) - want) < 1e-12
    # identical planes: 1
    assert abs(oracle_ssim(t, t, weight) - 1) < 1e-12


def test_ssim_skips_partial_windows_and_uses_the_peak():
    p, t = ssim_case(32, 16, 3, affine=True)
    L = max(p.max(), t.max())
    vals, wts = ssim_windows(p, t, L)
    # 4 extra columns and 5 extra rows of other values below the peak: no
    # complete window covers them, so only L could change, and it does not
    rng = np.random.default_rng(9)
    P = rng.uniform(0, 0.9 * L, (21, 36))
    T = rng.uniform(0, 0.9 * L, (21, 36))
    P[:16, :32], T[:16, :32] = p, t
    assert abs(oracle_ssim(P, T, "target") - np.sum(vals * wts) / np.sum(wts)) < 1e-12
    # the data range is the larger peak of the two planes
    q = p.copy()
    q[0, 0] = 10 * L
    vq, wq = ssim_windows(q, t, 10 * L)
    assert abs(oracle_ssim(q, t, "target") - np.sum(vq * wq) / np.sum(wq)) < 1e-12


# ------------------------------------- zero order, rms radius, spots ------
def metric_fns(nx, ny, field=None, plane=None, **kw):
    inputs = {}
    if field is not None:
        inputs.update(re=field.real.ravel(), im=field.imag.ravel())
    if plane is not None:
        inputs["plane"] = plane.ravel()
    return oj.run("metric_fns", {"nx": nx, "ny": ny, **kw}, inputs)[1]


def test_zero_order_fraction():
    n = 32
    c = 0.3 - 1.2j
    assert abs(metric_fns(n, n, np.full((n, n), c))["zeroOrder"] - 1) < 1e-12
    j, i = np.mgrid[0:n, 0:n]
    checker = c * (-1.0) ** (i + j)
    assert abs(metric_fns(n, n, checker)["zeroOrder"]) < 1e-30
    rng = np.random.default_rng(4)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 0.5
    want = abs(u.sum()) ** 2 / (u.size * np.sum(np.abs(u) ** 2))
    assert abs(metric_fns(n, n, u)["zeroOrder"] - want) < 1e-12


def test_rms_radius_gaussian():
    """I = exp(-2 r^2 / w^2) has rms widths w/2 per axis: r = w / sqrt(2)."""
    n, dx, w = 256, 1e-6, 18e-6
    c = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(c, c)
    g = np.exp(-2 * (X ** 2 + (Y / 1.5) ** 2) / w ** 2)  # y width 1.5 w
    r = metric_fns(n, n, plane=g, dx=dx, dy=dx)["rmsRadius"]
    assert abs(r["x"] / (w / 2) - 1) < 1e-10
    assert abs(r["y"] / (1.5 * w / 2) - 1) < 1e-10
    assert abs(r["r"] - np.hypot(w / 2, 1.5 * w / 2)) < 1e-10 * w
    # about a given centre: the same beam moved there
    cx, cy = 12 * dx, -7 * dx
    g2 = np.exp(-2 * ((X - cx) ** 2 + ((Y - cy) / 1.5) ** 2) / w ** 2)
    r2 = metric_fns(n, n, plane=g2, dx=dx, dy=dx, cx=cx, cy=cy)["rmsRadius"]
    assert abs(r2["r"] / r["r"] - 1) < 1e-10


def test_spot_uniformity():
    n = 64
    cols = rows = 4
    bg = 0.05  # below minPeak * peak (0.1 and 0.15 here): never counted
    plane = np.full((n, n), bg)
    cells = n // cols
    for cj in range(rows):
        for ci in range(cols):
            y, x = cj * cells + 6, ci * cells + 7
            plane[y:y + 3, x:x + 3] = 2.0
    r = metric_fns(n, n, plane=plane, dx=1.0, spots={"cols": cols, "rows": rows})
    assert r["spotUniformity"] == 1.0 and np.allclose(r["spotCells"], 18.0, rtol=0, atol=1e-12)
    # two spots, 9 and 27 in power, the rest dark: 2 min / (max + min) = 0.5
    plane = np.full((n, n), bg)
    plane[5:8, 5:8] = 1.0
    plane[40:43, 50:53] = 3.0
    r = metric_fns(n, n, plane=plane, dx=1.0, spots={"cols": cols, "rows": rows})
    assert r["spotUniformity"] == 0.5
    assert sorted(v for v in r["spotCells"] if v > 0) == [9.0, 27.0]


# ------------------------------------------------- illumination -----------
def test_illumination_profiles():
    n, dx = 128, 8e-6
    radius = n * dx / 2 * 0.95
    out, _ = oj.run("illumination", {"type": "tophat", "nx": n, "ny": n, "dx": dx, "dy": dx})
    top = soft_disc(n, dx, radius, 2 * dx)
    assert np.max(np.abs(out["profile"] - top.ravel())) < 1e-14
    # the edge: 1 inside radius - soft, 1/2 half-way, 0 from the radius out
    c = (np.arange(n) - n / 2) * dx
    R = np.hypot(*np.meshgrid(c, c)).ravel()
    p = out["profile"]
    assert np.all(p[R <= radius - 2 * dx] == 1) and np.all(p[R >= radius] == 0)
    ring = (R > radius - 2 * dx) & (R < radius)
    assert np.all((p[ring] > 0) & (p[ring] < 1))
    out, _ = oj.run("illumination", {"type": "gaussian", "nx": n, "ny": n, "dx": dx, "dy": dx,
                                     "opts": {"waist": 0.3 * n * dx}})
    X, Y = np.meshgrid(c, c)
    g = top * np.exp(-(X ** 2 + Y ** 2) / (0.3 * n * dx) ** 2)
    assert np.max(np.abs(out["profile"] - g.ravel())) < 1e-14


# -------------------------------------- ComplexField sampling / resample ---
def sample_image(nx, ny, img, scale=1.0, fill=0.0):
    h, w = img.shape
    out, _ = oj.run("sample_image", {"nx": nx, "ny": ny, "width": w, "height": h, "scale": scale,
                                     "fill": fill}, {"img": img.ravel()})
    assert np.all(out["im"] == 0)
    return out["re"].reshape(ny, nx)


def test_sample_image_closed_forms():
    rng = np.random.default_rng(8)
    img = rng.uniform(0, 1, (16, 32))
    # an image of the grid's own size comes back unchanged
    assert np.max(np.abs(sample_image(32, 16, img) - img)) < 1e-15
    # half the resolution: each sample is the mean of its 2 x 2 block
    want = img.reshape(8, 2, 16, 2).mean(axis=(1, 3))
    assert np.max(np.abs(sample_image(16, 8, img) - want)) < 1e-14
    # inset by scale: the samples integrate the image exactly once
    for nx, ny, scale in ((32, 16, 0.5), (64, 32, 0.75), (16, 16, 0.6), (128, 64, 0.9)):
        g = sample_image(nx, ny, img, scale, 0.0)
        box = (32 / (nx * scale)) * (16 / (ny * scale))
        assert abs(np.sum(g) * box / np.sum(img) - 1) < 1e-12, (nx, ny, scale)
        # a fill only enters samples whose box leaves the image
        gf = sample_image(nx, ny, img, scale, 0.7)
        c = (np.arange(nx) + 0.5) / nx
        cx = (c - (1 - scale) / 2) * 32 / scale
        cy = ((np.arange(ny) + 0.5) / ny - (1 - scale) / 2) * 16 / scale
        bw, bh = 32 / (nx * scale), 16 / (ny * scale)
        ix = (cx - bw / 2 >= 0) & (cx + bw / 2 <= 32)
        iy = (cy - bh / 2 >= 0) & (cy + bh / 2 <= 16)
        inside = np.outer(iy, ix)
        assert inside.sum() > 0 and np.max(np.abs(gf[inside] - g[inside])) < 1e-13  # (rounding of the box area)
        # outside, the reference weights the fill by area - (hix - lox)(hiy - loy),
        # whose factors go negative beyond the image (reproduced, not corrected)
        assert np.any(gf[~inside] != g[~inside])


def test_resample_bilinear_reproduces_affine_planes():
    n, dx = 32, 4e-6
    j, i = np.mgrid[0:n, 0:n].astype(float)
    re = 0.3 + 2.0 * i - 0.5 * j
    im = -1.0 + 0.25 * i + 1.5 * j
    for n2, dx2 in ((64, 2.5e-6), (16, 6e-6), (128, 1.5e-6)):
        out, _ = oj.run("field_resample", {"nx": n, "ny": n, "dx": dx, "dy": dx, "nx2": n2, "ny2": n2,
                                           "dx2": dx2, "dy2": dx2}, {"re": re.ravel(), "im": im.ravel()})
        g = (out["re"] + 1j * out["im"]).reshape(n2, n2)
        f = (np.arange(n2) - n2 / 2) * dx2 / dx + n / 2 - 0.5  # fractional source index
        fy, fx = np.meshgrid(f, f, indexing="ij")
        inside = (fx >= 0) & (fx <= n - 1) & (fy >= 0) & (fy <= n - 1)
        want = (0.3 + 2.0 * fx - 0.5 * fy) + 1j * (-1.0 + 0.25 * fx + 1.5 * fy)
        assert np.max(np.abs(g[inside] - want[inside])) < 1e-11
        outside = (fx <= -1) | (fx >= n) | (fy <= -1) | (fy >= n)
        assert np.all(g[outside] == 0)
        assert inside.sum() > 0.2 * n2 * n2

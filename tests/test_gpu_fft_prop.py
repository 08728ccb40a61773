"""HIP fft2d and propagators vs the FP64 oracle (ref: src/fft.js, src/propagate.js).

Gates: FP32 <= 1e-5 relative L2 against the FP64 path, FP64 check mode
<= 1e-12; FP16/BF16 storage (FP32 arithmetic) is reported against looser,
storage-rounding tolerances stated per test.  At the full batch size the
size-independent properties are checked (round trip, Parseval, linearity).
"""
import numpy as np
import pytest

import oracle_js as oj
from diffractiveopticalelement_amd import native as nat

pytestmark = pytest.mark.gpu

LAM = 532e-9
TOL = {"f64": 1e-12, "f32": 1e-5, "f16": 2e-3, "bf16": 1.5e-2}


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rand_field(rng, ny, nx):
    return rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))


def gpu_fft(ctx, u, dtype, inverse=False):
    f = ctx.upload(u.real, u.imag, dtype)
    ctx.fft2d(f, inverse)
    re, im = f.download()
    f.free()
    return re + 1j * im


@pytest.mark.parametrize("dtype", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("nx,ny", [(2, 2), (4, 8), (8, 8), (16, 2), (64, 64), (128, 32), (256, 256)])
def test_fft2d_vs_oracle(ctx, dtype, nx, ny):
    print(f"fft2d {dtype} {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny)
    u = rand_field(rng, ny, nx)
    for inverse in (False, True):
        out, _ = oj.run("fft2d", {"nx": nx, "ny": ny, "inverse": inverse},
                        {"re": u.real.ravel(), "im": u.imag.ravel()})
        ref = (out["re"] + 1j * out["im"]).reshape(ny, nx)
        got = gpu_fft(ctx, u, dtype, inverse)[0]
        assert rel(got, ref) < TOL[dtype], (dtype, nx, ny, inverse)


@pytest.mark.parametrize("nx,ny", [(512, 512), (1024, 1024), (2048, 2048), (4096, 4096), (8192, 16), (16, 8192)])
def test_fft2d_large_vs_numpy(ctx, nx, ny):
    # numpy's FFT stands in for the oracle at sizes its O(N log N) JS loop is slow at
    print(f"fft2d large {nx}x{ny}", flush=True)
    rng = np.random.default_rng(1)
    u = rand_field(rng, ny, nx)
    ref = np.fft.fft2(u)
    for dtype in ("f64", "f32"):
        assert rel(gpu_fft(ctx, u, dtype)[0], ref) < TOL[dtype], dtype
    back = gpu_fft(ctx, ref, "f32", inverse=True)[0]
    assert rel(back, u) < TOL["f32"]


def test_fft2d_host_dropin_and_errors(ctx):
    rng = np.random.default_rng(3)
    u = rand_field(rng, 16, 32)
    re = np.ascontiguousarray(u.real.ravel())
    im = np.ascontiguousarray(u.imag.ravel())
    ctx.fft2d_host(re, im, 32, 16, False, "f64")
    assert rel((re + 1j * im).reshape(16, 32), np.fft.fft2(u)) < 1e-13
    # the reference's length error (ref: fft.js:149-151) surfaces as a status
    with pytest.raises(nat.DoeError) as ei:
        ctx.fft2d_host(np.zeros(10), np.zeros(10), 4, 4, False, "f64")
    assert "does not match 4x4" in str(ei.value)
    # any grid can be stored; a transform needs power-of-two sides (ref: field.js:30-32)
    f = ctx.field(12, 16)
    with pytest.raises(nat.DoeError, match="powers of two"):
        ctx.fft2d(f)
    f.free()
    with pytest.raises(nat.DoeError, match="65536 rows"):
        ctx.field(0, 16)


def test_fft2d_batch_linearity_and_parseval(ctx):
    # full-size batch: 16 x 2048^2 FP32 (the propagation workload's grid)
    print("fft2d batch 16x2048^2", flush=True)
    B, N = 16, 2048
    rng = np.random.default_rng(7)
    u = rng.standard_normal((B, N, N)).astype(np.float64)
    v = rng.standard_normal((B, N, N)).astype(np.float64)
    f = ctx.upload(u, v, "f32")
    ctx.fft2d(f)
    re, im = f.download()
    X = re + 1j * im
    # Parseval per field and a spot check of two fields against numpy
    e_time = np.sum(u * u + v * v, axis=(1, 2))
    e_freq = np.sum(np.abs(X) ** 2, axis=(1, 2)) / (N * N)
    assert np.max(np.abs(e_freq / e_time - 1)) < 1e-6
    for b in (0, B - 1):
        assert rel(X[b], np.fft.fft2(u[b] + 1j * v[b])) < 1e-5
    ctx.fft2d(f, inverse=True)
    re, im = f.download()
    assert rel(re + 1j * im, u + 1j * v) < 1e-6
    f.free()


# ---------------------------------------------------------------- ASM ----
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n,dx,z,band,evan", [
    (64, 8e-6, 25e-3, False, True),
    (128, 8e-6, 0.05, True, True),
    (64, 1e-6, 0.01, True, False),   # evanescent waves kept as a phase
    (256, 8e-6, 0.03, True, True),
])
def test_asm_vs_oracle(ctx, dtype, n, dx, z, band, evan):
    print(f"asm {dtype} {n} z={z} band={band} evan={evan}", flush=True)
    rng = np.random.default_rng(n)
    u = rand_field(rng, n, n)
    out, _ = oj.run("asm", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z,
                            "bandLimit": band, "evanescent": evan, "softEdge": 0.02},
                    {"re": u.real.ravel(), "im": u.imag.ravel()})
    ref = (out["re"] + 1j * out["im"]).reshape(n, n)
    fi = ctx.upload(u.real, u.imag, dtype)
    fo = ctx.field(n, n, 1, dtype)
    ctx.asm(fi, fo, dx, dx, LAM, z, band, evan, 0.02)
    re, im = fo.download()
    assert rel(re[0] + 1j * im[0], ref) < TOL[dtype]
    fi.free()
    fo.free()


def test_asm_multi_uneven_sizes(ctx):
    # batches of different sides (one launch per pass for all of them): each
    # equals doe_asm on its own batch (FP32 rounding apart where the line
    # transforms are split differently), and the small ones the FP64 oracle
    sides = [256, 64, 1024, 512, 128, 2048]
    batch = [2, 3, 1, 2, 1, 1]
    dxs = [8e-6, 8e-6, 4e-6, 6e-6, 8e-6, 4e-6]
    print(f"asm multi sides={sides} batches={batch}", flush=True)
    rng = np.random.default_rng(77)
    fins, fouts, zs, us = [], [], [], []
    for n, b in zip(sides, batch):
        u = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
        us.append(u)
        fins.append(ctx.upload(u.real, u.imag, "f32"))
        fouts.append(ctx.field(n, n, b, "f32"))
        zs.append(rng.uniform(0.01, 0.08, b))
    ctx.asm_multi(fins, fouts, dxs, dxs, LAM, zs, band_limit=True)
    for k, (n, b) in enumerate(zip(sides, batch)):
        re, im = fouts[k].download()
        got = re + 1j * im
        g1 = ctx.field(n, n, b, "f32")
        ctx.asm(fins[k], g1, dxs[k], dxs[k], LAM, zs[k], band_limit=True)
        r1, i1 = g1.download()
        assert rel(got, r1 + 1j * i1) < 1e-6, (n, rel(got, r1 + 1j * i1))
        g1.free()
        if n <= 128:
            out, _ = oj.run("asm", {"nx": n, "ny": n, "dx": dxs[k], "dy": dxs[k], "lambda": LAM,
                                    "z": float(zs[k][0]), "bandLimit": True, "evanescent": True,
                                    "softEdge": 0.02},
                            {"re": us[k][0].real.ravel(), "im": us[k][0].imag.ravel()})
            ref = (out["re"] + 1j * out["im"]).reshape(n, n)
            assert rel(got[0], ref) < TOL["f32"], n
    # in place, and an unsupported side is refused
    ctx.asm_multi(fins[:2], fins[:2], dxs[:2], dxs[:2], LAM, zs[:2], band_limit=True)
    re, im = fins[1].download()
    r0, i0 = fouts[1].download()
    assert rel(re + 1j * im, r0 + 1j * i0) < 1e-6
    big = ctx.field(4096, 4096, 1, "f32")
    with pytest.raises(nat.DoeError):
        ctx.asm_multi([big], [big], [4e-6], [4e-6], LAM, [[0.05]])
    big.free()
    for f in fins + fouts:
        f.free()


def test_asm_transfer_table_vs_oracle(ctx):
    n, dx, z = 128, 8e-6, 0.05
    out, res = oj.run("transfer", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z,
                                   "bandLimit": True, "softEdge": 0.01})
    hr, hi, frac = ctx.asm_transfer(n, n, dx, dx, LAM, z, True, True, 0.01, "f64")
    assert np.max(np.abs((hr + 1j * hi).ravel() - (out["re"] + 1j * out["im"]))) < 1e-12
    assert frac == res["propagatingFraction"]


def test_asm_zero_distance_is_identity(ctx):
    rng = np.random.default_rng(4)
    u = rand_field(rng, 32, 32)
    fi = ctx.upload(u.real, u.imag, "f64")
    fo = ctx.field(32, 32, 1, "f64")
    ctx.asm(fi, fo, 8e-6, 8e-6, LAM, 0.0)
    re, im = fo.download()
    assert np.array_equal(re[0] + 1j * im[0], u)


def test_asm_batch_per_field_distances(ctx):
    # 16 x 2048^2 FP32 band-limited ASM forward then inverse with per-field z
    print("asm batch 16x2048^2", flush=True)
    B, N, dx, lam = 16, 2048, 4e-6, 633e-9
    rng = np.random.default_rng(11)
    zs = rng.uniform(0.01, 0.1, B)
    c = (np.arange(N) - N / 2) * dx
    w = 0.3 * N * dx
    amp = np.exp(-(c[None, :] ** 2 + c[:, None] ** 2) / w ** 2)
    ph = rng.uniform(0, 2 * np.pi, (B, N, N))
    u = amp[None] * np.exp(1j * ph)
    f = ctx.upload(u.real, u.imag, "f32")
    g = ctx.field(N, N, B, "f32")
    # unitary (no band limit): power is conserved per field
    ctx.asm(f, g, dx, dx, lam, zs, band_limit=False)
    re, im = g.download()
    p_in = np.sum(np.abs(u) ** 2, axis=(1, 2))
    p_out = np.sum(re ** 2 + im ** 2, axis=(1, 2))
    assert np.max(np.abs(p_out / p_in - 1)) < 1e-5
    ctx.asm(g, g, dx, dx, lam, -zs, band_limit=False)
    re, im = g.download()
    assert rel(re + 1j * im, u) < 1e-5
    # field 0 against numpy's independent pipeline (FP64)
    f1 = (np.arange(N) - N / 2) / (N * dx)
    FX, FY = np.meshgrid(f1, f1)
    arg = 1 / lam ** 2 - FX ** 2 - FY ** 2
    H = np.where(arg >= 0, np.exp(2j * np.pi * zs[0] * np.sqrt(np.abs(arg))), 0)
    ref = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(
        np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(u[0]))) * H)))
    ctx.asm(f, g, dx, dx, lam, zs, band_limit=False)
    re, im = g.download()
    assert rel(re[0] + 1j * im[0], ref) < 1e-5
    f.free()
    g.free()


def test_chunked_batches_match_whole_batch(ctx):
    # a batch run chunk by chunk (per-field parameters offset per chunk) is
    # bit-identical to the whole batch in one launch per pass
    print("chunked batch 5x512^2", flush=True)
    B, N, dx, lam = 5, 512, 8e-6, 633e-9
    rng = np.random.default_rng(12)
    zs = rng.uniform(0.01, 0.1, B)
    u = rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))
    old = ctx.chunk_bytes()
    outs = {}
    try:
        for chunk in (0, 2 * N * N * 8):  # whole batch; chunks of 2 fields (2, 2, 1)
            ctx.set_chunk_bytes(chunk)
            f = ctx.upload(u.real, u.imag, "f32")
            g = ctx.field(N, N, B, "f32")
            res = []
            ctx.fft2d(f)
            res.append(f.download())
            ctx.fft2d(f, inverse=True)
            res.append(f.download())
            ctx.asm(f, g, dx, dx, lam, zs)
            res.append(g.download())
            ctx.asm(f, g, dx, dx, lam, zs, band_limit=False)
            res.append(g.download())
            ctx.fresnel(f, g, dx, dx, lam, zs)
            res.append(g.download())
            ctx.fresnel(g, g, dx, dx, lam, zs, inverse=True, in_dx=lam * zs[0] / (N * dx),
                        in_dy=lam * zs[0] / (N * dx))
            res.append(g.download())
            outs[chunk] = res
            f.free()
            g.free()
    finally:
        ctx.set_chunk_bytes(old)
    a, b = outs.values()
    for (ra, ia), (rb, ib) in zip(a, b):
        assert np.array_equal(ra, rb) and np.array_equal(ia, ib)
    # the last field (last chunk) used its own distance: numpy reference
    f1 = (np.arange(N) - N / 2) / (N * dx)
    FX, FY = np.meshgrid(f1, f1)
    arg = 1 / lam ** 2 - FX ** 2 - FY ** 2
    H = np.where(arg >= 0, np.exp(2j * np.pi * zs[-1] * np.sqrt(np.abs(arg))), 0)
    ref = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(
        np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(u[-1]))) * H)))
    re, im = a[3]
    assert rel(re[-1] + 1j * im[-1], ref) < 1e-5


# ------------------------------------------------------------- Fresnel ----
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n,dx,z", [(8, 20e-6, 50e-3), (128, 8e-6, 30e-3), (256, 4e-6, 0.01)])
def test_fresnel_vs_oracle(ctx, dtype, n, dx, z):
    print(f"fresnel {dtype} {n}", flush=True)
    rng = np.random.default_rng(n + 1)
    u = rand_field(rng, n, n)
    args = {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z}
    out, res = oj.run("fresnel", args, {"re": u.real.ravel(), "im": u.imag.ravel()})
    ref = (out["re"] + 1j * out["im"]).reshape(n, n)
    fi = ctx.upload(u.real, u.imag, dtype)
    fo = ctx.field(n, n, 1, dtype)
    odx, ody = ctx.fresnel(fi, fo, dx, dx, LAM, z)
    assert odx[0] == res["dx"] and ody[0] == res["dy"]
    re, im = fo.download()
    assert rel(re[0] + 1j * im[0], ref) < TOL[dtype]
    # inverse against the oracle's inverse of the same input
    back, _ = oj.run("fresnel_inv", {**args, "in_dx": res["dx"], "in_dy": res["dy"]}, out)
    fb = ctx.field(n, n, 1, dtype)
    ctx.fresnel(fo, fb, dx, dx, LAM, z, inverse=True, in_dx=odx[0], in_dy=ody[0])
    br, bi = fb.download()
    assert rel(br[0] + 1j * bi[0], (back["re"] + 1j * back["im"]).reshape(n, n)) < TOL[dtype]
    # round trip: the reference's own gate (ref test/propagate.test.js:137-146)
    assert rel(br[0] + 1j * bi[0], u) < (1e-8 if dtype == "f64" else 1e-5)
    for x in (fi, fo, fb):
        x.free()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_lens_vs_oracle(ctx, dtype):
    n, dx = 128, 8e-6
    rng = np.random.default_rng(5)
    u = rand_field(rng, n, n)
    for f in (0.02, 0.05):
        out, res = oj.run("lens", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "f": f},
                          {"re": u.real.ravel(), "im": u.imag.ravel()})
        fi = ctx.upload(u.real, u.imag, dtype)
        fo = ctx.field(n, n, 1, dtype)
        odx, _ = ctx.lens(fi, fo, dx, dx, LAM, f)
        assert odx[0] == res["dx"]
        re, im = fo.download()
        assert rel((re[0] + 1j * im[0]).ravel(), out["re"] + 1j * out["im"]) < TOL[dtype]
        fi.free()
        fo.free()


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_half_storage_propagators(ctx, dtype):
    # FP16/BF16 storage with FP32 arithmetic: error is dominated by storage
    # rounding (2^-11 / 2^-8 relative per value and pass)
    n, dx, z = 256, 8e-6, 0.03
    rng = np.random.default_rng(9)
    u = rand_field(rng, n, n)
    out, _ = oj.run("asm", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z,
                            "bandLimit": True, "softEdge": 0.02},
                    {"re": u.real.ravel(), "im": u.imag.ravel()})
    fi = ctx.upload(u.real, u.imag, dtype)
    fo = ctx.field(n, n, 1, dtype)
    ctx.asm(fi, fo, dx, dx, LAM, z)
    re, im = fo.download()
    assert rel((re[0] + 1j * im[0]).ravel(), out["re"] + 1j * out["im"]) < 3 * TOL[dtype]
    # Fresnel forward and inverse, and the lens transform, in the same storage
    args = {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "z": z}
    out, res = oj.run("fresnel", args, {"re": u.real.ravel(), "im": u.imag.ravel()})
    odx, ody = ctx.fresnel(fi, fo, dx, dx, LAM, z)
    re, im = fo.download()
    assert rel((re[0] + 1j * im[0]).ravel(), out["re"] + 1j * out["im"]) < 3 * TOL[dtype]
    back, _ = oj.run("fresnel_inv", {**args, "in_dx": res["dx"], "in_dy": res["dy"]}, out)
    fb = ctx.field(n, n, 1, dtype)
    ctx.fresnel(fo, fb, dx, dx, LAM, z, inverse=True, in_dx=odx[0], in_dy=ody[0])
    br, bi = fb.download()
    assert rel((br[0] + 1j * bi[0]).ravel(), back["re"] + 1j * back["im"]) < 3 * TOL[dtype]
    out, _ = oj.run("lens", {"nx": n, "ny": n, "dx": dx, "dy": dx, "lambda": LAM, "f": 0.05},
                    {"re": u.real.ravel(), "im": u.imag.ravel()})
    ctx.lens(fi, fo, dx, dx, LAM, 0.05)
    re, im = fo.download()
    assert rel((re[0] + 1j * im[0]).ravel(), out["re"] + 1j * out["im"]) < 3 * TOL[dtype]
    for x in (fi, fo, fb):
        x.free()


def test_parameter_blocks_in_flight(ctx):
    # Calls with different per-field parameters queue without draining the
    # stream (a ring of parameter slots, each refilled only after the work
    # behind its previous block has ended): 20 asynchronous ASM / Fresnel
    # calls with distinct distances, more than the ring holds, each equal
    # bit for bit to the same call run on its own; then three of them in
    # turn, whose blocks the slots still hold.
    print("parameter blocks in flight", flush=True)
    n, lam, dx = 128, 633e-9, 4e-6
    rng = np.random.default_rng(31)
    u = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    fin = ctx.upload(u.real, u.imag, "f32")
    zs = [(0.01 + 0.003 * i, 0.02 + 0.002 * i) for i in range(20)]
    outs = [ctx.field(n, n, 2, "f32") for _ in zs]
    ctx.set_flags(asynchronous=True)
    try:
        for i, z in enumerate(zs):
            if i % 3 == 2:
                ctx.fresnel(fin, outs[i], dx, dx, lam, z)
            else:
                ctx.asm(fin, outs[i], dx, dx, lam, z, band_limit=bool(i % 2))
        ctx.sync()
    finally:
        ctx.set_flags(asynchronous=False)
    got = [o.download() for o in outs]
    # blocks that repeat (a few parameter sets in turn) are read where a slot
    # still holds them: the same outputs again, in a cycle of three calls
    ctx.set_flags(asynchronous=True)
    try:
        for r in range(12):
            i = r % 3
            o = outs[i]
            o.upload(np.zeros((2, n, n)), np.zeros((2, n, n)))
            if i % 3 == 2:
                ctx.fresnel(fin, o, dx, dx, lam, zs[i])
            else:
                ctx.asm(fin, o, dx, dx, lam, zs[i], band_limit=bool(i % 2))
        ctx.sync()
    finally:
        ctx.set_flags(asynchronous=False)
    for i in range(3):
        again = outs[i].download()
        assert np.array_equal(again[0], got[i][0]) and np.array_equal(again[1], got[i][1]), i
    ref = ctx.field(n, n, 2, "f32")
    for i, z in enumerate(zs):
        if i % 3 == 2:
            ctx.fresnel(fin, ref, dx, dx, lam, z)
        else:
            ctx.asm(fin, ref, dx, dx, lam, z, band_limit=bool(i % 2))
        want = ref.download()
        assert np.array_equal(got[i][0], want[0]) and np.array_equal(got[i][1], want[1]), i
    for o in outs:
        o.free()
    ref.free()
    fin.free()


def test_asm_transfer_fraction_on_repeated_calls(ctx):
    # the propagating fraction is counted afresh by every call, whether or
    # not the call's parameter block is one a previous call uploaded
    n, dx = 64, 8e-6
    first = ctx.asm_transfer(n, n, dx, dx, LAM, 0.05, True, True, 0.01, "f64")[2]
    other = ctx.asm_transfer(n, n, dx, dx, LAM, 0.02, True, True, 0.01, "f64")[2]
    for _ in range(3):
        assert ctx.asm_transfer(n, n, dx, dx, LAM, 0.05, True, True, 0.01, "f64")[2] == first
        assert ctx.asm_transfer(n, n, dx, dx, LAM, 0.02, True, True, 0.01, "f64")[2] == other
    assert 0 < first <= 1 and 0 < other <= 1

"""Per-kernel timing (DOE_CTX_PROFILE): the row and column passes carry their
record's start / stop events on their own dispatches, a reducing pass's
start on the pass and stop on its finalize kernel.  Profiling changes no
result: every output equals the unprofiled call bit for bit.
"""
import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def test_profiled_fft2d_records_every_pass(ctx):
    print("profiled fft2d 1024^2 x 2", flush=True)
    n, b = 1024, 2
    rng = np.random.default_rng(41)
    u = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    f = ctx.upload(u.real, u.imag, "f32")
    ctx.fft2d(f)
    want = f.download()
    f.upload(u.real, u.imag)
    ctx.profile_report()  # clear
    ctx.set_flags(asynchronous=False, profile=True)
    try:
        for _ in range(3):
            f.upload(u.real, u.imag)
            ctx.fft2d(f)
        rep = ctx.profile_report()
    finally:
        ctx.set_flags()
    got = f.download()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    for tag in ("fft_row_fwd", "fft_col_fwd"):
        cnt, ms, pts = rep[tag]
        assert cnt == 3 and pts == 3 * b * n * n, (tag, rep[tag])
        # a pass over 2 x 1024^2 FP32 moves 32 MiB: between 5 us (> 6 TB/s)
        # and 10 ms per launch
        assert 5e-3 < ms / cnt < 10.0, (tag, ms)
    f.free()


def test_profiled_reducing_passes_match_the_graph_replay(ctx):
    # far-field WGS: the column pass reduces (pass + finalize kernel); with
    # profiling the loop runs launch by launch, without it from a captured
    # graph, and both give the same iterates
    print("profiled far-field WGS 256^2, 20 iterations", flush=True)
    n, iters, dx = 256, 20, 8e-6
    c = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(c, c)
    ill = (np.hypot(X, Y) <= 0.45 * n * dx).astype(float)
    t = ((np.abs(X) < 0.2 * n * dx) & (np.abs(Y) < 0.1 * n * dx)).astype(float)
    t *= np.sqrt(np.sum(ill ** 2) / np.sum(t ** 2))
    phi0 = np.random.default_rng(3).uniform(-np.pi, np.pi, (n, n))
    plan = nat.GsPlan(mode=nat.MODE_FAR_FIELD, algorithm=nat.ALG_WGS, iterations=iters, nx=n, ny=n,
                      dx=dx, dy=dx, lambda_=532e-9, z=0.03, beta=0.9, feedback=0.0, escape=0,
                      dtype=nat.DTYPES["f32"])
    want = ctx.gs_run(plan, ill, t, None, phi0)
    ctx.profile_report()
    ctx.set_flags(asynchronous=False, profile=True)
    try:
        got = ctx.gs_run(plan, ill, t, None, phi0)
        rep = ctx.profile_report()
    finally:
        ctx.set_flags()
    for k in ("phase", "error", "efficiency", "weights"):
        assert np.array_equal(got[k], want[k]), k
    cnt, ms, _ = rep["ff_col_project"]
    assert cnt == iters and ms > 0
    cnt, ms, _ = rep["ff_row_phase"]
    assert cnt == iters and ms > 0

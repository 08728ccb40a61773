"""doe_random_phase against the oracle's randomPhase (ref: src/doe.js:65-86).

Gate: FP64 equal bit for bit (xorshift32 states are integers, the phase map
is exact up to its one rounding in the product with pi); FP32 equal to the
FP64 stream rounded once.  Sizes straddle the 64-sample chunk each thread
steps through, and the 1024^2 stream reaches jump distances of 2^20.
"""
import numpy as np
import pytest

from diffractiveopticalelement_amd import native as nat
import oracle_js as oj

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oj.node_available(), reason="node (the oracle's runtime) missing")]


@pytest.fixture(scope="module")
def ctx():
    c = nat.Context(0)
    yield c
    c.close()


def device_phase(ctx, nx, ny, seed, dtype="f64", batch=1):
    f = ctx.field(nx, ny, batch, dtype)
    ctx.random_phase(f, seed)
    re, _ = f.download()
    f.free()
    return re.ravel()


@pytest.mark.parametrize("nx,ny,seed", [
    (1, 1, 0), (63, 1, 1), (64, 1, 12345), (65, 1, 4294967295), (17, 19, 2 ** 31),
    (256, 256, 12345), (1024, 1024, 7), (1000, 3, 0),
])
def test_random_phase_bit_exact(ctx, nx, ny, seed):
    print("randomPhase", nx, ny, seed, flush=True)
    out, _ = oj.run("random_phase", {"nx": nx, "ny": ny, "seed": seed})
    ref = np.asarray(out["phase"])
    np.testing.assert_array_equal(device_phase(ctx, nx, ny, seed), ref)
    np.testing.assert_array_equal(device_phase(ctx, nx, ny, seed, "f32"), ref.astype(np.float32))


def test_random_phase_batch_is_one_stream(ctx):
    out, _ = oj.run("random_phase", {"nx": 8, "ny": 24, "seed": 99})
    np.testing.assert_array_equal(device_phase(ctx, 8, 8, 99, batch=3), np.asarray(out["phase"]))


def test_random_phase_rejects_half_width(ctx):
    f = ctx.field(8, 8, 1, "f16")
    with pytest.raises(nat.DoeError, match="FP64 or FP32"):
        ctx.random_phase(f, 1)
    f.free()

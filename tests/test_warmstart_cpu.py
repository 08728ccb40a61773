"""Host logic of the warm-start comparison: target families, design inputs and
the summary arithmetic (no GPU)."""
import numpy as np
import pytest

from diffractiveopticalelement_amd import warmstart as ws


@pytest.mark.parametrize("kind", ws.FAMILIES)
def test_targets_are_seeded_and_centred(kind):
    n = 128
    a = ws.make_target(kind, n, np.random.default_rng(3))
    b = ws.make_target(kind, n, np.random.default_rng(3))
    assert np.array_equal(a, b) and a.max() > 0
    # blob tails reach up to 2 sigma (<= 0.08 n) past the central half
    m = int(np.ceil(0.08 * n)) if kind == "blobs" else 0
    outside = a.copy()
    outside[n // 4 - m:3 * n // 4 + m, n // 4 - m:3 * n // 4 + m] = 0
    assert outside.max() == 0


def test_design_inputs_power_match_and_roi():
    n = 64
    ill = ws.illumination(n)
    I = ws.make_target("spots", n, np.random.default_rng(0))
    t, roi = ws.design_inputs(I, ill)
    assert np.isclose(np.sum(t ** 2), np.sum(ill ** 2))
    assert set(np.unique(roi)) <= {0.0, 1.0}
    assert np.all(roi[I > 0.05 * I.max()] == 1)  # the support lies inside the ROI


def test_quadratic_phasors_are_unit_and_centred():
    x = np.zeros((1, 16, 16))
    x[0, 8, 8] = 1.0
    z = ws.quadratic_phasors(x, 64)
    assert np.allclose(z[..., 0] ** 2 + z[..., 1] ** 2, 1, atol=1e-6)
    assert np.isclose(z[0, 8, 8, 0], 1) and np.isclose(z[0, 8, 8, 1], 0)  # zero phase at the centroid


def test_summary_arithmetic():
    kinds = ["spots", "blobs", "lines", "glyphs"]
    plain = {"kinds": kinds, "ms_per_design_iteration": 0.02, "inference_ms_batch": 0.0,
             "iterations_to_thr": [100, None, None, None],
             "best_screen_rmse": [0.001, 0.4, 0.02, 0.3]}
    warm = {"kinds": kinds, "ms_per_design_iteration": 0.02, "inference_ms_batch": 0.4,
            "iterations_to_thr": [40, None, None, None],
            "best_screen_rmse": [0.002, 0.6, 0.015, 0.3],
            "common_threshold": [0.01, 0.6, 0.02, 0.3],
            "iterations_to_common": [40, 2, 50, 10],
            "plain_iterations_to_common": [100, 3, 80, 20]}
    s = ws.summarize({"plain": plain, "warm": warm}, 0.01)
    sp = s["warm"]["families"]["spots"]
    assert sp["reach_thr"] == 1 and sp["median_iterations_to_thr"] == 40
    # 40 iterations x 0.02 ms + 0.4 ms inference / 4 designs
    assert sp["median_gpu_ms_to_thr"] == pytest.approx(0.9)
    assert sp["median_plain_over_mode_gpu_time"] == pytest.approx(100 * 0.02 / (40 * 0.02 + 0.1), rel=1e-3)
    assert s["plain"]["families"]["blobs"]["reach_thr"] == 0
    assert s["plain"]["families"]["blobs"]["median_gpu_ms_to_thr"] is None
    assert s["warm"]["families"]["lines"]["median_best_rmse"] == pytest.approx(0.015)

"""Bit-for-bit comparison of FP64 arrays: -0 and +0 differ, every NaN
equals every NaN (np.testing.assert_array_equal treats -0 == +0)."""
import numpy as np


def assert_same_bits(got, want, msg=""):
    g = np.ascontiguousarray(np.asarray(got, dtype=np.float64)).ravel()
    w = np.ascontiguousarray(np.asarray(want, dtype=np.float64)).ravel()
    assert g.shape == w.shape, (g.shape, w.shape, msg)
    gb = g.view(np.uint64)
    wb = w.view(np.uint64)
    bad = (gb != wb) & ~(np.isnan(g) & np.isnan(w))
    if bad.any():
        k = int(np.flatnonzero(bad)[0])
        raise AssertionError(f"{int(bad.sum())} of {g.size} values differ in their bits; first at {k}: "
                             f"{g[k]!r} vs {w[k]!r} {msg}")

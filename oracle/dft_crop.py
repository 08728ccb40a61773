"""ORACLE (test infrastructure only) — naive DFT of a crop of the spectrum,
in FP64 (ref: src/fft.js:208-237 dft2dNaive: X[ky, kx] = sum_y sum_x
x[y, x] exp(-2 pi i (kx x / nx + ky y / ny)), unscaled; the inverse uses
exp(+...) and 1 / (nx ny)).

dft2dNaive is O(N^4); at 4096^2 only a crop of outputs is affordable.  The
double sum separates, X[K] = W_y[K_y, :] @ x @ W_x[:, K_x], so a size^2 crop
costs ny nx size + ny size^2 complex products.  Twiddle angles are reduced
exactly, (n k) mod N, before the FP64 exp, as the reference's direct sum
would need for large n k.

Only tests/ may import it, as the checker.  Pinned by
tests/test_oracle_cpu.py: equal to the JS oracle's dft2dNaive on whole small
grids and to numpy.fft on crops of 512^2 grids.
"""
from __future__ import annotations

import numpy as np


def twiddles(n: int, k: np.ndarray, inverse: bool) -> np.ndarray:
    """exp(-+2 pi i (j k mod n) / n) for j < n (rows) and the given k (columns)."""
    j = np.arange(n, dtype=np.int64)
    m = np.outer(j, np.asarray(k, dtype=np.int64)) % n
    sgn = 1.0 if inverse else -1.0
    return np.exp(sgn * 2j * np.pi * m / n)


def dft2_crop(x: np.ndarray, ky0: int, kx0: int, size_y: int, size_x: int | None = None,
              inverse: bool = False) -> np.ndarray:
    """Spectrum rows ky0..ky0+size_y-1, columns kx0..kx0+size_x-1 of the
    2-D DFT of x (ny x nx, complex), indices taken modulo the grid."""
    ny, nx = x.shape
    size_x = size_y if size_x is None else size_x
    ky = (ky0 + np.arange(size_y)) % ny
    kx = (kx0 + np.arange(size_x)) % nx
    wx = twiddles(nx, kx, inverse)          # nx x size_x
    wy = twiddles(ny, ky, inverse).T        # size_y x ny
    out = wy @ (np.asarray(x, dtype=np.complex128) @ wx)
    return out / (nx * ny) if inverse else out

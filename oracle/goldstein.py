"""Goldstein adaptive phase filter restated in numpy (test infrastructure: the oracle of the GPU
filter; nothing in core_amd imports it).

The reference has two implementations that differ in how overlapping patches are combined:
  * insardev/insardev/utils_goldstein.py:12-154 `goldstein_numpy` (its CPU path): the sum of the
    triangle-weighted patch outputs is divided by the sum of the triangle weights of the valid
    patches covering the pixel; a pixel whose weight sum is < 1e-6 becomes NaN;
  * utils_goldstein.py:157-330 `goldstein_pytorch` (its GPU path, chosen by
    Batch.py:1969-2035 `_goldstein` for any non-CPU device): the same sum divided by the number
    of valid patches covering the pixel (at least 1); a pixel covered by no valid patch is NaN.
`filter(..., mode='count')` follows the GPU path, `mode='weight'` the CPU path.  Everything else
is common: patches of psize_y x psize_x at steps psize/2 from the origin over the image padded at
the bottom/right (utils_goldstein.py:48-57, :224-233), NaN phase and corr read as 0, a patch is
valid when at least int(psize_y psize_x threshold) of its pixels are valid (:77, :93-95),
alpha = 1 - sum(w corr) / sum(w) with the separable triangle window w (:62-68, :89-92), spectrum
times max(|F|, 1e-10)^alpha (:98-105), inverse FFT, times w, invalid patches zeroed (:112-118).
"""
from __future__ import annotations

import numpy as np


def triangle_window(psize_y: int, psize_x: int, dtype=np.float32) -> np.ndarray:
    """The separable triangle window (utils_goldstein.py:62-68): a quadrant rising from 0 to 1 and
    its mirror images."""
    def ramp(p):
        h = p // 2
        return 1.0 - np.abs(np.arange(h) - (p / 2.0 - 1.0)) / (p / 2.0 - 1.0)
    wy, wx = ramp(psize_y), ramp(psize_x)
    q = np.outer(wy, wx)
    return np.block([[q, q[:, ::-1]], [q[::-1, :], q[::-1, ::-1]]]).astype(dtype)


def patch_grid(ny: int, nx: int, psize_y: int, psize_x: int, mode: str = "count"):
    """Patch origins (rows, cols): steps s = psize/2 from 0.  The numpy version pads by
    psize - s and takes floor((n - s) / s) + 1 patches per axis (utils_goldstein.py:53-57,
    :85-86); the torch version pads further so that the fold is exact, i.e. ceil((n - s) / s) + 1
    patches (:224-230): one more row/column when s does not divide n - s, overlapping the last
    n - s mod s image lines (valid only for a low threshold)."""
    sy, sx = psize_y // 2, psize_x // 2
    if ny < sy or nx < sx:
        raise ValueError(f"image {ny}x{nx} smaller than half a {psize_y}x{psize_x} patch")
    if mode == "weight":
        npy, npx = (ny - sy) // sy + 1, (nx - sx) // sx + 1
    else:
        npy, npx = -(-(ny - sy) // sy) + 1, -(-(nx - sx) // sx) + 1
    return np.arange(npy) * sy, np.arange(npx) * sx


def filter(phase, corr, psize_y=32, psize_x=32, threshold=0.5, mode="count", dtype=np.float64):
    """Goldstein filter of a complex (ny, nx) array with correlation corr; returns complex.

    dtype float64 (accuracy oracle) or float32 (the reference's own precision)."""
    if mode not in ("count", "weight"):
        raise ValueError("mode must be 'count' or 'weight'")
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    phase = np.asarray(phase)
    corr = np.asarray(corr)
    ny, nx = phase.shape
    nan_mask = np.isnan(phase)
    w = triangle_window(psize_y, psize_x, dtype)
    wsum = w.sum(dtype=np.float64)
    ys, xs = patch_grid(ny, nx, psize_y, psize_x, mode)
    Hp, Wp = ys[-1] + psize_y, xs[-1] + psize_x
    ph = np.zeros((Hp, Wp), cdt)
    ph[:ny, :nx] = np.where(nan_mask, 0, phase)
    cr = np.zeros((Hp, Wp), dtype)
    cr[:ny, :nx] = np.where(np.isnan(corr), 0, corr)
    valid = np.zeros((Hp, Wp), bool)
    valid[:ny, :nx] = ~nan_mask
    min_valid = int(psize_y * psize_x * threshold)
    # all patches at once (as the reference does): statistics, FFT2, spectral weight, inverse
    from numpy.lib.stride_tricks import sliding_window_view
    sy, sx = psize_y // 2, psize_x // 2
    win = (psize_y, psize_x)
    P = sliding_window_view(ph, win)[::sy, ::sx].reshape(-1, psize_y, psize_x)
    C = sliding_window_view(cr, win)[::sy, ::sx].reshape(-1, psize_y, psize_x)
    V = sliding_window_view(valid, win)[::sy, ::sx].reshape(-1, psize_y, psize_x)
    pvalid = V.sum(axis=(1, 2)) >= min_valid
    alpha = 1.0 - np.einsum('ij,kij->k', w.astype(np.float64), C.astype(np.float64)) / wsum
    F = np.fft.fft2(P, axes=(1, 2))
    F = F * np.power(np.maximum(np.abs(F), 1e-10), alpha[:, None, None])
    Q = (np.fft.ifft2(F, axes=(1, 2)).astype(cdt) * w)
    out = np.zeros((Hp, Wp), cdt)
    cnt = np.zeros((Hp, Wp), np.float64)
    k = 0
    for y0 in ys:  # overlap-add in raster order, as the reference
        for x0 in xs:
            if pvalid[k]:
                sl = (slice(y0, y0 + psize_y), slice(x0, x0 + psize_x))
                out[sl] += Q[k]
                cnt[sl] += w if mode == "weight" else 1.0
            k += 1
    insufficient = cnt < 1e-6
    res = out / (np.maximum(cnt, 1e-10) if mode == "weight" else np.maximum(cnt, 1.0))
    res = res[:ny, :nx].astype(np.complex64 if dtype == np.float32 else np.complex128)
    res[nan_mask | insufficient[:ny, :nx]] = np.nan + 1j * np.nan
    return res

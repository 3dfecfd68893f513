"""NaN-aware Gaussian filter restated in numpy (test infrastructure: the oracle of the GPU filter;
nothing in core_amd imports it).

Follows the reference's GPU path, insardev/insardev/utils_gaussian.py:
  * `nanconv` = nanconvolve2d_gaussian_pytorch (:14-116): 1-D kernels of size 2 ceil(s truncate) + 1,
    exp(-x^2 / 2 s^2) normalised to sum 1 (:51-55); a separable convolution along y then x with
    replicate borders (:65-72); data and weights with NaN read as 0, weight = valid (no weight
    array) or valid x weight (:77-90); numerator / (denominator + 1e-10) where the denominator
    exceeds `threshold` (no weight array) or 1e-10 (with weights), NaN elsewhere (:93-100); complex
    data filtered as real and imaginary parts (:104-107);
  * `gaussian` = gaussian_numpy's GPU branch (:221-445): sigma as a number or (y, x); with
    pixel_sizes the decimation path (:307-316, :353-421) when both scales exceed 1 and
    sigma_y 5.3 dy / resolution > 24: anti-alias filter, mean over scale_y x scale_x blocks of the
    trimmed image, the main filter on the decimated grid, nearest upsampling, replicate padding
    to the input shape.  As in the reference, that path takes the float32 cast of the data (the
    real part of complex data) and ignores the weights.
"""
from __future__ import annotations

import math

import numpy as np

CUTOFF = 5.3


def kernel_1d(s: float, truncate: float = 4.0) -> np.ndarray:
    size = int(2 * math.ceil(s * truncate) + 1)
    x = np.arange(size, dtype=np.float64) - size // 2
    k = np.exp(-0.5 * (x / s) ** 2)
    return k / k.sum()


def _conv_replicate(a: np.ndarray, k: np.ndarray, axis: int) -> np.ndarray:
    """Correlation with the symmetric kernel k along `axis`, replicate borders (F.pad 'replicate'
    then a valid conv2d)."""
    p = len(k) // 2
    pad = [(0, 0), (0, 0)]
    pad[axis] = (p, p)
    ap = np.pad(a, pad, mode="edge")
    out = np.zeros_like(a, dtype=np.float64)
    n = a.shape[axis]
    for t, kt in enumerate(k):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(t, t + n)
        out += kt * ap[tuple(sl)]
    return out


def _nanconv_real(arr, wgt, ky, kx, threshold):
    arr = np.asarray(arr, np.float64)
    valid = (~np.isnan(arr)).astype(np.float64)
    data = np.nan_to_num(arr, nan=0.0)
    w = valid if wgt is None else valid * np.nan_to_num(np.asarray(wgt, np.float64), nan=0.0)
    num = _conv_replicate(_conv_replicate(data * w, ky, 0), kx, 1)
    den = _conv_replicate(_conv_replicate(w, ky, 0), kx, 1)
    lim = threshold if wgt is None else 1e-10
    return np.where(den > lim, num / (den + 1e-10), np.nan)


def _sig2(sigma):
    if not isinstance(sigma, (list, tuple, np.ndarray)):
        return float(sigma), float(sigma)
    return float(sigma[0]), float(sigma[1])


def nanconv(data, weight=None, sigma=None, truncate=4.0, threshold=0.5):
    if sigma is None:
        return data
    sy, sx = _sig2(sigma)
    ky, kx = kernel_1d(sy, truncate), kernel_1d(sx, truncate)
    if np.iscomplexobj(data):
        return _nanconv_real(data.real, weight, ky, kx, threshold) + 1j * _nanconv_real(data.imag, weight, ky, kx,
                                                                                       threshold)
    return _nanconv_real(data, weight, ky, kx, threshold)


def decimation(shape_sigma_y, pixel_sizes, resolution=67.0):
    """(use, scale_y, scale_x, sigma_aa, sigma_dec) of gaussian_numpy's GPU branch."""
    dy, dx = pixel_sizes
    scale_y, scale_x = int(np.round(resolution / dy)), int(np.round(resolution / dx))
    wavelength = shape_sigma_y * CUTOFF * dy
    use = scale_y > 1 and scale_x > 1 and wavelength / resolution > 24
    sigma_aa = (resolution / CUTOFF / dy, resolution / CUTOFF / dx)
    wl_dec = math.sqrt(max(wavelength ** 2 - resolution ** 2, 0.0))
    sigma_dec = (wl_dec / CUTOFF / (dy * scale_y), wl_dec / CUTOFF / (dx * scale_x))
    return use, scale_y, scale_x, sigma_aa, sigma_dec


def gaussian(data, weight=None, sigma=None, truncate=4.0, threshold=0.5, pixel_sizes=None, resolution=67.0):
    if sigma is None:
        return data
    sy, sx = _sig2(sigma)
    if pixel_sizes is not None:
        use, scy, scx, s_aa, s_dec = decimation(sy, pixel_sizes, resolution)
        if use:
            real = np.asarray(data).real.astype(np.float64) if np.iscomplexobj(data) else np.asarray(data, np.float64)
            aa = _nanconv_real(real, None, kernel_1d(s_aa[0], truncate), kernel_1d(s_aa[1], truncate), threshold)
            ny, nx = aa.shape
            nyt, nxt = ny // scy * scy, nx // scx * scx
            dec = aa[:nyt, :nxt].reshape(nyt // scy, scy, nxt // scx, scx).mean(axis=(1, 3))
            res = _nanconv_real(dec, None, kernel_1d(s_dec[0], truncate), kernel_1d(s_dec[1], truncate), threshold)
            up = np.repeat(np.repeat(res, scy, axis=0), scx, axis=1)
            return np.pad(up, ((0, ny - up.shape[0]), (0, nx - up.shape[1])), mode="edge")
    return nanconv(data, weight, (sy, sx), truncate, threshold)

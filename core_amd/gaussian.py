"""NaN-aware Gaussian filter on the MI355X (insardev/insardev/utils_gaussian.py).

`nanconvolve2d_gaussian(data, weight, sigma, truncate=4.0, threshold=0.5, device='auto')` is the
reference's `nanconvolve2d_gaussian_pytorch` (:14-116), its GPU path; `gaussian(...)` is
`gaussian_numpy` (:221-445) as it runs on a GPU device: the same arguments (sigma a number or
(y, x) in pixels, weight optional, (1, y, x) blocks, complex or real data cast to
complex64/float32), including the decimation path for large sigma with `pixel_sizes`
(anti-alias filter, block mean, main filter on the coarse grid, nearest upsampling, edge padding;
like the reference, that path filters the float32 cast of the data and ignores the weight).  The
convolutions run in the library's kernels (irls2d_gaussian); the block mean and the upsampling of
the decimation path are torch tensor operations on the device.
"""
from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np
import torch

from . import _lib
from .unwrap2d import resolve_device

CUTOFF = 5.3


def _sig2(sigma):
    if not isinstance(sigma, (list, tuple, np.ndarray)):
        return float(sigma), float(sigma)
    return float(sigma[0]), float(sigma[1])


def gaussian_device(data: torch.Tensor, weight: torch.Tensor | None, sigma, truncate=4.0, threshold=0.5) -> torch.Tensor:
    """Filter a 2-D float32 or complex64 HIP tensor (weight: float32 of the same shape or None)."""
    if data.dim() != 2 or data.device.type != 'cuda' or data.dtype not in (torch.float32, torch.complex64):
        raise ValueError("data must be a 2-D float32 or complex64 HIP tensor")
    if weight is not None and (weight.dtype != torch.float32 or tuple(weight.shape) != tuple(data.shape)
                               or weight.device != data.device):
        raise ValueError("weight must be float32 of the data's shape on its device")
    sy, sx = _sig2(sigma)
    data = data.contiguous()
    weight = weight.contiguous() if weight is not None else None
    ny, nx = data.shape
    ch = 2 if data.dtype == torch.complex64 else 1
    L = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.irls2d_gaussian_workspace_size(ny, nx, ch, sy, sx, float(truncate), ctypes.byref(nbytes)),
               "irls2d_gaussian_workspace_size")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=data.device)
    out = torch.empty_like(data)
    stream = torch.cuda.current_stream(data.device).cuda_stream
    _lib.check(L.irls2d_gaussian(ny, nx, ch, data.data_ptr(), weight.data_ptr() if weight is not None else None,
                                 out.data_ptr(), sy, sx, float(truncate), float(threshold), ws.data_ptr(),
                                 int(nbytes.value), stream), "irls2d_gaussian")
    return out


def nanconvolve2d_gaussian(data_np, weight_np=None, sigma=None, truncate=4.0, threshold=0.5, device='auto'):
    """utils_gaussian.py nanconvolve2d_gaussian_pytorch on the HIP device (numpy in and out)."""
    if sigma is None:
        return data_np
    dev = resolve_device(device)
    cplx = np.iscomplexobj(data_np)
    d = torch.from_numpy(np.ascontiguousarray(data_np.astype(np.complex64 if cplx else np.float32))).to(dev)
    w = None if weight_np is None else torch.from_numpy(np.ascontiguousarray(weight_np.astype(np.float32))).to(dev)
    return gaussian_device(d, w, sigma, truncate, threshold).cpu().numpy()


def gaussian(data_np, weight_np=None, sigma=None, truncate=4.0, threshold=0.5, device='auto', pixel_sizes=None,
             resolution=67.0):
    """utils_gaussian.py gaussian_numpy as it runs on a GPU device."""
    if sigma is None:
        return data_np
    sy, sx = _sig2(sigma)
    squeeze = False
    if data_np.ndim == 3 and data_np.shape[0] == 1:
        data_np = data_np[0]
        squeeze = True
    cplx = np.iscomplexobj(data_np)
    data_np = data_np.astype(np.complex64 if cplx else np.float32, copy=False)
    if weight_np is not None:
        weight_np = weight_np.astype(np.float32, copy=False)
    dev = resolve_device(device)
    use = False
    if pixel_sizes is not None:
        dy, dx = pixel_sizes
        scy, scx = int(np.round(resolution / dy)), int(np.round(resolution / dx))
        wavelength = sy * CUTOFF * dy
        use = scy > 1 and scx > 1 and wavelength / resolution > 24
    if use:
        s_aa = (resolution / CUTOFF / dy, resolution / CUTOFF / dx)
        wl_dec = math.sqrt(wavelength ** 2 - resolution ** 2)
        s_dec = (wl_dec / CUTOFF / (dy * scy), wl_dec / CUTOFF / (dx * scx))
        if cplx:
            warnings.warn("gaussian: the decimation path filters the real part of complex data, as the reference's "
                          "float32 cast does", np.exceptions.ComplexWarning)
        real = np.ascontiguousarray(data_np.real if cplx else data_np, dtype=np.float32)
        t = gaussian_device(torch.from_numpy(real).to(dev), None, s_aa, truncate, threshold)
        ny, nx = t.shape
        nyt, nxt = ny // scy * scy, nx // scx * scx
        dec = t[:nyt, :nxt].reshape(nyt // scy, scy, nxt // scx, scx).mean(dim=(1, 3)).contiguous()
        res = gaussian_device(dec, None, s_dec, truncate, threshold)
        up = res.repeat_interleave(scy, dim=0).repeat_interleave(scx, dim=1)
        up = torch.nn.functional.pad(up[None, None], (0, nx - up.shape[1], 0, ny - up.shape[0]), mode='replicate')
        result = up[0, 0].cpu().numpy()
    else:
        result = nanconvolve2d_gaussian(data_np, weight_np, (sy, sx), truncate, threshold, dev)
    if squeeze:
        result = result[np.newaxis, ...]
    return result

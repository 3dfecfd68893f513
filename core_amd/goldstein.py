"""Goldstein adaptive phase filter on the MI355X (insardev/insardev/utils_goldstein.py).

`goldstein(phase, corr, psize=32, threshold=0.5, device='auto')` is the reference's
`BatchComplex._goldstein` (Batch.py:1969-2035) for one image: the same arguments (psize an int or
{'y': .., 'x': ..}, None = no filtering), the same (1, y, x) handling and dtypes (complex64 in and
out, float32 correlation).  On a GPU device the reference runs `goldstein_pytorch`
(utils_goldstein.py:157-330), whose overlapping patches are averaged over the number of valid
patches: that is `mode='count'` (default).  `mode='weight'` gives its CPU path `goldstein_numpy`
(:12-154), which divides by the triangle-window weight sums.  Patch sizes 16, 32 and 64 per axis.
`goldstein_device` works on device tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .unwrap2d import resolve_device

MODES = {"count": 0, "weight": 1}


def _psizes(psize):
    if isinstance(psize, dict):
        return int(psize['y']), int(psize['x'])
    return int(psize), int(psize)


def goldstein_device(phase: torch.Tensor, corr: torch.Tensor, psize=32, threshold=0.5, mode="count",
                     out: torch.Tensor | None = None) -> torch.Tensor:
    """Filter a complex64 (ny, nx) HIP tensor with its float32 correlation; returns complex64."""
    if mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    if phase.dtype != torch.complex64 or phase.dim() != 2 or phase.device.type != 'cuda':
        raise ValueError("phase must be a 2-D complex64 HIP tensor")
    if corr.dtype != torch.float32 or tuple(corr.shape) != tuple(phase.shape) or corr.device != phase.device:
        raise ValueError("corr must be a float32 tensor of the phase's shape on its device")
    phase = phase.contiguous()
    corr = corr.contiguous()
    ny, nx = phase.shape
    py, px = _psizes(psize)
    L = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.irls2d_goldstein_workspace_size(ny, nx, py, px, ctypes.byref(nbytes)),
               "irls2d_goldstein_workspace_size")
    ws = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=phase.device)
    if out is None:
        out = torch.empty_like(phase)
    stream = torch.cuda.current_stream(phase.device).cuda_stream
    _lib.check(L.irls2d_goldstein(ny, nx, phase.data_ptr(), corr.data_ptr(), out.data_ptr(), py, px,
                                  float(threshold), MODES[mode], ws.data_ptr(), int(nbytes.value), stream),
               "irls2d_goldstein")
    return out


def goldstein(phase_np, corr_np, psize=32, threshold=0.5, device='auto', mode="count"):
    """BatchComplex._goldstein (Batch.py:1969-2035) on the HIP device: numpy in, numpy out."""
    if psize is None:
        return phase_np
    squeeze = False
    if phase_np.ndim == 3 and phase_np.shape[0] == 1:
        phase_np = phase_np[0]
        corr_np = corr_np[0] if corr_np.ndim == 3 else corr_np
        squeeze = True
    if phase_np.dtype != np.complex64:
        phase_np = phase_np.astype(np.complex64)
    if corr_np.dtype != np.float32:
        corr_np = corr_np.astype(np.float32)
    dev = resolve_device(device)
    ph = torch.from_numpy(np.ascontiguousarray(phase_np)).to(dev)
    cr = torch.from_numpy(np.ascontiguousarray(corr_np)).to(dev)
    result = goldstein_device(ph, cr, psize=psize, threshold=threshold, mode=mode).cpu().numpy()
    if squeeze:
        result = result[np.newaxis, ...]
    return result

"""Epilogue of the solve on the device (insardev/insardev/utils_unwrap2d.py:1659-1746).

The same steps as core_amd.postprocess.epilogue, with the per-pixel work in the library's kernels
(irls2d_ep_*) and the two decisions that follow numpy's semantics taken here on small arrays:
the median of the cycle counts (from their histogram, as np.median on the float32 values) and
the size order of the components (np.argsort, the reference's own call).  Results equal the
host epilogue's: conncomp exactly, unwrapped phase to float32 rounding of the sin/cos sums.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from . import _lib

TWO_PI = 2 * np.pi
_I32_MAX, _I32_MIN = 2 ** 31 - 1, -2 ** 31
# k saturates at +-2^30 in the kernels (irls2d_ep_scan); above HIST_MAX_BINS bins (or at the
# saturation bound) the median comes from the float32 k values on the host instead.
K_SAT = 2 ** 30
HIST_MAX_BINS = 1 << 24


def _bind(L):
    if getattr(L, "_ep_bound", False):
        return
    vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
    L.irls2d_ep_blocks.argtypes = [i64, i64]
    L.irls2d_ep_blocks.restype = i64
    L.irls2d_ep_scan.argtypes = [i64, i64, i64, vp, vp, vp, vp]
    L.irls2d_ep_hist.argtypes = [i64, i64, i64, vp, vp, ctypes.c_int32, ctypes.c_int32, vp, vp]
    L.irls2d_ep_label.argtypes = [i64, i64, i64, vp, vp, f32, vp, vp, vp]
    L.irls2d_ep_number.argtypes = [i64, i64, vp, vp, vp, vp]
    L.irls2d_ep_sums.argtypes = [i64, i64, i64, vp, vp, vp, vp, vp, vp, vp]
    L.irls2d_ep_apply.argtypes = [i64, i64, i64, vp, vp, vp, vp, vp, vp, vp]
    for f in ("scan", "hist", "label", "number", "sums", "apply"):
        getattr(L, "irls2d_ep_" + f).restype = ctypes.c_int
    L._ep_bound = True


def _median_from_hist(hist, kmin):
    """np.median of float32 values given as counts per integer value (kmin + index)."""
    n = int(hist.sum())
    cum = np.cumsum(hist)
    def value_at(pos):  # pos-th smallest (0-based)
        return np.float32(kmin + int(np.searchsorted(cum, pos, side="right")))
    if n % 2:
        return value_at(n // 2)
    a, b = value_at(n // 2 - 1), value_at(n // 2)
    return np.float32(np.mean(np.array([a, b], np.float32)))


def _median_host(phase, u, offset, P, H, W):
    """np.median(np.round((u - psi) / 2 pi)) over the valid pixels in float32, as
    utils_unwrap2d.py:1671-1674 computes it (used when k does not fit a histogram)."""
    ph = phase[offset:offset + H * P].view(H, P)[:, :W].cpu().numpy()
    un = u[offset:offset + H * P].view(H, P)[:, :W].cpu().numpy()
    v = ~np.isnan(ph)
    kv = np.round((un[v] - ph[v]) / (2 * np.pi))
    return np.float32(np.median(kv))


def epilogue_device(phase, u, offset, pitch, H, W, conncomp_size, stream=None):
    """Post-process image (offset, pitch, H, W) of the device arenas `phase` and `u` (float32
    torch tensors; u is updated in place).  Returns (unwrapped float32 [H, W], conncomp uint16
    [H, W]) as numpy arrays, or None after a RuntimeWarning when a valid pixel is non-finite."""
    L = _lib.load()
    _bind(L)
    dev = u.device
    s = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
    ph_p = phase.data_ptr() + 4 * offset
    u_p = u.data_ptr() + 4 * offset
    H, W, P = int(H), int(W), int(pitch)

    stats = torch.tensor([0, _I32_MAX, _I32_MIN, 0], dtype=torch.int32, device=dev)
    _lib.check(L.irls2d_ep_scan(H, W, P, ph_p, u_p, stats.data_ptr(), s), "irls2d_ep_scan")
    nonfin, kmin, kmax, nvalid = (int(v) for v in stats.cpu())
    if nonfin:
        warnings.warn(f'IRLS produced {nonfin}/{nvalid} NaN/inf values - returning NaN for this component',
                      RuntimeWarning)
        return None
    if nvalid == 0:
        return np.full((H, W), np.nan, np.float32), np.zeros((H, W), np.uint16)
    nb = kmax - kmin + 1
    if kmin <= -K_SAT or kmax >= K_SAT or nb > HIST_MAX_BINS:
        # a diverged (finite) solution: the reference's np.median of the float32 k values
        kmed = _median_host(phase, u, offset, P, H, W)
    else:
        hist = torch.zeros(nb, dtype=torch.int32, device=dev)
        _lib.check(L.irls2d_ep_hist(H, W, P, ph_p, u_p, kmin, nb, hist.data_ptr(), s), "irls2d_ep_hist")
        kmed = _median_from_hist(hist.cpu().numpy().astype(np.int64), kmin)
    # the reference subtracts k_median * 2 pi in float32 (utils_unwrap2d.py:1675)
    shift = np.float32(kmed * TWO_PI) if kmed != 0 else np.float32(0.0)

    nblk = int(L.irls2d_ep_blocks(H, P))
    parent = torch.empty(H * P, dtype=torch.int32, device=dev)
    roots = torch.empty(nblk, dtype=torch.int32, device=dev)
    _lib.check(L.irls2d_ep_label(H, W, P, ph_p, u_p, float(shift), parent.data_ptr(), roots.data_ptr(), s),
               "irls2d_ep_label")
    counts = roots.cpu().numpy().astype(np.int64)
    n_comp = int(counts.sum())
    offsets = torch.from_numpy((np.cumsum(counts) - counts).astype(np.int32)).to(dev)
    root_label = torch.empty(H * P, dtype=torch.int32, device=dev)
    _lib.check(L.irls2d_ep_number(H, P, parent.data_ptr(), offsets.data_ptr(), root_label.data_ptr(), s),
               "irls2d_ep_number")
    n_labels = n_comp + 1
    sums = torch.zeros(4 * n_labels, dtype=torch.float64, device=dev)
    cnt_t = torch.zeros(n_labels, dtype=torch.int64, device=dev)
    _lib.check(L.irls2d_ep_sums(H, W, P, ph_p, u_p, parent.data_ptr(), root_label.data_ptr(), sums.data_ptr(),
                                cnt_t.data_ptr(), s), "irls2d_ep_sums")
    sm = sums.cpu().numpy().reshape(n_labels, 4)
    cnt = cnt_t.cpu().numpy()
    # circular-mean alignment and size relabel, as core_amd.postprocess.epilogue (:1691-1746)
    cf = cnt.astype(np.float64)
    cf[cf == 0] = 1
    d = np.arctan2(sm[:, 0] / cf, sm[:, 1] / cf) - np.arctan2(sm[:, 2] / cf, sm[:, 3] / cf)
    adj = np.arctan2(np.sin(d), np.cos(d))
    adj[0] = 0
    sizes = cnt[1:n_labels]
    order = np.argsort(sizes)[::-1]
    keep = sizes[order] >= conncomp_size
    old_to_new = np.zeros(n_labels, np.uint16)
    old_to_new[order[keep] + 1] = np.arange(1, int(keep.sum()) + 1, dtype=np.uint16)
    adj_t = torch.from_numpy(adj).to(dev)
    relabel_t = torch.from_numpy(old_to_new.view(np.int16)).to(dev)
    cc = torch.empty(H * P, dtype=torch.int16, device=dev)
    _lib.check(L.irls2d_ep_apply(H, W, P, u_p, parent.data_ptr(), root_label.data_ptr(), adj_t.data_ptr(),
                                 relabel_t.data_ptr(), cc.data_ptr(), s), "irls2d_ep_apply")
    un = u[offset:offset + H * P].view(H, P)[:, :W].cpu().numpy()
    ccn = cc.view(H, P)[:, :W].cpu().numpy().view(np.uint16)
    return un, ccn

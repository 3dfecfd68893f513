"""Host prologue and epilogue around the device solve.

Follows insardev/insardev/utils_unwrap2d.py:
  * all-NaN shortcut and bounding-box crop  (:1196-1220)
  * non-finite guard                        (:1659-1665)
  * NaN restore and 2 pi median shift        (:1667-1676)
  * per-component circular-mean alignment    (:1678-1720, float64 accumulators as in :24-75)
  * size relabel into uint16 conncomp        (:1722-1746)
  * un-crop                                  (:1748-1757)
The weight/phase fill of :1226-1241 runs on the device (k_setup).  4-connected labelling uses
scipy.ndimage.label in place of cv2.connectedComponents (both number components in raster
order of their first pixel).
"""
from __future__ import annotations

import warnings

import numpy as np
import scipy.ndimage as ndi

_STRUCT4 = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool)
TWO_PI = 2 * np.pi


def crop_box(phase: np.ndarray):
    """(nan_mask, (r0, r1, c0, c1)) or (nan_mask, None) when every pixel is NaN."""
    nan_mask = np.isnan(phase)
    vr = ~nan_mask.all(axis=1)
    if not vr.any():
        return nan_mask, None
    vc = ~nan_mask.all(axis=0)
    r0 = int(np.argmax(vr)); r1 = vr.size - int(np.argmax(vr[::-1]))
    c0 = int(np.argmax(vc)); c1 = vc.size - int(np.argmax(vc[::-1]))
    return nan_mask, (r0, r1, c0, c1)


def epilogue(u: np.ndarray, phase_c: np.ndarray, nan_mask: np.ndarray, conncomp_size: int):
    """Post-process the cropped solver output in place.

    Returns (unwrapped float32, conncomp uint16) on the cropped grid, or None after a
    RuntimeWarning when a valid pixel is non-finite.
    """
    unwrapped = u
    valid = ~nan_mask
    uv = unwrapped[valid]
    finite = np.isfinite(uv)
    if not finite.all():
        nbad = int(finite.size - np.count_nonzero(finite))
        warnings.warn(f'IRLS produced {nbad}/{int(valid.sum())} NaN/inf values - returning NaN for this component',
                      RuntimeWarning)
        return None
    unwrapped[nan_mask] = np.nan
    conncomp = np.zeros(unwrapped.shape, np.uint16)
    if uv.size == 0:
        return unwrapped, conncomp
    k = np.round((uv - phase_c[valid]) / TWO_PI)
    kmed = np.median(k)
    if kmed != 0:
        uv = uv - kmed * TWO_PI
        unwrapped[valid] = uv
    labels, nlab = ndi.label(valid, structure=_STRUCT4)
    n_labels = nlab + 1
    lv = labels[valid]
    pv = phase_c[valid]
    uv32 = uv.astype(np.float32, copy=False)
    s_in = np.bincount(lv, weights=np.sin(pv).astype(np.float64), minlength=n_labels)
    c_in = np.bincount(lv, weights=np.cos(pv).astype(np.float64), minlength=n_labels)
    s_out = np.bincount(lv, weights=np.sin(uv32).astype(np.float64), minlength=n_labels)
    c_out = np.bincount(lv, weights=np.cos(uv32).astype(np.float64), minlength=n_labels)
    cnt = np.bincount(lv, minlength=n_labels).astype(np.int64)
    cf = cnt.astype(np.float64)
    cf[cf == 0] = 1
    d = np.arctan2(s_in / cf, c_in / cf) - np.arctan2(s_out / cf, c_out / cf)
    adj = np.arctan2(np.sin(d), np.cos(d))
    adj[0] = 0
    # float32 += float64 adjustment, rounded once (utils_unwrap2d.py:69-75)
    unwrapped[valid] = (uv32.astype(np.float64) + adj[lv]).astype(np.float32)
    sizes = cnt[1:n_labels]
    order = np.argsort(sizes)[::-1]
    keep = sizes[order] >= conncomp_size
    old_to_new = np.zeros(n_labels, np.uint16)
    old_to_new[order[keep] + 1] = np.arange(1, int(keep.sum()) + 1, dtype=np.uint16)
    conncomp = old_to_new[labels]
    return unwrapped, conncomp


def uncrop(unwrapped, conncomp, box, shape):
    r0, r1, c0, c1 = box
    if (r1 - r0, c1 - c0) == tuple(shape):
        return unwrapped, conncomp
    out = np.full(shape, np.nan, np.float32)
    out[r0:r1, c0:c1] = unwrapped
    cc = np.zeros(shape, np.uint16)
    cc[r0:r1, c0:c1] = conncomp
    return out, cc

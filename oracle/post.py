"""Oracle restatement of the reference's host prologue/epilogue (test infrastructure).

Follows insardev/insardev/utils_unwrap2d.py:
  prologue  :1196-1241  (all-NaN shortcut, bbox crop, NaN fill, weight fill/clip)
  epilogue  :1659-1757  (non-finite guard, NaN restore, k-median shift,
                         4-connected components, per-component circular-mean
                         alignment with float64 accumulators (:24-75), size
                         relabel, un-crop)
cv2.connectedComponents(connectivity=4) is restated with scipy.ndimage.label
and a 4-connected structuring element; both number components in raster order
of their first pixel.  The per-label accumulation is a plain loop for small
grids and np.bincount for large ones (same float64 sums up to summation order).
"""
from __future__ import annotations

import warnings

import numpy as np
import scipy.ndimage as ndi

STRUCT4 = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool)


def bbox(nan_mask: np.ndarray):
    """Row/column bounds of the valid region (utils_unwrap2d.py:1204-1209)."""
    vr = np.any(~nan_mask, axis=1)
    vc = np.any(~nan_mask, axis=0)
    r0 = int(np.argmax(vr)); r1 = len(vr) - int(np.argmax(vr[::-1]))
    c0 = int(np.argmax(vc)); c1 = len(vc) - int(np.argmax(vc[::-1]))
    return r0, r1, c0, c1


def prologue(phase: np.ndarray, weight):
    """Crop + fill (utils_unwrap2d.py:1196-1241).

    Returns (crop, phase_cropped_view, nan_mask_cropped, phase_filled f32,
    weight_filled f32) or None when every pixel is NaN.
    """
    nan_mask = np.isnan(phase)
    if np.all(nan_mask):
        return None
    r0, r1, c0, c1 = bbox(nan_mask)
    ph = phase[r0:r1, c0:c1]
    nm = nan_mask[r0:r1, c0:c1]
    wt = weight[r0:r1, c0:c1] if weight is not None else None
    pf = ph.copy()
    np.copyto(pf, 0.0, where=nm)
    if wt is not None:
        wf = wt.copy()
        np.copyto(wf, 0.0, where=np.isnan(wf))
        np.copyto(wf, 0.0, where=nm)
        np.clip(wf, 0.01, 1.0, out=wf)
    else:
        wf = np.subtract(1.0, nm, dtype=np.float32)
    return (r0, r1, c0, c1), ph, nm, pf.astype(np.float32), wf.astype(np.float32)


def label4(valid: np.ndarray):
    """4-connected labels, background 0, components 1..n in raster order."""
    lab, n = ndi.label(valid, structure=STRUCT4)
    return lab.astype(np.int32), n + 1


def accumulate(labels, psi, u, n_labels, loop_limit=4096):
    """Per-label float64 sums of sin/cos of psi and u plus counts (utils_unwrap2d.py:24-66)."""
    lf = labels.ravel()
    pf = psi.ravel().astype(np.float32)
    uf = u.ravel().astype(np.float32)
    if lf.size <= loop_limit:
        s_in = np.zeros(n_labels); c_in = np.zeros(n_labels)
        s_out = np.zeros(n_labels); c_out = np.zeros(n_labels)
        cnt = np.zeros(n_labels, np.int64)
        for i in range(lf.size):
            k = lf[i]
            if k > 0:
                s_in[k] += float(np.sin(pf[i])); c_in[k] += float(np.cos(pf[i]))
                s_out[k] += float(np.sin(uf[i])); c_out[k] += float(np.cos(uf[i]))
                cnt[k] += 1
        return s_in, c_in, s_out, c_out, cnt
    m = lf > 0
    k = lf[m]
    s_in = np.bincount(k, weights=np.sin(pf[m]).astype(np.float64), minlength=n_labels)
    c_in = np.bincount(k, weights=np.cos(pf[m]).astype(np.float64), minlength=n_labels)
    s_out = np.bincount(k, weights=np.sin(uf[m]).astype(np.float64), minlength=n_labels)
    c_out = np.bincount(k, weights=np.cos(uf[m]).astype(np.float64), minlength=n_labels)
    cnt = np.bincount(k, minlength=n_labels).astype(np.int64)
    return s_in, c_in, s_out, c_out, cnt


def epilogue(u: np.ndarray, phase_c: np.ndarray, nan_mask: np.ndarray, conncomp_size: int):
    """Post-processing of the solver output on the cropped grid (utils_unwrap2d.py:1659-1746).

    Returns (unwrapped f32, conncomp uint16) on the cropped grid, or None when the
    solver output has non-finite valid pixels (the caller warns and returns NaNs).
    """
    unwrapped = np.asarray(u, dtype=np.float32).copy()
    valid = ~nan_mask
    if not np.isfinite(unwrapped[valid]).all():
        nbad = int(np.sum(~np.isfinite(unwrapped[valid])))
        warnings.warn(f'IRLS produced {nbad}/{int(valid.sum())} NaN/inf values - returning NaN for this component',
                      RuntimeWarning)
        return None
    unwrapped[nan_mask] = np.nan
    if np.any(valid):
        diff = unwrapped[valid] - phase_c[valid]
        kv = np.round(diff / (2 * np.pi))
        kmed = np.median(kv)
        unwrapped[valid] = unwrapped[valid] - kmed * 2 * np.pi
    conncomp = np.zeros(unwrapped.shape, np.uint16)
    if np.any(valid):
        labels, n_labels = label4(valid)
        s_in, c_in, s_out, c_out, cnt = accumulate(labels, phase_c, unwrapped, n_labels)
        cf = cnt.astype(np.float64)
        cf[cf == 0] = 1
        in_mean = np.arctan2(s_in / cf, c_in / cf)
        out_mean = np.arctan2(s_out / cf, c_out / cf)
        d = in_mean - out_mean
        adj = np.arctan2(np.sin(d), np.cos(d))
        adj[0] = 0
        uf = unwrapped.ravel()
        lf = labels.ravel()
        sel = lf > 0
        uf[sel] = uf[sel] + adj[lf[sel]]
        sizes = cnt[1:n_labels]
        order = np.argsort(sizes)[::-1]
        old_to_new = np.zeros(n_labels, np.uint16)
        nl = 1
        for oi in order:
            if sizes[oi] >= conncomp_size:
                old_to_new[oi + 1] = nl
                nl += 1
        conncomp = old_to_new[labels].astype(np.uint16)
    return unwrapped, conncomp


def uncrop(unwrapped, conncomp, crop, shape):
    r0, r1, c0, c1 = crop
    if (r1 - r0, c1 - c0) == tuple(shape):
        return unwrapped, conncomp
    out = np.full(shape, np.nan, np.float32)
    out[r0:r1, c0:c1] = unwrapped
    cc = np.zeros(shape, np.uint16)
    cc[r0:r1, c0:c1] = conncomp
    return out, cc

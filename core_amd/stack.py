"""Stack-level unwrapping on the MI355X: the reference's `Stack_unwrap2d` entry points for whole
SBAS stacks (insardev/insardev/Stack_unwrap2d.py).

The reference runs one dask task per pair, each taking a process-wide GPU lock and calling
`utils_unwrap2d.irls_unwrap_2d` (:17-29, :32-71, :753-990).  Here every pair of a burst variable
goes through ONE batched device solve (`irls_unwrap_2d_batch`), with the same per-pair results.

* `unwrap_stack(phase[pair, y, x] or [y, x], weight)` -> (unwrapped float32, conncomp uint16):
  the array core.
* `unwrap2d_irls`, `unwrap2d`, `unwrap2d_dataset`: the reference's methods over burst mappings
  of xarray Datasets (keys -> Dataset), with the same arguments, defaults, errors and output
  structure (dims ('pair', 'y', 'x') or ('y', 'x'), y/x coords, the pair dimension's
  non-dimension coords such as ref/rep/BPR, `units = 'radians'`, dataset attrs, CRS when
  rioxarray is present, the input's spatial chunks when it was dask-backed).  xarray is an
  optional dependency: the functions import it on use (or take the module as `xr`).  Unlike the
  reference they compute eagerly: each burst variable is one device call over all its pairs.
* `unwrap_chunked` / `unwrap2d_chunk`: the per-chunk unwrapping with overlap margins of
  Stack_unwrap2d.py:992-1136 (dask map_overlap, boundary 'none'), with every tile of every pair
  in one batched device solve.
* `filter_small_components`: the first step of the reference's component linking
  (Stack_unwrap2d.py:343-375), which NaNs out labelled components below max(conncomp_size, 4).
  `unwrap2d(conncomp=False)` then links the remaining components by integer 2 pi offsets
  (core_amd.linking, Stack_unwrap2d.py:377-397, utils_unwrap2d.py:195-444, :886-1040).
"""
from __future__ import annotations

import importlib

import numpy as np

from .linking import link_components
from .unwrap2d import irls_unwrap_2d_batch

__all__ = ["unwrap_stack", "filter_small_components", "unwrap2d_irls", "unwrap2d", "unwrap2d_dataset",
           "chunk_windows", "unwrap_chunked", "unwrap2d_chunk"]


def unwrap_stack(phase, weight=None, device='auto', max_iter=50, tol=1e-2, cg_max_iter=10, cg_tol=1e-3,
                 epsilon=1e-2, conncomp_size=30, debug=False, graph=False):
    """Unwrap a (pair, y, x) or (y, x) stack in one device call.

    Defaults are unwrap2d_irls's (Stack_unwrap2d.py:753-755).  Per pair the result equals
    irls_unwrap_2d (an all-NaN pair gives NaN and zeros, as the reference task does,
    Stack_unwrap2d.py:54-59).  Returns (unwrapped float32, conncomp uint16) of the input's shape.
    """
    ph = np.asarray(phase)
    if ph.ndim not in (2, 3):
        raise ValueError(f"Expected 2D or 3D phase, got shape {ph.shape}")
    squeeze = ph.ndim == 2
    ph3 = ph[None] if squeeze else ph
    w3 = None
    if weight is not None:
        w3 = np.asarray(weight)
        w3 = w3[None] if w3.ndim == 2 else w3
        if w3.shape != ph3.shape:
            raise ValueError(f"weight shape {w3.shape} does not match phase shape {ph3.shape}")
    res = irls_unwrap_2d_batch([p for p in ph3], None if w3 is None else [w for w in w3], device=device,
                               max_iter=max_iter, tol=tol, cg_max_iter=cg_max_iter, cg_tol=cg_tol, epsilon=epsilon,
                               conncomp_size=conncomp_size, debug=debug, graph=graph)
    unw = np.stack([r[0].astype(np.float32) for r in res])
    cc = np.stack([r[1].astype(np.uint16) for r in res])
    return (unw[0], cc[0]) if squeeze else (unw, cc)


def filter_small_components(unwrapped, conncomp, conncomp_size=1000):
    """NaN out the labelled components smaller than max(conncomp_size, 4) pixels.

    Restates Stack_unwrap2d.py:343-375 for one 2-D grid: with the size-ordered IRLS labels, the
    components with label >= 1 and at least min_size pixels are kept, label-0 pixels (invalid, or
    components already below the solver's conncomp_size) are left as they are, every other
    labelled pixel becomes NaN.  Returns (filtered copy, number of kept components); a grid
    without valid pixels is returned unchanged.
    """
    u = np.asarray(unwrapped)
    if not np.any(~np.isnan(u)):
        return u, 0
    labeled = np.asarray(conncomp).astype(np.int32)
    min_size = max(conncomp_size, 4)
    n_total = int(labeled.max())
    sizes = np.bincount(labeled.ravel(), minlength=n_total + 1)
    kept = np.array([i for i in range(1, n_total + 1) if sizes[i] >= min_size], dtype=np.int32)
    is_linked = np.isin(labeled, kept) | (labeled == 0)
    out = u.copy()
    out[~is_linked] = np.nan
    return out, int(kept.size)


def _xr(xr):
    return xr if xr is not None else importlib.import_module("xarray")


def _pair_coords(da):
    """Non-dimension coords along 'pair' (ref, rep, BPR, ...), Stack_unwrap2d.py:873-885."""
    out = {}
    if 'pair' in da.dims:
        for k, v in da.coords.items():
            if k != 'pair' and getattr(v, 'dims', None) == ('pair',):
                out[k] = ('pair', np.asarray(v.values if hasattr(v, 'values') else v))
    return out


def _spatial_chunks(da):
    chunks = getattr(getattr(da, 'data', None), 'chunks', None)
    if isinstance(chunks, tuple) and len(chunks) >= 2:
        return chunks[-2], chunks[-1]
    return None


def unwrap2d_irls(phase, weight=None, device='auto', max_iter=50, tol=1e-2, cg_max_iter=10, cg_tol=1e-3,
                  epsilon=1e-2, conncomp_size=30, semaphore=8, debug=False, xr=None):
    """Stack_unwrap2d.unwrap2d_irls (:753-990) over a burst mapping {key: xr.Dataset}.

    Returns a tuple (unwrapped {key: Dataset}, conncomp {key: Dataset}) like the reference's
    Batches((Batch, BatchUnit)).  `semaphore` only limits CPU tasks in the reference and has no
    effect here (no CPU path).
    """
    xr = _xr(xr)
    unwrap_result, conncomp_result = {}, {}
    for key in phase.keys():
        ds = phase[key]
        wds = weight[key] if weight is not None and key in weight else None
        data_vars = [v for v in ds.data_vars if 'y' in ds[v].dims and 'x' in ds[v].dims]
        if debug:
            print(f'unwrap2d_irls: burst {key!r}, variables {data_vars}')
        uvars, cvars = {}, {}
        for var in data_vars:
            pda = ds[var]
            wda = wds[var] if wds is not None else None
            chunks = _spatial_chunks(pda)
            arr = np.asarray(pda.values, dtype=np.float32)
            warr = None if wda is None else np.asarray(wda.values, dtype=np.float32)
            unw, cc = unwrap_stack(arr, warr, device=device, max_iter=max_iter, tol=tol, cg_max_iter=cg_max_iter,
                                   cg_tol=cg_tol, epsilon=epsilon, conncomp_size=conncomp_size, debug=debug)
            dims = ('pair', 'y', 'x') if 'pair' in pda.dims else ('y', 'x')
            coords = {'y': pda.y, 'x': pda.x}
            uda = xr.DataArray(unw, dims=dims, coords=coords)
            uda.attrs['units'] = 'radians'
            cda = xr.DataArray(cc, dims=dims, coords=coords)
            pc = _pair_coords(pda)
            if pc:
                uda = uda.assign_coords(**pc)
                cda = cda.assign_coords(**pc)
            if chunks is not None and (len(chunks[0]) > 1 or len(chunks[1]) > 1):
                uda = uda.chunk({'y': chunks[0], 'x': chunks[1]})
                cda = cda.chunk({'y': chunks[0], 'x': chunks[1]})
            uvars[var] = uda
            cvars[var] = cda
        unwrap_result[key] = xr.Dataset(uvars, attrs=dict(ds.attrs))
        conncomp_result[key] = xr.Dataset(cvars, attrs=dict(ds.attrs))
        rio = getattr(ds, 'rio', None)
        crs = getattr(rio, 'crs', None) if rio is not None else None
        if crs is not None:
            unwrap_result[key].rio.write_crs(crs, inplace=True)
            conncomp_result[key].rio.write_crs(crs, inplace=True)
    return unwrap_result, conncomp_result


def unwrap2d(phase, weight=None, conncomp=False, conncomp_size=1_000, conncomp_gap=None, conncomp_linksize=5,
             conncomp_linkcount=30, device='auto', debug=False, xr=None, **kwargs):
    """Stack_unwrap2d.unwrap2d (:525-643) over a burst mapping {key: xr.Dataset}.

    conncomp=True returns (unwrapped, conncomp).  conncomp=False (the default) links each grid's
    components as the reference's `_link_components` does (Stack_unwrap2d.py:293-454): components
    below max(conncomp_size, 4) pixels become NaN, the others are shifted by the integer 2 pi
    offsets of core_amd.linking.link_components, and the unwrapped mapping is returned.
    """
    if not conncomp and conncomp_linksize > conncomp_size:
        raise ValueError(
            f'conncomp_linksize ({conncomp_linksize}) cannot be greater than conncomp_size ({conncomp_size}). '
            f'Components must have at least conncomp_linksize pixels for reliable offset estimation.')
    xr = _xr(xr)
    unwrapped, labels = unwrap2d_irls(phase, weight, device=device, conncomp_size=conncomp_size, debug=debug,
                                      xr=xr, **kwargs)
    if conncomp:
        return unwrapped, labels
    result = {}
    for key, ds in unwrapped.items():
        lds = labels[key]
        vars_ = {}
        for var in ds.data_vars:
            uda = ds[var]
            u = np.asarray(uda.values)
            cc = np.asarray(lds[var].values)
            u3, c3 = (u[None], cc[None]) if u.ndim == 2 else (u, cc)
            out = np.empty_like(u3, dtype=np.float32)
            for p in range(u3.shape[0]):
                out[p] = link_components(u3[p], c3[p], conncomp_size=conncomp_size, conncomp_gap=conncomp_gap,
                                         conncomp_linksize=conncomp_linksize, conncomp_linkcount=conncomp_linkcount,
                                         debug=debug)
            out = out[0] if u.ndim == 2 else out
            nda = xr.DataArray(out, dims=uda.dims, coords=uda.coords)
            nda.attrs.update(dict(uda.attrs))
            vars_[var] = nda
        result[key] = xr.Dataset(vars_, attrs=dict(ds.attrs))
    return result


def unwrap2d_dataset(phase, weight=None, conncomp=False, conncomp_size=1000, conncomp_gap=None,
                     conncomp_linksize=5, conncomp_linkcount=30, device='auto', debug=False, xr=None, **kwargs):
    """Stack_unwrap2d.unwrap2d_dataset (:645-751): one merged xr.Dataset in, Dataset(s) out."""
    xr = _xr(xr)
    if not isinstance(phase, xr.Dataset):
        raise TypeError(f"phase must be xr.Dataset, got {type(phase).__name__}")
    if weight is not None and not isinstance(weight, xr.Dataset):
        raise TypeError(f"weight must be xr.Dataset, got {type(weight).__name__}")
    res = unwrap2d({'': phase}, None if weight is None else {'': weight}, conncomp=conncomp,
                   conncomp_size=conncomp_size, conncomp_gap=conncomp_gap, conncomp_linksize=conncomp_linksize,
                   conncomp_linkcount=conncomp_linkcount, device=device, debug=debug, xr=xr, **kwargs)
    if conncomp:
        return res[0][''], res[1]['']
    return res['']


def _depth(ov, chunk):
    """Stack_unwrap2d.py:1084-1086: fractions of the first chunk or pixels, at least 1."""
    if ov is None:
        ov = 0
    return max(1, int(ov * chunk) if isinstance(ov, float) else int(ov))


def _chunk_bounds(n, c):
    """Regular chunks of size c along a dimension of length n (the last one shorter)."""
    edges = list(range(0, n, c)) + [n]
    return list(zip(edges[:-1], edges[1:]))


def chunk_windows(H, W, chunks, overlap=None):
    """Tiles of a dask map_overlap with boundary 'none' (Stack_unwrap2d.py:1081-1111).

    chunks = (cy, cx) (the spatial chunk sizes, as dask's first chunk of each axis); overlap as
    the reference's argument (None, a pixel int, a chunk fraction float, or a per-axis tuple).
    Returns [(core window, extended window)] with windows as (y0, y1, x0, x1): a chunk grows by
    the depth towards every neighbour chunk and not beyond the array; the unwrapped extended
    tile is trimmed back to the core."""
    cy, cx = int(chunks[0]), int(chunks[1])
    ov_y, ov_x = overlap if isinstance(overlap, tuple) else (overlap, overlap)
    dy, dx = _depth(ov_y, cy), _depth(ov_x, cx)
    out = []
    for y0, y1 in _chunk_bounds(H, cy):
        for x0, x1 in _chunk_bounds(W, cx):
            ey0 = y0 - dy if y0 > 0 else y0
            ey1 = y1 + dy if y1 < H else y1
            ex0 = x0 - dx if x0 > 0 else x0
            ex1 = x1 + dx if x1 < W else x1
            out.append(((y0, y1, x0, x1), (max(0, ey0), min(H, ey1), max(0, ex0), min(W, ex1))))
    return out


def unwrap_chunked(phase, weight=None, chunks=None, overlap=None, device='auto', max_iter=50, tol=1e-2,
                   cg_max_iter=10, cg_tol=1e-3, epsilon=1e-2, conncomp_size=30, debug=False):
    """Per-chunk unwrapping with overlap of a (pair, y, x) or (y, x) array.

    Same tiles and per-tile calls as Stack_unwrap2d.unwrap2d_chunk (each extended tile of each
    pair unwrapped independently, all-NaN tiles passed through, the margins trimmed), with all
    tiles of all pairs solved in one batched device call.  chunks = (cy, cx); None = one chunk.
    Returns the unwrapped float32 array (no conncomp, as the reference)."""
    ph = np.asarray(phase)
    if ph.ndim not in (2, 3):
        raise ValueError(f"Expected 2D or 3D phase, got shape {ph.shape}")
    squeeze = ph.ndim == 2
    ph3 = ph[None] if squeeze else ph
    w3 = None if weight is None else (np.asarray(weight)[None] if squeeze else np.asarray(weight))
    H, W = ph3.shape[1:]
    wins = chunk_windows(H, W, chunks if chunks is not None else (H, W), overlap)
    tiles, wt, where = [], [], []
    for p in range(ph3.shape[0]):
        for core, ext in wins:
            ey0, ey1, ex0, ex1 = ext
            tiles.append(ph3[p, ey0:ey1, ex0:ex1])
            wt.append(None if w3 is None else w3[p, ey0:ey1, ex0:ex1])
            where.append((p, core, ext))
    if debug:
        print(f'unwrap_chunked: {len(tiles)} tiles of {len(wins)} windows per pair in one batched solve')
    res = irls_unwrap_2d_batch(tiles, None if w3 is None else wt, device=device, max_iter=max_iter, tol=tol,
                               cg_max_iter=cg_max_iter, cg_tol=cg_tol, epsilon=epsilon,
                               conncomp_size=conncomp_size, debug=debug)
    out = np.empty(ph3.shape, np.float32)
    for (p, core, ext), (u, _) in zip(where, res):
        y0, y1, x0, x1 = core
        ey0, _, ex0, _ = ext
        out[p, y0:y1, x0:x1] = u[y0 - ey0:y1 - ey0, x0 - ex0:x1 - ex0]
    return out[0] if squeeze else out


def unwrap2d_chunk(phase, weight=None, overlap=None, device='auto', semaphore=8, debug=False, chunks=None,
                   xr=None, **kwargs):
    """Stack_unwrap2d.unwrap2d_chunk (:992-1136) over a burst mapping {key: xr.Dataset}.

    The spatial chunks are the DataArray's dask chunks when it has them, else `chunks`
    ((cy, cx); None = one chunk).  kwargs: max_iter (50), tol (1e-2), cg_max_iter (10), cg_tol
    (1e-3), epsilon (1e-2), conncomp_size (30), the reference's defaults (:1063-1069).  Returns
    {key: Dataset} of unwrapped float32 phase (units 'radians', attrs and CRS kept)."""
    xr = _xr(xr)
    params = dict(max_iter=kwargs.get('max_iter', 50), tol=kwargs.get('tol', 1e-2),
                  cg_max_iter=kwargs.get('cg_max_iter', 10), cg_tol=kwargs.get('cg_tol', 1e-3),
                  epsilon=kwargs.get('epsilon', 1e-2), conncomp_size=kwargs.get('conncomp_size', 30))
    result = {}
    for key in phase.keys():
        ds = phase[key]
        wds = weight[key] if weight is not None and key in weight else None
        data_vars = [v for v in ds.data_vars if 'y' in ds[v].dims and 'x' in ds[v].dims]
        uvars = {}
        for var in data_vars:
            pda = ds[var]
            wda = wds[var] if wds is not None else None
            dch = getattr(getattr(pda, 'data', None), 'chunks', None)
            ch = (dch[-2][0], dch[-1][0]) if isinstance(dch, tuple) and len(dch) >= 2 else chunks
            arr = np.asarray(pda.values, dtype=np.float32)
            warr = None if wda is None else np.asarray(wda.values, dtype=np.float32)
            unw = unwrap_chunked(arr, warr, chunks=ch, overlap=overlap, device=device, debug=debug, **params)
            uda = xr.DataArray(unw, dims=pda.dims, coords=pda.coords)
            uda.attrs['units'] = 'radians'
            uvars[var] = uda
        result[key] = xr.Dataset(uvars, attrs=dict(ds.attrs))
        rio = getattr(ds, 'rio', None)
        crs = getattr(rio, 'crs', None) if rio is not None else None
        if crs is not None:
            result[key].rio.write_crs(crs, inplace=True)
    return result

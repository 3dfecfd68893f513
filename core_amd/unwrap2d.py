"""Drop-in replacement of insardev's 2-D IRLS unwrap call on MI355X.

`irls_unwrap_2d` keeps the signature, defaults, return types, warnings and errors of
insardev/insardev/utils_unwrap2d.py:1082-1766; the solve runs in the gfx950 library
(include/irls2d.h) through ctypes, with device memory and streams from torch.
`irls_unwrap_2d_batch` unwraps many interferograms of uneven shapes in one library call
(one launch per solver step for the whole stack), which is what the reference's dask graph
does one pair at a time under a process-wide GPU lock (Stack_unwrap2d.py:17-29, :32-71).
`solve_device` is the device-level entry (tensors already in HBM) used by the benchmark.
"""
from __future__ import annotations

import threading
import time
import weakref

import numpy as np
import torch

from . import _lib
from . import postprocess as pp
from .epilogue_gpu import epilogue_device

__all__ = ["irls_unwrap_2d", "irls_unwrap_2d_batch", "solve_device", "precondition_device", "resolve_device",
           "Workspace"]

# The reference serialises GPU solves per process with one lock (Stack_unwrap2d.py:20, :63);
# the same lock guards the shared workspaces here.
_gpu_lock = threading.RLock()


def resolve_device(device) -> torch.device:
    """Map the reference's device argument (utils_unwrap2d.py:1173-1186, utils_torch.py:12-62).

    'auto' and 'cuda[:k]' select the HIP device (PyTorch-ROCm names it 'cuda').  This build has
    only the gfx950 path: 'cpu', 'mps' and 'tpu' raise ValueError instead of falling back.
    """
    if isinstance(device, torch.device):
        dev = device
    elif isinstance(device, str):
        if device == 'auto':
            if not torch.cuda.is_available():
                raise RuntimeError("irls_unwrap_2d: device='auto' found no HIP GPU; this build has no CPU path")
            dev = torch.device('cuda', torch.cuda.current_device())
        elif device in ('cpu', 'mps', 'tpu'):
            raise ValueError(f"irls_unwrap_2d: device={device!r} is not supported by the MI355X build "
                             "(use 'cuda' or 'auto')")
        else:
            dev = torch.device(device)
    else:
        raise ValueError(f"irls_unwrap_2d: bad device {device!r}")
    if dev.type != 'cuda':
        raise ValueError(f"irls_unwrap_2d: device {dev} is not a HIP device")
    if dev.index is None:
        dev = torch.device('cuda', torch.cuda.current_device())
    return dev


def _release_address(ptr: int):
    """Drop the library's cached graphs that refer to the workspace at `ptr`."""
    try:
        _lib.load().irls2d_release(ptr)
    except Exception:  # interpreter shutdown: the library may already be gone
        pass


class Workspace:
    """Device workspace for one set of shapes (reused across calls with the same shapes).

    The library's cached graphs for this workspace are released when the object is collected
    (or by `release()`), before its device memory returns to torch's allocator."""

    _cache: dict = {}

    def __init__(self, shapes, device):
        self.shapes = [(int(h), int(w)) for h, w in shapes]
        self.device = device
        self.offsets, self.pitches, self.total = _lib.arena_layout(self.shapes)
        self.nbytes = _lib.workspace_size(self.shapes)
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.info = torch.zeros(4 * len(self.shapes), dtype=torch.int32, device=device)
        self._finalizer = weakref.finalize(self, _release_address, self.buf.data_ptr())

    @classmethod
    def get(cls, shapes, device):
        key = (device.index, tuple((int(h), int(w)) for h, w in shapes))
        ws = cls._cache.get(key)
        if ws is not None and _lib.workspace_size(ws.shapes) > ws.nbytes:  # plans changed (autotune)
            cls._cache.pop(key).release()
            ws = None
        if ws is None:
            if len(cls._cache) >= 4:
                for k in list(cls._cache)[:len(cls._cache) - 3]:
                    old = cls._cache.pop(k)
                    old.release()
            ws = cls(shapes, device)
            cls._cache[key] = ws
        return ws

    def release(self):
        self._finalizer.detach()
        _lib.check(_lib.load().irls2d_release(self.buf.data_ptr()), "irls2d_release")

    @classmethod
    def clear(cls):
        for ws in cls._cache.values():
            ws.release()
        cls._cache.clear()


def solve_device(phase: torch.Tensor, weight, shapes, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4,
                 epsilon=1e-2, out: torch.Tensor | None = None, graph: bool = False, workspace=None,
                 time_kernels: bool = False):
    """Run the solver on device arenas.

    phase, weight: float32 device tensors holding the images at `Workspace.offsets`
    (weight may be None).  Returns (u arena tensor, info int32[n, 4] device tensor).
    """
    if phase.dtype != torch.float32 or phase.device.type != 'cuda' or not phase.is_contiguous():
        raise ValueError("phase must be a contiguous float32 HIP tensor")
    if weight is not None and (weight.dtype != torch.float32 or weight.device != phase.device
                               or not weight.is_contiguous() or weight.numel() != phase.numel()):
        raise ValueError("weight must match phase (contiguous float32 on the same device)")
    L = _lib.load()
    ws = workspace or Workspace.get(shapes, phase.device)
    if phase.numel() < ws.total:
        raise ValueError(f"phase arena has {phase.numel()} elements, layout needs {ws.total}")
    if out is None:
        out = torch.empty(ws.total, dtype=torch.float32, device=phase.device)
    elif out.numel() < ws.total or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out arena too small or not contiguous float32")
    n, H, W = _lib.shapes_arrays(ws.shapes)
    stream = torch.cuda.current_stream(phase.device).cuda_stream
    rc = L.irls2d_solve(n, H, W, phase.data_ptr(), weight.data_ptr() if weight is not None else None,
                        out.data_ptr(), int(max_iter), float(tol), int(cg_max_iter), float(cg_tol),
                        float(epsilon), ws.buf.data_ptr(), ws.nbytes, ws.info.data_ptr(), stream,
                        (_lib.FLAG_GRAPH if graph else 0) | (_lib.FLAG_TIME_KERNELS if time_kernels else 0))
    _lib.check(rc, "irls2d_solve")
    return out, ws.info.view(n, 4)


def precondition_device(r: torch.Tensor, shapes, out: torch.Tensor | None = None, workspace=None):
    """z = IDCT2(-DCT2(r) / lambda), DC zeroed, per image of an arena (the reference's
    apply_preconditioner closure, utils_unwrap2d.py:1390-1403).  Returns the z arena."""
    if r.dtype != torch.float32 or r.device.type != 'cuda' or not r.is_contiguous():
        raise ValueError("r must be a contiguous float32 HIP tensor")
    L = _lib.load()
    ws = workspace or Workspace.get(shapes, r.device)
    if r.numel() < ws.total:
        raise ValueError(f"r arena has {r.numel()} elements, layout needs {ws.total}")
    if out is None:
        out = torch.empty(ws.total, dtype=torch.float32, device=r.device)
    elif out.numel() < ws.total or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out arena too small or not contiguous float32")
    n, H, W = _lib.shapes_arrays(ws.shapes)
    stream = torch.cuda.current_stream(r.device).cuda_stream
    _lib.check(L.irls2d_precondition(n, H, W, r.data_ptr(), out.data_ptr(), ws.buf.data_ptr(), ws.nbytes, stream),
               "irls2d_precondition")
    return out


def image_view(arena, ws, k):
    """Rows of image k inside an arena (numpy array or torch tensor) laid out by `ws`."""
    h, w = ws.shapes[k]
    o, p = ws.offsets[k], ws.pitches[k]
    return arena[o:o + h * p].reshape(h, p)[:, :w]


def pack(arrays, ws, device):
    """Copy 2-D float32 arrays into a pinned host arena and upload it."""
    host = torch.zeros(ws.total, dtype=torch.float32, pin_memory=True)
    hv = host.numpy()
    for k, a in enumerate(arrays):
        image_view(hv, ws, k)[...] = a
    return host.to(device, non_blocking=True)


def memory_groups(shapes, budget_bytes):
    """Split image shapes (in order) into consecutive groups whose batched workspace plus the
    phase / weight / output arenas stay within `budget_bytes` (a single image that alone exceeds
    it still forms its own group).  Returns lists of indices."""
    groups, cur, cur_bytes = [], [], 0
    for i, (h, w) in enumerate(shapes):
        pitch = (int(w) + 63) // 64 * 64
        need = _lib.workspace_size([(h, w)]) + 3 * int(h) * pitch * 4
        if cur and cur_bytes + need > budget_bytes:
            groups.append(cur)
            cur, cur_bytes = [], 0
        cur.append(i)
        cur_bytes += need
    if cur:
        groups.append(cur)
    return groups


def irls_unwrap_2d_batch(phases, weights=None, device='auto', max_iter=50, tol=1e-3, cg_max_iter=20,
                         cg_tol=1e-4, epsilon=1e-2, conncomp_size=30, debug=False, graph=False,
                         memory_fraction=0.6, epilogue='device'):
    """Unwrap a list of 2-D interferograms, all images of a group in one device solve.

    Each image gets exactly the per-image semantics of `irls_unwrap_2d` (crop, solve, epilogue);
    returns a list of (unwrapped float32, conncomp uint16) tuples in input order.  The images are
    split into consecutive groups whose device memory (workspace and arenas) fits within
    `memory_fraction` of the free HBM, so stacks of any length can be passed.  epilogue='device'
    (default) runs the post-processing of utils_unwrap2d.py:1659-1746 in the library's kernels,
    'host' in numpy/scipy (core_amd.postprocess); both give the same conncomp labels and the same
    phase to float32 rounding of the per-component sin/cos sums.
    """
    if epilogue not in ('device', 'host'):
        raise ValueError("epilogue must be 'device' or 'host'")
    phases = list(phases)
    if weights is None:
        weights = [None] * len(phases)
    weights = list(weights)
    if len(weights) != len(phases):
        raise ValueError("weights must match phases")
    dev = resolve_device(device)
    results = [None] * len(phases)
    jobs = []
    for k, (ph, wt) in enumerate(zip(phases, weights)):
        ph = np.asarray(ph)
        if ph.ndim != 2:
            raise ValueError(f"irls_unwrap_2d expects 2D phase array, got shape {ph.shape}")
        nan_mask, box = pp.crop_box(ph)
        if box is None:
            results[k] = (np.full_like(ph, np.nan), np.zeros(ph.shape, np.uint16))
            continue
        r0, r1, c0, c1 = box
        jobs.append((k, ph, None if wt is None else np.asarray(wt), nan_mask[r0:r1, c0:c1], box))
    if not jobs:
        return results
    shapes_all = [(j[4][1] - j[4][0], j[4][3] - j[4][2]) for j in jobs]
    free, _total = torch.cuda.mem_get_info(dev)
    groups = memory_groups(shapes_all, int(free * memory_fraction))
    for g in groups:
        # one group: keep its workspace cached for the next call; several: free each before the next
        _solve_group([jobs[i] for i in g], results, dev, max_iter, tol, cg_max_iter, cg_tol, epsilon,
                     conncomp_size, debug, graph, cache=len(groups) == 1, epilogue=epilogue)
    return results


def _solve_group(jobs, results, dev, max_iter, tol, cg_max_iter, cg_tol, epsilon, conncomp_size, debug, graph,
                 cache=True, epilogue='device'):
    t0 = time.time()
    shapes = [(j[4][1] - j[4][0], j[4][3] - j[4][2]) for j in jobs]
    has_w = [j[2] is not None for j in jobs]
    with _gpu_lock, torch.cuda.device(dev):
        ws = Workspace.get(shapes, dev) if cache else Workspace(shapes, dev)
        try:
            eps, info_h, u_host, t1, t2 = _run_group(jobs, ws, dev, shapes, has_w, max_iter, tol, cg_max_iter,
                                                     cg_tol, epsilon, conncomp_size, graph, epilogue)
        finally:
            if not cache:
                ws.release()  # its device memory returns to torch's allocator with the object
    if epilogue != 'device':
        uh = u_host.numpy()
        eps = []
        for idx, (k, ph, wt, nm, box) in enumerate(jobs):
            r0, r1, c0, c1 = box
            eps.append(pp.epilogue(image_view(uh, ws, idx).copy(), ph[r0:r1, c0:c1], nm, conncomp_size))
    for (k, ph, wt, nm, box), (h, w), inf, ep in zip(jobs, shapes, info_h, eps):
        if ep is None:
            results[k] = (np.full(ph.shape, np.nan, np.float32), np.zeros(ph.shape, np.uint16))
            continue
        results[k] = pp.uncrop(ep[0], ep[1], box, ph.shape)
        if debug:
            print(f'  image {k}: {h}x{w} irls_iters={inf[1]} cg_iters={inf[2]} status={inf[0]} valid={inf[3]}')
    if debug:
        t3 = time.time()
        print(f'TIMING irls_unwrap_2d_batch ({len(jobs)} images) on {dev}: upload {(t1 - t0) * 1e3:.1f} ms, '
              f'solve+download {(t2 - t1) * 1e3:.1f} ms, epilogue {(t3 - t2) * 1e3:.1f} ms')


def _run_group(jobs, ws, dev, shapes, has_w, max_iter, tol, cg_max_iter, cg_tol, epsilon, conncomp_size, graph,
               epilogue):
    """Upload, solve and (device epilogue) post-process one group on `ws`; returns the per-image
    epilogue results (device) or the downloaded u arena (host epilogue)."""
    ph_t = pack([j[1][j[4][0]:j[4][1], j[4][2]:j[4][3]] for j in jobs], ws, dev)
    w_t = None
    if any(has_w):
        wl = []
        for j, hw in zip(jobs, has_w):
            r0, r1, c0, c1 = j[4]
            if hw:
                wl.append(j[2][r0:r1, c0:c1])
            else:  # uniform weight 1 - nan (utils_unwrap2d.py:1239-1241) via weight 1 on valid pixels
                wl.append(np.ones((r1 - r0, c1 - c0), np.float32))
        w_t = pack(wl, ws, dev)
    t1 = time.time()
    u_t, info = solve_device(ph_t, w_t, shapes, max_iter, tol, cg_max_iter, cg_tol, epsilon,
                             graph=graph, workspace=ws)
    info_h = info.cpu().numpy().copy()
    eps, u_host = None, None
    if epilogue == 'device':
        # per-pixel epilogue work on the device; only the images come back to the host
        eps = [epilogue_device(ph_t, u_t, ws.offsets[idx], ws.pitches[idx], h, w, conncomp_size)
               for idx, (h, w) in enumerate(shapes)]
    else:
        u_host = torch.empty(ws.total, dtype=torch.float32, pin_memory=True)
        u_host.copy_(u_t, non_blocking=True)
    torch.cuda.synchronize(dev)
    t2 = time.time()
    return eps, info_h, u_host, t1, t2


def irls_unwrap_2d(phase, weight=None, device='auto', max_iter=50, tol=1e-3,
                   cg_max_iter=20, cg_tol=1e-4, epsilon=1e-2, conncomp_size=30, debug=False):
    """Unwrap 2D phase with the DCT-preconditioned IRLS L1 solver on an MI355X.

    Same contract as insardev.utils_unwrap2d.irls_unwrap_2d (utils_unwrap2d.py:1082-1766):
    returns (unwrapped float32[H, W], conncomp uint16[H, W]); ValueError for non-2D input;
    all-NaN input gives (NaN, zeros); a non-finite solve warns (RuntimeWarning) and returns
    (NaN, zeros).
    """
    if debug:
        print(f"DEBUG irls_unwrap_2d: input phase.shape={phase.shape}, "
              f"weight.shape={weight.shape if weight is not None else None}")
    if phase.ndim != 2:
        raise ValueError(f"irls_unwrap_2d expects 2D phase array, got shape {phase.shape}")
    return irls_unwrap_2d_batch([phase], None if weight is None else [weight], device=device, max_iter=max_iter,
                                tol=tol, cg_max_iter=cg_max_iter, cg_tol=cg_tol, epsilon=epsilon,
                                conncomp_size=conncomp_size, debug=debug)[0]

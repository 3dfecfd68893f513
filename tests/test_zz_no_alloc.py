"""GPU test of irls2d_solve_static (the entry point that allocates no device memory): through the
C ABI, its output and info equal irls2d_solve's with a caller workspace, bitwise, on two calls in
a row (the second on a smaller batch, so the static scratch holds the first call's contents), and
shapes beyond the static scratch are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0, epsilon=1e-2)


def _inputs(shapes, seed0, dev):
    from core_amd.unwrap2d import Workspace, pack
    ims = [S.generate(S.Workload(h, w, seed0 + i, ramp_max=30.0, fault=i == 0, mask_frac=0.05))[:2]
           for i, (h, w) in enumerate(shapes)]
    ws = Workspace(shapes, dev)
    return ws, pack([a for a, _ in ims], ws, dev), pack([b for _, b in ims], ws, dev)


def _static(shapes, ph, co, out, info):
    from core_amd import _lib
    L = _lib.load()
    n, H, W = _lib.shapes_arrays(shapes)
    stream = torch.cuda.current_stream().cuda_stream
    return L.irls2d_solve_static(n, H, W, ph.data_ptr(), co.data_ptr(), out.data_ptr(), FIXED["max_iter"],
                                 FIXED["tol"], FIXED["cg_max_iter"], FIXED["cg_tol"], FIXED["epsilon"],
                                 info.data_ptr(), stream)


def _images(u, ws):
    from core_amd.unwrap2d import image_view
    return [image_view(u, ws, k).cpu().numpy() for k in range(len(ws.shapes))]


@pytest.mark.parametrize("shapes,seed0", [([(300, 257), (96, 130)], 40), ([(64, 48)], 50)])
def test_static_equals_workspace_solve(hip, shapes, seed0):
    from core_amd import _lib, solve_device
    print(f"solve_static {shapes}", flush=True)
    assert _lib.workspace_size(shapes) <= _lib.STATIC_SCRATCH_BYTES
    ws, ph, co = _inputs(shapes, seed0, hip)
    u_ref, info_ref = solve_device(ph, co, shapes, workspace=ws, **FIXED)
    u_ref = _images(u_ref, ws)
    info_ref = info_ref.cpu().numpy().copy()
    for rep in range(2):
        out = torch.full((ws.total,), float("nan"), device=hip)
        info = torch.full((4 * len(shapes),), -1, dtype=torch.int32, device=hip)
        _lib.check(_static(shapes, ph, co, out, info), "irls2d_solve_static")
        torch.cuda.synchronize()
        for a, b in zip(_images(out, ws), u_ref):
            assert np.array_equal(a, b, equal_nan=True), rep
        assert np.array_equal(info.view(-1, 4).cpu().numpy(), info_ref)
    ws.release()


def test_static_rejects_shapes_beyond_scratch(hip):
    from core_amd import _lib
    shapes = [(2048, 2048)]
    assert _lib.workspace_size(shapes) > _lib.STATIC_SCRATCH_BYTES
    ph = torch.zeros(8, device=hip)
    rc = _static(shapes, ph, ph, ph, torch.zeros(4, dtype=torch.int32, device=hip))
    assert rc == _lib.EWORKSPACE
    assert b"static scratch" in _lib.load().irls2d_last_error()

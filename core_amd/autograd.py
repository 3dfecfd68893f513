"""Differentiable unwrap: torch.autograd.Function around the gfx950 solver and its adjoint.

forward(phase, weight) -> u_c, the solver output centred by the masked mean (the reference's
centring, utils_unwrap2d.py:1508-1513, 1538-1541), for a batch [B, H, W] of equal-shape
interferograms resident on the HIP device.  Invalid (NaN) pixels of u_c carry the solver's
weakly coupled values; mask them in the loss.

backward differentiates the converged IRLS fixed point implicitly (no unrolling): it re-solves
(D^T W' D) lambda = C^T dL/du_c with the same DCT-preconditioned CG and returns dL/dphase and
dL/dweight (include/irls2d.h: irls2d_adjoint).  Only the last iterate u* and the inputs are
saved for backward.
"""
from __future__ import annotations

import torch

from . import _lib
from .unwrap2d import Workspace, image_view, solve_device


def _to_arena(x, ws):
    a = torch.zeros(ws.total, dtype=torch.float32, device=x.device)
    for k in range(x.shape[0]):
        image_view(a, ws, k).copy_(x[k])
    return a


def _from_arena(a, ws, B):
    return torch.stack([image_view(a, ws, k) for k in range(B)])


def adjoint_device(phase, weight, u, grad, epsilon=1e-2, cg_iters=200, cg_tol=0.0):
    """The implicit adjoint at a given fixed point: [B, H, W] HIP tensors phase, weight (or None),
    u (the solver output u*, uncentred) and grad = dL/d(C u) in; (dL/dphase, dL/dweight or None)
    out (include/irls2d.h: irls2d_adjoint)."""
    B, H, W = phase.shape
    ws = Workspace.get([(H, W)] * B, phase.device)
    ph = _to_arena(phase.float(), ws)
    wt = _to_arena(weight.float(), ws) if weight is not None else None
    ua = _to_arena(u.float(), ws)
    ga = _to_arena(grad.contiguous().float(), ws)
    gph = torch.empty(ws.total, dtype=torch.float32, device=phase.device)
    gw = torch.empty(ws.total, dtype=torch.float32, device=phase.device) if weight is not None else None
    n, Hs, Ws = _lib.shapes_arrays(ws.shapes)
    rc = _lib.load().irls2d_adjoint(
        n, Hs, Ws, ph.data_ptr(), wt.data_ptr() if wt is not None else None, ua.data_ptr(), ga.data_ptr(),
        gph.data_ptr(), gw.data_ptr() if gw is not None else None, float(epsilon), int(cg_iters), float(cg_tol),
        ws.buf.data_ptr(), ws.nbytes, ws.info.data_ptr(), torch.cuda.current_stream(phase.device).cuda_stream, 0)
    _lib.check(rc, "irls2d_adjoint")
    return _from_arena(gph, ws, B), (_from_arena(gw, ws, B) if gw is not None else None)


class IrlsUnwrap2d(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phase, weight, max_iter=10, cg_max_iter=50, epsilon=1e-2, adj_cg_iter=200, adj_cg_tol=0.0,
                graph=False):
        if phase.dim() != 3:
            raise ValueError("phase must be [B, H, W]")
        B, H, W = phase.shape
        ws = Workspace.get([(H, W)] * B, phase.device)
        ph = _to_arena(phase.detach().float(), ws)
        wt = _to_arena(weight.detach().float(), ws) if weight is not None else None
        u, _ = solve_device(ph, wt, ws.shapes, max_iter=max_iter, tol=0.0, cg_max_iter=cg_max_iter, cg_tol=0.0,
                            epsilon=epsilon, graph=graph, workspace=ws)
        u = _from_arena(u, ws, B)
        valid = ~torch.isnan(phase.detach())
        cnt = valid.sum(dim=(1, 2), keepdim=True).clamp_min(1)
        mean = (torch.where(valid, u, torch.zeros_like(u))).sum(dim=(1, 2), keepdim=True) / cnt
        u_c = u - mean
        ctx.save_for_backward(phase.detach(), weight.detach() if weight is not None else None, u)
        ctx.params = (epsilon, adj_cg_iter, adj_cg_tol)
        ctx.has_weight = weight is not None
        return u_c

    @staticmethod
    def backward(ctx, grad_out):
        phase, weight, u = ctx.saved_tensors
        epsilon, iters, tol = ctx.params
        g_phase, g_weight = adjoint_device(phase, weight if ctx.has_weight else None, u, grad_out, epsilon, iters,
                                           tol)
        return g_phase, g_weight, None, None, None, None, None, None


def irls_unwrap_2d_diff(phase, weight=None, max_iter=10, cg_max_iter=50, epsilon=1e-2, adj_cg_iter=200,
                        adj_cg_tol=0.0, graph=False):
    """Differentiable batched unwrap: [B, H, W] HIP tensors in, centred unwrapped phase out."""
    return IrlsUnwrap2d.apply(phase, weight, max_iter, cg_max_iter, epsilon, adj_cg_iter, adj_cg_tol, graph)

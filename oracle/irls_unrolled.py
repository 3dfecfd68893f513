"""Unrolled, differentiable float64 torch restatement of the reference solver (test infrastructure).

The reference solver (insardev/insardev/utils_unwrap2d.py:1247-1632) written with out-of-place
torch operations in float64 so that autograd differentiates through every IRLS and PCG
iteration: targets (:1255-1262), edge weights (:1268-1275), eigenvalues (:1284-1289), the DCT
preconditioner (:1390-1403, the same Makhoul DCT pair as oracle.irls_torch32), PCG (:1405-1474)
with the clamped alpha/beta, the L2 start (:1476-1513) and the IRLS loop (:1529-1632).  One
deliberate difference: the reference turns alpha and beta into Python floats (`.item()`), which
would cut their dependence on the inputs out of the gradient; here they stay tensors.  The break
tests run on detached values, as in the reference.

With enough iterations to converge on a small grid, the gradient of the centred output through
this unrolled solve is the derivative of the IRLS fixed point, which the implicit adjoint
(oracle.adjoint_fp64, irls2d_adjoint) computes without unrolling.
"""
from __future__ import annotations

import math

import torch

from .irls_torch32 import dct, idct


def _wrap(x):
    return torch.atan2(torch.sin(x), torch.cos(x))


def _fwd_x(u):  # forward difference along columns, 0 in the last column
    return torch.cat([u[:, 1:] - u[:, :-1], torch.zeros_like(u[:, :1])], dim=1)


def _fwd_y(u):
    return torch.cat([u[1:, :] - u[:-1, :], torch.zeros_like(u[:1, :])], dim=0)


def _div(ex, ey):
    """D^T of the forward differences: out = -(ex_{j-1} - ex_j) ... (Neumann), as in :1378-1387."""
    ex = torch.cat([ex[:, :-1], torch.zeros_like(ex[:, :1])], dim=1)
    ey = torch.cat([ey[:-1, :], torch.zeros_like(ey[:1, :])], dim=0)
    left = torch.cat([torch.zeros_like(ex[:, :1]), ex[:, :-1]], dim=1)
    up = torch.cat([torch.zeros_like(ey[:1, :]), ey[:-1, :]], dim=0)
    return left - ex + up - ey


def solve(phase, weight, valid, max_iter=60, cg_max_iter=200, epsilon=0.5, cg_tol=0.0):
    """phase, weight: float64 [H, W] tensors (weight already filled and clipped: the caller does
    the clip so that its derivative is part of the graph); valid: bool [H, W].  Returns the
    centred final iterate u - mean_valid(u)."""
    H, W = phase.shape
    vf = valid.to(phase.dtype)
    tx = _wrap(_fwd_x(phase))
    ty = _wrap(_fwd_y(phase))
    tx = torch.cat([tx[:, :-1], torch.zeros_like(tx[:, :1])], dim=1)
    ty = torch.cat([ty[:-1, :], torch.zeros_like(ty[:1, :])], dim=0)
    vx = torch.cat([vf[:, :-1] * vf[:, 1:], torch.zeros_like(vf[:, :1])], dim=1)
    vy = torch.cat([vf[:-1, :] * vf[1:, :], torch.zeros_like(vf[:1, :])], dim=0)
    wx = torch.cat([(weight[:, :-1] + weight[:, 1:]) / 2, torch.zeros_like(weight[:, :1])], dim=1) * vx
    wy = torch.cat([(weight[:-1, :] + weight[1:, :]) / 2, torch.zeros_like(weight[:1, :])], dim=0) * vy
    ii = torch.arange(H, dtype=torch.float64)
    jj = torch.arange(W, dtype=torch.float64)
    lam = 2 * torch.cos(math.pi * ii / H)[:, None] + 2 * torch.cos(math.pi * jj / W)[None, :] - 4
    lam[0, 0] = 1.0
    dc = torch.ones_like(lam)
    dc[0, 0] = 0.0
    eps_sq = epsilon * epsilon

    def dct2(x):
        return dct(dct(x).T).T

    def idct2(X):
        return idct(idct(X.T).T)

    def precond(r):
        return idct2(-dct2(r) / lam * dc)

    def lap(x, wxi, wyi):
        return _div(wxi * _fwd_x(x), wyi * _fwd_y(x))

    def cg(b, wxi, wyi, x):
        r = b - lap(x, wxi, wyi)
        if not bool(torch.isfinite(r).all()):
            return x
        z = precond(r)
        p = z
        rz = torch.sum(r * z)
        for _ in range(cg_max_iter):
            Ap = lap(p, wxi, wyi)
            pAp = torch.sum(p * Ap)
            pv = float(pAp.detach())
            if abs(pv) < 1e-15 or not math.isfinite(pv):
                break
            alpha = torch.clamp(rz / pAp, -1e6, 1e6)
            x = x + alpha * p
            r = r - alpha * Ap
            rn = float(torch.sqrt(torch.sum(r * r)).detach())
            if rn < cg_tol or not math.isfinite(rn):
                break
            z = precond(r)
            rz_new = torch.sum(r * z)
            if abs(float(rz.detach())) < 1e-15:
                break
            beta = torch.clamp(rz_new / rz, -1e6, 1e6)
            p = z + beta * p
            rz = rz_new
        return x

    vc = vf.sum()

    def centre(u):
        return u - (u * vf).sum() / vc

    u = centre(precond(_div(wx * tx, wy * ty)))
    for _ in range(max_iter):
        u = centre(u)
        rx = _fwd_x(u) - tx
        ry = _fwd_y(u) - ty
        wxi = torch.clamp(wx / torch.sqrt(rx * rx + eps_sq), 1e-6, 1e6)
        wyi = torch.clamp(wy / torch.sqrt(ry * ry + eps_sq), 1e-6, 1e6)
        # the last column / row edges carry weight wx = 0 -> clamped to 1e-6 with target 0, as in
        # the reference (their difference is not used by lap or the right-hand side)
        b = _div(wxi * tx, wyi * ty)
        u = cg(b, wxi, wyi, u)
    return centre(u)

"""torch-CPU float32 restatement of the reference solver, op for op (test infrastructure).

This is the "reference CPU path": the reference's own function
(insardev/insardev/utils_unwrap2d.py:1082-1766) cannot be imported here
(numba, torch_dct and cv2 are absent), so its float32 torch operations are
restated in the same order on the CPU device.  It serves two purposes:
  * the integer cycle count k = round((phi - psi) / 2 pi) of the reference's
    float32 trajectory, against which the GPU result is compared;
  * the CPU timing baseline.

The DCT pair restates the published algorithm of torch-dct (unpinned,
insardev/setup.py:49; called at utils_unwrap2d.py:1334): Makhoul's DCT-II via
one complex FFT of the even/odd reordered sequence, scaled by 2, and the inverse
via a Hermitian irfft, with the twiddles formed in the input dtype.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import post


def dct(x: torch.Tensor) -> torch.Tensor:
    """Unnormalised DCT-II along the last dim: X_k = 2 sum_n x_n cos(pi k (2n+1) / 2N)."""
    shape = x.shape
    N = shape[-1]
    x = x.contiguous().view(-1, N)
    v = torch.cat([x[:, ::2], x[:, 1::2].flip([1])], dim=1)
    Vc = torch.view_as_real(torch.fft.fft(v, dim=1))
    k = -torch.arange(N, dtype=x.dtype, device=x.device) * math.pi / (2 * N)
    Wr = torch.cos(k)
    Wi = torch.sin(k)
    V = Vc[:, :, 0] * Wr - Vc[:, :, 1] * Wi
    return (2 * V).view(*shape)


def idct(X: torch.Tensor) -> torch.Tensor:
    """Exact inverse of `dct` along the last dim."""
    shape = X.shape
    N = shape[-1]
    Xv = X.contiguous().view(-1, N) / 2
    k = torch.arange(N, dtype=X.dtype, device=X.device) * math.pi / (2 * N)
    Wr = torch.cos(k)
    Wi = torch.sin(k)
    Vtr = Xv
    Vti = torch.cat([Xv[:, :1] * 0, -Xv.flip([1])[:, :-1]], dim=1)
    Vr = Vtr * Wr - Vti * Wi
    Vi = Vtr * Wi + Vti * Wr
    V = torch.complex(Vr, Vi)
    v = torch.fft.irfft(V, n=N, dim=1)
    x = v.new_zeros(v.shape)
    x[:, ::2] += v[:, :N - (N // 2)]
    x[:, 1::2] += v.flip([1])[:, :N // 2]
    return x.view(*shape)


def solve(phi_np, w_np, valid_np, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4, epsilon=1e-2,
          stats=None, dtype=torch.float32, device='cpu', inject=None):
    """utils_unwrap2d.py:1247-1632 on a filled, cropped grid, float32 on the CPU device.

    `dtype` / `device` run the same torch operations in another precision or on a torch device
    (the reference itself runs them on whatever device torch is given): float64 gives an FP64
    trajectory of the same algorithm for grids too large for the numpy oracle.  `inject` = (IRLS
    iteration, PCG iteration): p[0, 0] = +inf once alpha is formed there (the library's test hook
    irls2d_debug_inject)."""
    phi = torch.from_numpy(np.ascontiguousarray(phi_np, np.float32)).to(device=device, dtype=dtype)
    w = torch.from_numpy(np.ascontiguousarray(w_np, np.float32)).to(device=device, dtype=dtype)
    valid = torch.from_numpy(np.ascontiguousarray(valid_np, bool)).to(device)
    H, W = phi.shape

    dx_t = torch.zeros_like(phi)
    dy_t = torch.zeros_like(phi)
    dx_t[:, :-1] = phi[:, 1:] - phi[:, :-1]
    dy_t[:-1, :] = phi[1:, :] - phi[:-1, :]
    torch.atan2(torch.sin(dx_t), torch.cos(dx_t), out=dx_t)
    torch.atan2(torch.sin(dy_t), torch.cos(dy_t), out=dy_t)
    del phi

    wx = torch.zeros_like(w)
    wy = torch.zeros_like(w)
    wx[:, :-1] = (w[:, :-1] + w[:, 1:]) / 2
    wy[:-1, :] = (w[:-1, :] + w[1:, :]) / 2
    wx[:, :-1] *= (valid[:, :-1] & valid[:, 1:]).to(dtype)
    wy[:-1, :] *= (valid[:-1, :] & valid[1:, :]).to(dtype)
    del w

    ii = torch.arange(H, dtype=torch.float64, device=device)
    jj = torch.arange(W, dtype=torch.float64, device=device)
    lam = (2 * torch.cos(torch.pi * ii / H).unsqueeze(1) + 2 * torch.cos(torch.pi * jj / W).unsqueeze(0) - 4).to(dtype)
    lam[0, 0] = 1.0

    eps_sq = epsilon * epsilon

    def dct2(x):
        return dct(dct(x).T.contiguous()).T.contiguous()

    def idct2(X):
        return idct(idct(X.T.contiguous()).T.contiguous())

    def lap(x, wxi, wyi):
        dx = torch.zeros_like(x)
        dy = torch.zeros_like(x)
        dx[:, :-1] = x[:, 1:] - x[:, :-1]
        dy[:-1, :] = x[1:, :] - x[:-1, :]
        dx.mul_(wxi)
        dy.mul_(wyi)
        out = torch.zeros_like(x)
        out[:, 1:].sub_(dx[:, :-1])
        out[:, :-1].add_(dx[:, :-1])
        out[1:, :].sub_(dy[:-1, :])
        out[:-1, :].add_(dy[:-1, :])
        return out.neg_()

    def precond(r):
        X = dct2(r)
        X.div_(lam)
        X.neg_()
        X[0, 0] = 0.0
        return idct2(X)

    def cg(b, wxi, wyi, x, inject_cg=-1):
        n = 0
        r = b.clone()
        r.sub_(lap(x, wxi, wyi))
        if not torch.isfinite(r).all():
            return x, 0
        z = precond(r)
        p = z.clone()
        rz = torch.dot(r.view(-1), z.view(-1))
        for i in range(cg_max_iter):
            n = i + 1
            Ap = lap(p, wxi, wyi)
            pAp = torch.dot(p.view(-1), Ap.view(-1))
            if pAp.abs() < 1e-15 or not torch.isfinite(pAp):
                break
            alpha = torch.clamp(rz / pAp, -1e6, 1e6).item()
            if i == inject_cg:
                p[0, 0] = float('inf')
            x.add_(p, alpha=alpha)
            r.sub_(Ap, alpha=alpha)
            if not torch.isfinite(x).all():
                break
            rn = r.norm()
            if rn < cg_tol or not torch.isfinite(rn):
                break
            z = precond(r)
            rz_new = torch.dot(r.view(-1), z.view(-1))
            if rz.abs() < 1e-15:
                break
            beta = torch.clamp(rz_new / rz, -1e6, 1e6)
            p.mul_(beta.item()).add_(z)
            rz = rz_new
        return x, n

    rho = torch.zeros_like(dx_t)
    rho[:, 1:] += wx[:, :-1] * dx_t[:, :-1]
    rho[:, :-1] -= wx[:, :-1] * dx_t[:, :-1]
    rho[1:, :] += wy[:-1, :] * dy_t[:-1, :]
    rho[:-1, :] -= wy[:-1, :] * dy_t[:-1, :]
    u = precond(rho)
    if not torch.isfinite(u).all():
        u.nan_to_num_(nan=0.0, posinf=0.0, neginf=0.0)
    vf = valid.to(dtype)
    vc = vf.sum()
    if vc > 0:
        u.sub_((u * vf).sum() / vc)

    u_best = u.clone()
    best = float('inf')
    n_it = 0
    for it in range(max_iter):
        n_it = it + 1
        u_prev = u.clone()
        if vc > 0:
            u.sub_((u * vf).sum() / vc)
        rx = torch.zeros_like(u)
        ry = torch.zeros_like(u)
        rx[:, :-1] = u[:, 1:] - u[:, :-1]
        ry[:-1, :] = u[1:, :] - u[:-1, :]
        rx.sub_(dx_t)
        ry.sub_(dy_t)
        res = (torch.dot(rx.view(-1), rx.view(-1)) + torch.dot(ry.view(-1), ry.view(-1))).item()
        if res < best and torch.isfinite(u).all():
            best = res
            u_best.copy_(u)
        wxi = torch.clamp(wx / torch.sqrt(rx * rx + eps_sq), 1e-6, 1e6)
        wyi = torch.clamp(wy / torch.sqrt(ry * ry + eps_sq), 1e-6, 1e6)
        b = torch.zeros_like(u)
        b[:, 1:].addcmul_(wxi[:, :-1], dx_t[:, :-1])
        b[:, :-1].addcmul_(wxi[:, :-1], dx_t[:, :-1], value=-1)
        b[1:, :].addcmul_(wyi[:-1, :], dy_t[:-1, :])
        b[:-1, :].addcmul_(wyi[:-1, :], dy_t[:-1, :], value=-1)
        del rx, ry
        u, _ = cg(b, wxi, wyi, u, inject[1] if inject is not None and inject[0] == it else -1)
        if not torch.isfinite(u).all():
            u.copy_(u_best)
            break
        u_prev.sub_(u)
        rel = (torch.norm(u_prev) / (torch.norm(u) + 1e-10)).item()
        if rel < tol:
            break
    if not torch.isfinite(u).all():
        u = u_best
    if stats is not None:
        stats['irls_iters'] = n_it
    out = u.cpu().numpy()
    return out.astype(np.float32) if dtype == torch.float32 else out


def irls_unwrap_2d(phase, weight=None, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4,
                   epsilon=1e-2, conncomp_size=30, return_raw=False, dtype=torch.float32, device='cpu',
                   inject=None):
    """The reference call on the CPU device (utils_unwrap2d.py:1082-1766); `dtype` / `device` as
    in `solve` (the epilogue always runs on the host, as the reference's does)."""
    if phase.ndim != 2:
        raise ValueError(f"irls_unwrap_2d expects 2D phase array, got shape {phase.shape}")
    pro = post.prologue(phase, weight)
    if pro is None:
        out = (np.full_like(phase, np.nan), np.zeros(phase.shape, np.uint16))
        return out + (None,) if return_raw else out
    crop, ph_c, nm, pf, wf = pro
    u = solve(pf, wf, ~nm, max_iter, tol, cg_max_iter, cg_tol, epsilon, dtype=dtype, device=device, inject=inject)
    ep = post.epilogue(u, ph_c, nm, conncomp_size)
    if ep is None:
        out = (np.full(phase.shape, np.nan, np.float32), np.zeros(phase.shape, np.uint16))
        return out + (u,) if return_raw else out
    unw, cc = post.uncrop(ep[0], ep[1], crop, phase.shape)
    return (unw, cc, u) if return_raw else (unw, cc)

"""FP64 numpy/scipy restatement of the reference's DCT-preconditioned IRLS solver.

Test infrastructure (the accuracy oracle).  Every step follows
insardev/insardev/utils_unwrap2d.py, cited per function.  The DCT is
scipy.fft.dct/idct(type=2, norm=None), which is the unnormalised DCT-II
X_k = 2 sum_n x_n cos(pi k (2n+1) / 2N) of the third-party torch-dct package the
reference calls (utils_unwrap2d.py:1334; torch-dct is unpinned in
insardev/setup.py:49) and its exact inverse.  The preconditioner
IDCT2(DCT2(r) / -lambda) does not depend on the DCT scaling as long as the
inverse is exact, so any exact DCT pair gives the same operator.

The solver works in `dtype` (float64 by default); `solve` returns the raw solver
output u (before the host epilogue) plus a trace of the control flow.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from . import post


def wrap_atan2(d):
    """wrap(d) = atan2(sin d, cos d) (utils_unwrap2d.py:1261-1262)."""
    return np.arctan2(np.sin(d), np.cos(d))


def targets(phi, dtype=np.float64):
    """Wrapped forward differences, zero on the last column/row (utils_unwrap2d.py:1255-1262)."""
    phi = phi.astype(dtype)
    tx = np.zeros_like(phi)
    ty = np.zeros_like(phi)
    tx[:, :-1] = phi[:, 1:] - phi[:, :-1]
    ty[:-1, :] = phi[1:, :] - phi[:-1, :]
    return wrap_atan2(tx), wrap_atan2(ty)


def edge_weights(w, valid, dtype=np.float64):
    """Edge weights (w_j + w_j+1)/2 on edges between two valid pixels (utils_unwrap2d.py:1268-1275)."""
    w = w.astype(dtype)
    wx = np.zeros_like(w)
    wy = np.zeros_like(w)
    wx[:, :-1] = (w[:, :-1] + w[:, 1:]) / 2
    wy[:-1, :] = (w[:-1, :] + w[1:, :]) / 2
    wx[:, :-1] *= (valid[:, :-1] & valid[:, 1:]).astype(dtype)
    wy[:-1, :] *= (valid[:-1, :] & valid[1:, :]).astype(dtype)
    return wx, wy


def eigenvalues(H, W, dtype=np.float64):
    """lambda_ij = 2 cos(pi i/H) + 2 cos(pi j/W) - 4 in float64, lambda_00 = 1 (utils_unwrap2d.py:1284-1289)."""
    ci = np.cos(np.pi * np.arange(H, dtype=np.float64) / H)
    cj = np.cos(np.pi * np.arange(W, dtype=np.float64) / W)
    lam = (2 * ci[:, None] + 2 * cj[None, :] - 4).astype(dtype)
    lam[0, 0] = 1.0
    return lam


def dct2(x):
    """Row DCT-II then column DCT-II (utils_unwrap2d.py:1337-1350)."""
    return sfft.dct(sfft.dct(x, type=2, axis=1), type=2, axis=0)


def idct2(X):
    """Column inverse then row inverse (utils_unwrap2d.py:1352-1365)."""
    return sfft.idct(sfft.idct(X, type=2, axis=0), type=2, axis=1)


def laplacian(x, wx, wy):
    """A x = D^T W D x with forward differences, Neumann borders (utils_unwrap2d.py:1367-1388)."""
    dx = np.zeros_like(x)
    dy = np.zeros_like(x)
    dx[:, :-1] = x[:, 1:] - x[:, :-1]
    dy[:-1, :] = x[1:, :] - x[:-1, :]
    dx *= wx
    dy *= wy
    out = np.zeros_like(x)
    out[:, 1:] -= dx[:, :-1]
    out[:, :-1] += dx[:, :-1]
    out[1:, :] -= dy[:-1, :]
    out[:-1, :] += dy[:-1, :]
    return -out


def divergence_rhs(tx, ty, wx, wy):
    """b = D^T W t (utils_unwrap2d.py:1576-1581, and rho at :1478-1483)."""
    b = np.zeros_like(tx)
    b[:, 1:] += wx[:, :-1] * tx[:, :-1]
    b[:, :-1] -= wx[:, :-1] * tx[:, :-1]
    b[1:, :] += wy[:-1, :] * ty[:-1, :]
    b[:-1, :] -= wy[:-1, :] * ty[:-1, :]
    return b


def precondition(r, lam):
    """z = IDCT2(-DCT2(r) / lambda) with the DC term zeroed (utils_unwrap2d.py:1390-1403)."""
    X = dct2(r)
    X = -(X / lam)
    X[0, 0] = 0.0
    return idct2(X)


def pcg(b, wx, wy, x, lam, max_iter, tol, inject_cg=-1):
    """PCG with the reference's clamps and break tests (utils_unwrap2d.py:1405-1474).

    Updates x in place; returns (x, n_iters, reason).  `inject_cg` mirrors the library's test hook
    (irls2d_debug_inject): in that iteration p[0, 0] = +inf once alpha is formed.
    """
    r = b - laplacian(x, wx, wy)
    if not np.isfinite(r).all():
        return x, 0, 'r0_nonfinite'
    z = precondition(r, lam)
    p = z.copy()
    rz = float(np.dot(r.ravel(), z.ravel()))
    n = 0
    reason = 'max_iter'
    for i in range(max_iter):
        n = i + 1
        Ap = laplacian(p, wx, wy)
        pAp = float(np.dot(p.ravel(), Ap.ravel()))
        if abs(pAp) < 1e-15 or not np.isfinite(pAp):
            reason = 'pAp'
            break
        alpha = float(np.clip(rz / pAp, -1e6, 1e6))
        if i == inject_cg:
            p[0, 0] = np.inf
        x += alpha * p
        r -= alpha * Ap
        if not np.isfinite(x).all():
            reason = 'x_nonfinite'
            break
        rn = float(np.linalg.norm(r))
        if rn < tol or not np.isfinite(rn):
            reason = 'rnorm'
            break
        z = precondition(r, lam)
        rz_new = float(np.dot(r.ravel(), z.ravel()))
        if abs(rz) < 1e-15:
            reason = 'rz'
            break
        beta = float(np.clip(rz_new / rz, -1e6, 1e6))
        p = z + beta * p
        rz = rz_new
    return x, n, reason


def solve(phi, w, valid, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4, epsilon=1e-2,
          dtype=np.float64, trace=None, inject=None):
    """The device part of irls_unwrap_2d (utils_unwrap2d.py:1251-1632) on a filled, cropped grid.

    phi, w: filled float arrays; valid: bool.  Returns u (dtype).
    `trace`, when a list, receives one dict per IRLS iteration.  `inject` = (IRLS iteration, PCG
    iteration) of the test hook (see `pcg`).
    """
    H, W = phi.shape
    tx, ty = targets(phi, dtype)
    wx, wy = edge_weights(w, valid, dtype)
    lam = eigenvalues(H, W, dtype)
    eps_sq = epsilon * epsilon

    rho = divergence_rhs(tx, ty, wx, wy)
    u = precondition(rho, lam)
    if not np.isfinite(u).all():
        u = np.nan_to_num(u, nan=0.0, posinf=0.0, neginf=0.0)
    vf = valid.astype(dtype)
    vc = vf.sum()
    if vc > 0:
        u -= (u * vf).sum() / vc

    u_best = u.copy()
    best = np.inf
    for it in range(max_iter):
        u_prev = u.copy()
        if vc > 0:
            u -= (u * vf).sum() / vc
        rx = np.zeros_like(u)
        ry = np.zeros_like(u)
        rx[:, :-1] = u[:, 1:] - u[:, :-1]
        ry[:-1, :] = u[1:, :] - u[:-1, :]
        rx -= tx
        ry -= ty
        res = float(np.dot(rx.ravel(), rx.ravel()) + np.dot(ry.ravel(), ry.ravel()))
        improved = res < best and np.isfinite(u).all()
        if improved:
            best = res
            u_best = u.copy()
        wxi = np.clip(wx / np.sqrt(rx * rx + eps_sq), 1e-6, 1e6)
        wyi = np.clip(wy / np.sqrt(ry * ry + eps_sq), 1e-6, 1e6)
        b = divergence_rhs(tx, ty, wxi, wyi)
        u, n_cg, reason = pcg(b, wxi, wyi, u, lam, cg_max_iter, cg_tol,
                              inject[1] if inject is not None and inject[0] == it else -1)
        if not np.isfinite(u).all():
            u = u_best.copy()
            if trace is not None:
                trace.append(dict(it=it, residual=res, cg=n_cg, reason=reason, stop='nonfinite'))
            break
        rel = float(np.linalg.norm(u_prev - u) / (np.linalg.norm(u) + 1e-10))
        if trace is not None:
            trace.append(dict(it=it, residual=res, cg=n_cg, reason=reason, rel=rel, improved=bool(improved)))
        if rel < tol:
            break
    if not np.isfinite(u).all():
        u = u_best
    return u


def irls_unwrap_2d(phase, weight=None, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4,
                   epsilon=1e-2, conncomp_size=30, dtype=np.float64, return_raw=False, inject=None):
    """Whole reference call (utils_unwrap2d.py:1082-1766) with the solver in `dtype`."""
    if phase.ndim != 2:
        raise ValueError(f"irls_unwrap_2d expects 2D phase array, got shape {phase.shape}")
    pro = post.prologue(phase, weight)
    if pro is None:
        out = (np.full_like(phase, np.nan), np.zeros(phase.shape, np.uint16))
        return out + (None,) if return_raw else out
    crop, ph_c, nm, pf, wf = pro
    u = solve(pf, wf, ~nm, max_iter, tol, cg_max_iter, cg_tol, epsilon, dtype, inject=inject)
    ep = post.epilogue(u, ph_c, nm, conncomp_size)
    if ep is None:
        out = (np.full(phase.shape, np.nan, np.float32), np.zeros(phase.shape, np.uint16))
        return out + (u,) if return_raw else out
    unw, cc = post.uncrop(ep[0], ep[1], crop, phase.shape)
    return (unw, cc, u) if return_raw else (unw, cc)

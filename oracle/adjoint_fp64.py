"""FP64 restatement of the implicit adjoint of the IRLS fixed point (test infrastructure).

The reference has no adjoint; this restates the mathematics the GPU adjoint implements, on top
of the FP64 solver restatement (oracle/irls_fp64.py, insardev/insardev/utils_unwrap2d.py:1082-1632).

At the IRLS fixed point u* the reference's iteration satisfies G(u) = D^T (w_hat(r) r) = 0 with
r = D u - t(psi) and w_hat = clamp(w / sqrt(r^2 + eps^2), 1e-6, 1e6).  For a loss L(C u*) with
C u = u - mean_valid(u) (the solver's centring):
    H = D^T W' D,  W'_e = w_e eps^2 / (r_e^2 + eps^2)^(3/2) (unclamped), w_hat_e (clamped)
    H lam = C^T dL/d(Cu)                    (mean-free right-hand side)
    dL/dpsi   = valid * (H lam)              (t = wrap(D psi), wrap' = 1 almost everywhere)
                = valid * C^T dL/d(Cu)         (H lam = C^T g exactly: the fixed point is
                                                translation-equivariant in psi)
    dL/dw_e   = -(D lam)_e r_e / sqrt(r_e^2 + eps^2)   (0 on clamped edges)
    dL/dcoh_i = 1[valid_i, 0.01 < coh_i < 1] * sum_{e ~ i} 0.5 m_e dL/dw_e
with m_e the both-valid edge mask (edge weight (w_i + w_j)/2 m_e, pixel weight clip(coh, 0.01, 1)).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import irls_fp64 as O


def diff_ops(H, W):
    def d1(n):
        m = sp.lil_matrix((n, n))
        for k in range(n - 1):
            m[k, k] = -1.0
            m[k, k + 1] = 1.0
        return m.tocsr()
    return sp.kron(sp.identity(H), d1(W)).tocsr(), sp.kron(d1(H), sp.identity(W)).tocsr()


def center(u, valid):
    return u - (u * valid).sum() / valid.sum()


def edge_terms(u, tx, ty, wx, wy, eps):
    """W' and s' = d(w_hat r)/dw on both edge sets; zero on the last column / row."""
    out = []
    for axis, t, w in ((1, tx, wx), (0, ty, wy)):
        r = np.zeros_like(u)
        if axis == 1:
            r[:, :-1] = u[:, 1:] - u[:, :-1]
        else:
            r[:-1, :] = u[1:, :] - u[:-1, :]
        r -= t
        q = r * r + eps * eps
        what = w / np.sqrt(q)
        clamped = (what < 1e-6) | (what > 1e6)
        wp = np.where(clamped, np.clip(what, 1e-6, 1e6), w * eps * eps / (q * np.sqrt(q)))
        sprime = np.where(clamped, 0.0, r / np.sqrt(q))
        if axis == 1:
            wp[:, -1] = 0.0
            sprime[:, -1] = 0.0
        else:
            wp[-1, :] = 0.0
            sprime[-1, :] = 0.0
        out += [wp, sprime]
    return out  # wpx, sx, wpy, sy


def adjoint(phase, weight, u, g, epsilon=1e-2, solver='direct', cg_iters=200, cg_tol=0.0):
    """Gradients (dL/dphase, dL/dweight) for dL/d(Cu) = g on a filled, uncropped grid.

    phase has NaN at invalid pixels; weight may be None.  solver='direct' uses a sparse LU of H
    (grounded at one pixel); solver='pcg' uses the reference's DCT-preconditioned CG from x = 0.
    """
    nan = np.isnan(phase)
    valid = ~nan
    phi = np.where(nan, 0.0, phase).astype(np.float64)
    if weight is not None:
        w = np.where(np.isnan(weight) | nan, 0.0, weight).astype(np.float64)
        w = np.clip(w, 0.01, 1.0)
    else:
        w = (~nan).astype(np.float64)
    tx, ty = O.targets(phi)
    wx, wy = O.edge_weights(w, valid)
    wpx, sx, wpy, sy = edge_terms(u.astype(np.float64), tx, ty, wx, wy, epsilon)
    Hh, Ww = phase.shape
    b = g.astype(np.float64) - g.sum() / valid.sum() * valid
    if solver == 'direct':
        Dx, Dy = diff_ops(Hh, Ww)
        A = (Dx.T @ sp.diags(wpx.ravel()) @ Dx + Dy.T @ sp.diags(wpy.ravel()) @ Dy).tocsc()
        # ground pixel 0 (constant null space); b is mean-free so the system is consistent
        keep = np.arange(1, Hh * Ww)
        lam = np.zeros(Hh * Ww)
        lam[keep] = spla.spsolve(A[keep][:, keep], b.ravel()[keep])
        lam = lam.reshape(Hh, Ww)
    else:
        lam = np.zeros_like(b)
        lam, _, _ = O.pcg(b, wpx, wpy, lam, O.eigenvalues(Hh, Ww), cg_iters, cg_tol)
    gphi = valid * O.laplacian(lam, wpx, wpy)
    dlx = np.zeros_like(lam); dly = np.zeros_like(lam)
    dlx[:, :-1] = lam[:, 1:] - lam[:, :-1]
    dly[:-1, :] = lam[1:, :] - lam[:-1, :]
    mx = np.zeros_like(lam); my = np.zeros_like(lam)
    mx[:, :-1] = valid[:, :-1] & valid[:, 1:]
    my[:-1, :] = valid[:-1, :] & valid[1:, :]
    gwe_x = -dlx * sx * mx * 0.5
    gwe_y = -dly * sy * my * 0.5
    gw = gwe_x.copy() + gwe_y
    gw[:, 1:] += gwe_x[:, :-1]
    gw[1:, :] += gwe_y[:-1, :]
    gweight = None
    if weight is not None:
        raw = np.where(np.isnan(weight) | nan, 0.0, weight)
        gweight = gw * (valid & (raw > 0.01) & (raw < 1.0))
    return gphi, gweight, lam


def fixed_point(phase, weight, u0, epsilon=1e-2, tol=1e-12, max_newton=200):
    """The IRLS fixed point G(u) = D^T (w_hat r) = 0, i.e. the minimiser of the convex
    Phi(u) = sum_{w_e > 0} w_e sqrt(r_e^2 + eps^2) + sum_{w_e = 0} 0.5e-6 r_e^2 (the 1e-6 floor of
    the clamp), by damped Newton (sparse direct solve grounded at pixel 0, backtracking on Phi),
    started from u0.  Smooth in (phase, weight) away from clamp boundaries, so finite differences
    of it check the implicit adjoint."""
    nan = np.isnan(phase)
    valid = ~nan
    phi = np.where(nan, 0.0, phase).astype(np.float64)
    if weight is not None:
        w = np.clip(np.where(np.isnan(weight) | nan, 0.0, weight).astype(np.float64), 0.01, 1.0)
    else:
        w = (~nan).astype(np.float64)
    tx, ty = O.targets(phi)
    wx, wy = O.edge_weights(w, valid)
    Hh, Ww = phase.shape
    Dx, Dy = diff_ops(Hh, Ww)
    keep = np.arange(1, Hh * Ww)
    ex = np.zeros_like(wx); ex[:, :-1] = 1.0   # existing edges
    ey = np.zeros_like(wy); ey[:-1, :] = 1.0
    eps2 = epsilon * epsilon

    def resid(u):
        rx = np.zeros_like(u); ry = np.zeros_like(u)
        rx[:, :-1] = u[:, 1:] - u[:, :-1]
        ry[:-1, :] = u[1:, :] - u[:-1, :]
        return rx - tx, ry - ty

    def phi_obj(u):
        rx, ry = resid(u)
        tot = 0.0
        for r, we, e in ((rx, wx, ex), (ry, wy, ey)):
            tot += np.sum(np.where(we > 0, we * np.sqrt(r * r + eps2), 0.5e-6 * r * r) * e)
        return tot

    def grad(u):
        rx, ry = resid(u)
        gx = np.where(wx > 0, wx * rx / np.sqrt(rx * rx + eps2), 1e-6 * rx) * ex
        gy = np.where(wy > 0, wy * ry / np.sqrt(ry * ry + eps2), 1e-6 * ry) * ey
        return Dx.T @ gx.ravel() + Dy.T @ gy.ravel()

    u = u0.astype(np.float64).copy()
    for _ in range(max_newton):
        G = grad(u)
        if np.abs(G).max() < tol:
            break
        wpx, _, wpy, _ = edge_terms(u, tx, ty, wx, wy, epsilon)
        A = (Dx.T @ sp.diags(wpx.ravel()) @ Dx + Dy.T @ sp.diags(wpy.ravel()) @ Dy).tocsc()
        d = np.zeros(Hh * Ww)
        d[keep] = spla.spsolve(A[keep][:, keep], G[keep])
        d = d.reshape(Hh, Ww)
        f0, slope, t = phi_obj(u), float(G @ d.ravel()), 1.0
        while phi_obj(u - t * d) > f0 - 1e-4 * t * slope and t > 1e-12:
            t *= 0.5
        u = u - t * d
    return u


def chord_solver(phase, weight, ustar, epsilon=1e-2):
    """Fixed points of perturbed inputs near u*: returns F(phase', weight') -> u solving
    G(u; phase', weight') = 0 by chord iterations u <- u - H0^-1 G(u) from u*, with H0 = D^T W' D at
    (phase, weight, u*) factorised once (grounded at pixel 0).  For perturbations of 1e-6 the
    contraction factor is O(1e-6): the iteration reaches the double-precision floor of G in a few
    steps, as the damped-Newton `fixed_point` does, at the cost of triangular solves only."""
    nan = np.isnan(phase)
    valid = ~nan
    Hh, Ww = phase.shape
    Dx, Dy = diff_ops(Hh, Ww)
    keep = np.arange(1, Hh * Ww)
    ex = np.zeros((Hh, Ww)); ex[:, :-1] = 1.0
    ey = np.zeros((Hh, Ww)); ey[:-1, :] = 1.0
    eps2 = epsilon * epsilon

    def setup(ph, wt):
        phi = np.where(nan, 0.0, ph).astype(np.float64)
        if wt is not None:
            w = np.clip(np.where(np.isnan(wt) | nan, 0.0, wt).astype(np.float64), 0.01, 1.0)
        else:
            w = (~nan).astype(np.float64)
        tx, ty = O.targets(phi)
        wx, wy = O.edge_weights(w, valid)
        return tx, ty, wx, wy

    def grad(u, tx, ty, wx, wy):
        rx = np.zeros_like(u); ry = np.zeros_like(u)
        rx[:, :-1] = u[:, 1:] - u[:, :-1]
        ry[:-1, :] = u[1:, :] - u[:-1, :]
        rx -= tx
        ry -= ty
        gx = np.where(wx > 0, wx * rx / np.sqrt(rx * rx + eps2), 1e-6 * rx) * ex
        gy = np.where(wy > 0, wy * ry / np.sqrt(ry * ry + eps2), 1e-6 * ry) * ey
        return Dx.T @ gx.ravel() + Dy.T @ gy.ravel()

    tx, ty, wx, wy = setup(phase, weight)
    wpx, _, wpy, _ = edge_terms(ustar, tx, ty, wx, wy, epsilon)
    A = (Dx.T @ sp.diags(wpx.ravel()) @ Dx + Dy.T @ sp.diags(wpy.ravel()) @ Dy).tocsc()
    lu = spla.splu(A[keep][:, keep].tocsc())

    def F(ph, wt, max_iter=30):
        # iterate to the rounding floor of G: finite differences with steps of 1e-6 over many
        # pixels need the fixed point to ~1e-15, not to a looser tolerance
        t = setup(ph, wt)
        u = ustar.astype(np.float64).copy()
        best = np.inf
        for _ in range(max_iter):
            G = grad(u, *t)
            g = float(np.abs(G[keep]).max())
            if g >= best or g == 0.0:
                break
            best = g
            d = np.zeros(Hh * Ww)
            d[keep] = lu.solve(G[keep])
            u = u - d.reshape(Hh, Ww)
        return u
    return F

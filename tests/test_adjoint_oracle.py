"""The FP64 implicit adjoint (oracle/adjoint_fp64.py) against FP64 central finite differences of
the IRLS fixed-point map (damped-Newton solve), random unit probe directions, step 1e-6."""
import numpy as np
import pytest

from core_amd import synthetic as S
from oracle import adjoint_fp64 as A, irls_fp64 as O, post


@pytest.mark.parametrize("noise,mask", [(0.3, 0.0), (1.2, 0.0), (1.0, 0.15)])
def test_adjoint_matches_finite_differences(noise, mask):
    wl = S.Workload(12, 14, 3, ramp_max=8.0, n_bumps=2, noise=noise)
    psi, coh, _ = S.generate(wl)
    if mask:
        psi[2:4, 5:8] = np.nan
        coh[2:4, 5:8] = np.nan
    psi = psi.astype(np.float64)
    coh = coh.astype(np.float64)
    _, _, nm, pf, wf = post.prologue(psi.astype(np.float32), coh.astype(np.float32))
    v = ~nm
    u0 = O.solve(pf, wf, v, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
    ustar = A.fixed_point(psi, coh, u0)
    F = lambda ps, wt: A.center(A.fixed_point(ps, wt, ustar), v)
    rng = np.random.default_rng(1)
    target = rng.normal(size=psi.shape)
    L = lambda uu: float(np.sum(((uu - target) * v) ** 2))
    u = F(psi, coh)
    gphi, gw, _ = A.adjoint(psi, coh, ustar, 2 * (u - target) * v)
    assert np.all(gphi[nm] == 0)
    h = 1e-6
    for _ in range(16):
        d = rng.normal(size=psi.shape)
        d[nm] = 0
        d /= np.linalg.norm(d)
        fd = (L(F(psi + h * d, coh)) - L(F(psi - h * d, coh))) / (2 * h)
        assert abs(fd - np.sum(gphi * d)) <= 1e-4 * max(1.0, abs(fd))
        dw = rng.normal(size=coh.shape)
        dw[nm] = 0
        dw /= np.linalg.norm(dw)
        fdw = (L(F(psi, coh + h * dw)) - L(F(psi, coh - h * dw))) / (2 * h)
        assert abs(fdw - np.sum(gw * dw)) <= 1e-4 * max(1.0, abs(fdw))


def test_adjoint_pcg_equals_direct():
    psi, coh, _ = S.generate(S.Workload(10, 12, 5, ramp_max=6.0, n_bumps=1, noise=1.0))
    psi = psi.astype(np.float64)
    coh = coh.astype(np.float64)
    _, _, nm, pf, wf = post.prologue(psi.astype(np.float32), coh.astype(np.float32))
    ustar = A.fixed_point(psi, coh, O.solve(pf, wf, ~nm, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0))
    g = np.random.default_rng(2).normal(size=psi.shape)
    a = A.adjoint(psi, coh, ustar, g)
    b = A.adjoint(psi, coh, ustar, g, solver='pcg', cg_iters=3000)
    np.testing.assert_allclose(b[0], a[0], atol=1e-7 * np.abs(a[0]).max())
    np.testing.assert_allclose(b[1], a[1], atol=1e-7 * np.abs(a[1]).max())


def test_phase_gradient_is_centred_incoming_gradient():
    """H lam = C^T g: the fixed point moves with psi (translation equivariance)."""
    psi, coh, _ = S.generate(S.Workload(11, 13, 8, ramp_max=6.0, n_bumps=1, noise=1.0))
    psi[4:6, 3:5] = np.nan
    psi = psi.astype(np.float64)
    coh = coh.astype(np.float64)
    _, _, nm, pf, wf = post.prologue(psi.astype(np.float32), coh.astype(np.float32))
    v = ~nm
    ustar = A.fixed_point(psi, coh, O.solve(pf, wf, v, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0))
    g = np.random.default_rng(3).normal(size=psi.shape)
    gphi, _, _ = A.adjoint(psi, coh, ustar, g)
    np.testing.assert_allclose(gphi, v * (g - g.sum() / v.sum()), atol=1e-9)
    d = np.zeros_like(psi)
    d[v] = 1e-3 * np.random.default_rng(4).normal(size=v.sum())
    u2 = A.fixed_point(psi + d, coh, ustar)
    np.testing.assert_allclose(A.center(u2, v), A.center(ustar + d, v), atol=1e-8)


@pytest.mark.parametrize("noise", [0.3, 1.0])
def test_chord_solver_equals_newton_fixed_point(noise):
    """The chord iterations used for the 256 x 256 finite differences land on the same perturbed
    fixed point as the damped-Newton solve (an independent method) to the rounding floor."""
    wl = S.Workload(32, 36, 5, ramp_max=10.0, n_bumps=2, noise=noise)
    psi, coh, _ = S.generate(wl)
    psi = psi.astype(np.float64)
    coh = coh.astype(np.float64)
    _, _, nm, pf, wf = post.prologue(psi.astype(np.float32), coh.astype(np.float32))
    u0 = O.solve(pf, wf, ~nm, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
    ustar = A.fixed_point(psi, coh, u0)
    F = A.chord_solver(psi, coh, ustar)
    rng = np.random.default_rng(7)
    d = rng.normal(size=psi.shape)
    for h in (1e-6, 1e-4):
        for ps, wt in ((psi + h * d, coh), (psi, coh + h * d)):
            a = A.center(F(ps, wt), ~nm)
            b = A.center(A.fixed_point(ps, wt, ustar), ~nm)
            assert np.abs(a - b).max() < 1e-11
    assert np.abs(A.center(F(psi, coh), ~nm) - A.center(ustar, ~nm)).max() < 1e-12

"""The implicit adjoint against autograd through the unrolled float64 torch restatement of the
reference solver (oracle/irls_unrolled.py), on small grids where 250 IRLS x 200 PCG iterations
converge: the gradient of the centred output through every unrolled iteration must equal the
implicit derivative of the fixed point (oracle/adjoint_fp64.py)."""
import numpy as np
import pytest
import torch

from core_amd import synthetic as S
from oracle import adjoint_fp64 as A, irls_unrolled as U, post

EPS = 0.5


def unrolled_grads(psi, coh, g, max_iter=250, cg_max_iter=200, epsilon=EPS):
    """(centred u, dL/dpsi, dL/dweight) for L = sum(g * u) by autograd through the unrolled solve;
    the weight enters through the reference's fill and clip(0.01, 1)."""
    nm = np.isnan(psi)
    ph = torch.tensor(np.where(nm, 0.0, psi.astype(np.float64)), requires_grad=True)
    wr = torch.tensor(np.where(nm | np.isnan(coh), 0.0, coh.astype(np.float64)), requires_grad=True)
    uc = U.solve(ph, torch.clamp(wr, 0.01, 1.0), torch.tensor(~nm), max_iter, cg_max_iter, epsilon)
    (uc * torch.tensor(g)).sum().backward()
    return uc.detach().numpy(), ph.grad.numpy(), wr.grad.numpy()


def case(noise, mask, seed=3):
    psi, coh, _ = S.generate(S.Workload(10, 12, seed, ramp_max=6.0, n_bumps=2, noise=noise))
    if mask:
        psi[2:4, 5:8] = np.nan
        coh[2:4, 5:8] = np.nan
    g = np.random.default_rng(4).normal(size=psi.shape)
    g[np.isnan(psi)] = 0.0   # the product returns NaN there: no loss term
    return psi, coh, g


@pytest.mark.parametrize("noise,mask", [(1.2, False), (1.0, True), (0.6, False)])
def test_implicit_adjoint_equals_unrolled_autograd(noise, mask):
    """Bars relative to the largest gradient entry: 2e-5 for dL/dpsi, 5e-4 for dL/dw.  The unrolled
    IRLS approaches its fixed point at rate 1 - eps^2/(r^2 + eps^2) on edges with |r| >> eps, so
    after 250 iterations it still sits ~1e-7 from stationarity (the Newton fixed point: 1e-16),
    which the weight gradient feels most.  (0.6: residue-free, so dL/dw is exactly 0.)"""
    psi, coh, g = case(noise, mask)
    v = ~np.isnan(psi)
    uc, ga_phi, ga_w = unrolled_grads(psi, coh, g)
    ustar = A.fixed_point(psi.astype(np.float64), coh.astype(np.float64), uc, epsilon=EPS)
    gphi, gw, _ = A.adjoint(psi.astype(np.float64), coh.astype(np.float64), ustar, g, epsilon=EPS)
    scale = max(np.abs(gw[v]).max(), np.abs(gphi[v]).max())
    assert np.abs(ga_phi[v] - gphi[v]).max() <= 2e-5 * np.abs(gphi[v]).max()
    assert np.abs(ga_w[v] - gw[v]).max() <= 5e-4 * scale
    assert np.abs(ga_phi[~v]).max(initial=0.0) <= 1e-12   # NaN pixels are not inputs


def test_unrolled_restatement_matches_fp32_oracle_forward():
    """The unrolled solver is the reference algorithm: with the same iterations its output equals
    the float32 restatement's (centred) to float32 accuracy."""
    from oracle import irls_torch32
    psi, coh, _ = case(0.8, True)
    nm = np.isnan(psi)
    _, _, _, pf, wf = post.prologue(psi.astype(np.float32), coh.astype(np.float32))
    u32 = irls_torch32.solve(pf, wf, ~nm, max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0, epsilon=1e-2)
    u32 = u32 - u32[~nm].mean()
    uc = U.solve(torch.tensor(pf.astype(np.float64)), torch.tensor(wf.astype(np.float64)), torch.tensor(~nm),
                 max_iter=10, cg_max_iter=50, epsilon=1e-2).numpy()
    assert np.abs(uc - u32)[~nm].max() < 1e-4

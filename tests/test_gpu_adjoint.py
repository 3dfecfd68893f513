"""GPU implicit adjoint (irls2d_adjoint / core_amd.autograd) against the FP64 oracle adjoint
(oracle/adjoint_fp64.py, pinned by finite differences in tests/test_adjoint_oracle.py), evaluated
at the same fixed point u*, the oracle with an exact sparse solve."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _case(H, W, seed, noise, mask=False):
    psi, coh, _ = S.generate(S.Workload(H, W, seed, ramp_max=10.0, n_bumps=2, noise=noise))
    if mask:
        psi[H // 3:H // 3 + 3, W // 2:W // 2 + 4] = np.nan
    return psi, coh


@pytest.mark.parametrize("H,W,seed,noise,mask", [(24, 28, 3, 0.4, False), (40, 50, 4, 1.1, True), (33, 61, 5, 0.8, False)])
def test_gpu_adjoint_matches_fp64(hip, H, W, seed, noise, mask):
    import torch
    from core_amd.autograd import irls_unwrap_2d_diff
    from oracle import adjoint_fp64 as A
    psi, coh = _case(H, W, seed, noise, mask)
    print(f"adjoint case {H}x{W} seed={seed}", flush=True)
    ph = torch.from_numpy(psi)[None].to(hip).requires_grad_(True)
    wt = torch.from_numpy(coh)[None].to(hip).requires_grad_(True)
    u = irls_unwrap_2d_diff(ph, wt, max_iter=10, cg_max_iter=50, adj_cg_iter=3000)
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(H, W)).astype(np.float32)
    g[np.isnan(psi)] = 0
    (u * torch.from_numpy(g)[None].to(hip)).sum().backward()
    gphi = ph.grad[0].cpu().numpy().astype(np.float64)
    gw = wt.grad[0].cpu().numpy().astype(np.float64)
    assert np.isfinite(gphi).all() and np.isfinite(gw).all()
    # oracle at the GPU's own fixed point
    from core_amd.autograd import IrlsUnwrap2d  # noqa: F401
    ustar = (u.detach()[0].cpu().numpy()).astype(np.float64)
    rphi, rw, _ = A.adjoint(psi.astype(np.float64), coh.astype(np.float64), ustar, g.astype(np.float64))
    scale = np.abs(rphi).max()
    ephi = np.abs(gphi - rphi).max() / scale
    ew = np.abs(gw - rw).max() / max(np.abs(rw).max(), 1e-2 * scale)
    print(f"  |dL/dphase| {scale:.2e} rel err {ephi:.2e};  |dL/dweight| {np.abs(rw).max():.2e} err {ew:.2e}")
    assert ephi < 1e-5
    assert ew < 2e-2


def test_autograd_batch_shapes_and_loss(hip):
    import torch
    from core_amd.autograd import irls_unwrap_2d_diff
    ims = [_case(32, 36, 10 + i, 0.5) for i in range(3)]
    ph = torch.from_numpy(np.stack([i[0] for i in ims])).to(hip).requires_grad_(True)
    wt = torch.from_numpy(np.stack([i[1] for i in ims])).to(hip).requires_grad_(True)
    u = irls_unwrap_2d_diff(ph, wt, adj_cg_iter=100)
    assert u.shape == (3, 32, 36)
    target = torch.zeros_like(u)
    loss = ((u - target) ** 2).sum()
    loss.backward()
    assert ph.grad.shape == ph.shape and wt.grad.shape == wt.shape
    assert torch.isfinite(ph.grad).all() and torch.isfinite(wt.grad).all()
    # centred output: masked mean zero per image
    assert float(u.detach().mean(dim=(1, 2)).abs().max()) < 1e-3


@pytest.mark.parametrize("noise,mask", [(1.2, False), (1.0, True)])
def test_gpu_adjoint_matches_unrolled_autograd(hip, noise, mask):
    """The GPU implicit adjoint against autograd through the unrolled float64 torch restatement of
    the reference solver (oracle/irls_unrolled.py), both run long enough to converge on 10 x 12."""
    import sys, os
    import torch
    from core_amd.autograd import irls_unwrap_2d_diff
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_adjoint_unrolled import EPS, case, unrolled_grads
    psi, coh, g = case(noise, mask)
    print(f"unrolled-autograd case noise={noise} mask={mask}", flush=True)
    v = ~np.isnan(psi)
    _, ga_phi, ga_w = unrolled_grads(psi, coh, g)
    ph = torch.from_numpy(psi)[None].to(hip).requires_grad_(True)
    wt = torch.from_numpy(coh)[None].to(hip).requires_grad_(True)
    u = irls_unwrap_2d_diff(ph, wt, max_iter=250, cg_max_iter=200, epsilon=EPS, adj_cg_iter=3000)
    (u * torch.from_numpy(g.astype(np.float32))[None].to(hip)).sum().backward()
    gphi = ph.grad[0].cpu().numpy().astype(np.float64)
    gw = wt.grad[0].cpu().numpy().astype(np.float64)
    scale = max(np.abs(ga_phi[v]).max(), np.abs(ga_w[v]).max())
    ephi = np.abs(gphi[v] - ga_phi[v]).max() / np.abs(ga_phi[v]).max()
    ew = np.abs(gw[v] - ga_w[v]).max() / scale
    print(f"  dL/dphase rel err {ephi:.2e}; dL/dweight err {ew:.2e} (of the largest gradient entry)", flush=True)
    assert ephi < 1e-4
    assert ew < 2e-3

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


@pytest.mark.timeout(900)
@pytest.mark.parametrize("noise", [0.3, 1.0])
def test_adjoint_256_against_fp64_finite_differences(hip, noise):
    """The differentiable-unwrap check at 256 x 256 (the 1024-case generator scaled: ramp 0-10 rad,
    6 bumps, coherence U[0.3, 1]; noise 0.3 as specified, and 1.0, which adds residues so that the
    weights matter): 64 random unit probe directions, FP64 central differences with step 1e-6 of
    the IRLS fixed-point map (oracle/adjoint_fp64.py: damped Newton to the fixed point, chord
    iterations for the perturbed inputs), loss sum((C u - target)^2) over valid pixels with the
    noiseless phase as the (centred) target.

    Each directional derivative is compared with its finite difference relative to
    max(|difference|, RMS directional derivative) -- the RMS over random unit directions is
    |gradient|_2 / sqrt(valid pixels); the differences carry an absolute rounding floor near 5e-7.
    FP64 oracle adjoint: within 1e-5.  GPU adjoint (float32, evaluated at the same fixed point):
    within 1e-4 for dL/dphase and 1e-3 for dL/dweight.  Without residues the fixed point has r = 0
    on every edge and dL/dweight vanishes; then both adjoints and the differences must be 0 to
    1e-5 of the phase scale."""
    import torch
    from core_amd.autograd import adjoint_device
    from core_amd.unwrap2d import Workspace, image_view, pack, solve_device
    from oracle import adjoint_fp64 as A
    wl0 = S.scaled(S.single_1024(), 256, 256)
    wl = S.Workload(256, 256, wl0.seed, wl0.ramp_max, wl0.n_bumps, coh_lo=wl0.coh_lo, coh_hi=wl0.coh_hi, noise=noise)
    psi, coh, true = S.generate(wl)
    print(f"adjoint FD 256x256 noise={noise}", flush=True)
    v = ~np.isnan(psi)
    ws = Workspace.get([psi.shape], hip)
    u0, _ = solve_device(pack([psi], ws, hip), pack([coh], ws, hip), [psi.shape], max_iter=10, tol=0.0,
                         cg_max_iter=50, cg_tol=0.0, workspace=ws)
    u0 = image_view(u0, ws, 0).cpu().numpy().astype(np.float64)
    p64, c64 = psi.astype(np.float64), coh.astype(np.float64)
    ustar = A.fixed_point(p64, c64, u0)
    F0 = A.chord_solver(p64, c64, ustar)
    F = lambda ps, wt: A.center(F0(ps, wt), v)
    target = A.center(true.astype(np.float64), v)
    u = A.center(ustar, v)
    g = 2 * (u - target) * v
    rphi, rw, _ = A.adjoint(p64, c64, ustar, g)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32))[None].to(hip)
    gph, gw = adjoint_device(t(psi), t(coh), t(ustar), t(g), cg_iters=3000)
    gph = gph[0].cpu().numpy().astype(np.float64)
    gw = gw[0].cpu().numpy().astype(np.float64)
    nv = np.sqrt(v.sum())
    rms_phi, rms_w = np.linalg.norm(rphi) / nv, np.linalg.norm(rw) / nv
    weights_matter = rms_w > 1e-3 * rms_phi
    print(f"  RMS directional derivative: phase {rms_phi:.2e} weight {rms_w:.2e}; GPU-vs-oracle max abs "
          f"phase {np.abs(gph - rphi).max():.1e} weight {np.abs(gw - rw).max():.1e}", flush=True)
    dL = lambda up, um: float(np.sum((up - um) * (up + um - 2 * target) * v))
    rng = np.random.default_rng(64)
    h = 1e-6
    worst = np.zeros(4)
    for probe in range(64):
        if probe % 16 == 0:
            print(f"  probe {probe}", flush=True)
        d = rng.normal(size=psi.shape)
        d[~v] = 0
        d /= np.linalg.norm(d)
        fd = dL(F(p64 + h * d, c64), F(p64 - h * d, c64)) / (2 * h)
        dw = rng.normal(size=psi.shape)
        dw[~v] = 0
        dw /= np.linalg.norm(dw)
        fdw = dL(F(p64, c64 + h * dw), F(p64, c64 - h * dw)) / (2 * h)
        den = max(abs(fd), rms_phi)
        denw = max(abs(fdw), rms_w) if weights_matter else rms_phi
        worst = np.maximum(worst, [abs(fd - np.sum(rphi * d)) / den, abs(fdw - np.sum(rw * dw)) / denw,
                                   abs(fd - np.sum(gph * d)) / den, abs(fdw - np.sum(gw * dw)) / denw])
    print(f"  64 probes, worst relative error vs FD: oracle phase {worst[0]:.1e} weight {worst[1]:.1e}; "
          f"GPU phase {worst[2]:.1e} weight {worst[3]:.1e} (weights matter: {weights_matter})", flush=True)
    assert worst[0] < 1e-5 and worst[1] < 1e-5
    assert worst[2] < 1e-4 and worst[3] < (1e-3 if weights_matter else 1e-5)
    if noise >= 1.0:
        assert weights_matter

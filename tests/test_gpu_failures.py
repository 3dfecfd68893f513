"""GPU tests of the reference's failure paths and early exits against the oracles.

* Non-finite x inside PCG (utils_unwrap2d.py:1444-1449) and the u_best restore that follows
  (:1551-1558, :1599-1604): no finite float32 input reaches them (the alpha / beta clamps and the
  rz / pAp breaks stop PCG first), so the condition is planted from the host with the library's
  test hook irls2d_debug_inject, mirrored by the oracles' `inject` argument.
* +inf / -inf phase pixels: NaN targets, the skipped PCG (non-finite r0, :1414-1419), and the
  epilogue's NaN circular mean on the affected components.
* Huge finite phases: cycle counts beyond the device histogram (the host median path).
* tol > 0: images of one batch stopping at different IRLS iterations.
"""
import warnings

import numpy as np
import pytest
import torch

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)


def _rms(u, ref, mask):
    d = u[mask].astype(np.float64) - ref[mask].astype(np.float64)
    return float(np.sqrt(np.mean(d * d))) if d.size else 0.0


def _kagree(a, b, psi, mask):
    ka = np.round((a[mask].astype(np.float64) - psi[mask]) / (2 * np.pi))
    kb = np.round((b[mask].astype(np.float64) - psi[mask]) / (2 * np.pi))
    return float(np.mean(ka == kb)) if ka.size else 1.0


def _compare(name, psi, u, cc, u64, cc64, u32, envelope=False):
    """Bars: RMS vs FP64 <= 1e-3 and k agreement with the float32 path >= 99.99%; with
    `envelope`, where the reference's own float32 path is further from FP64 than that (an early,
    unconverged iterate), RMS <= 2 x RMS(fp32 - fp64) and k agreement >= that of fp32 vs fp64."""
    fin = np.isfinite(u64)
    assert np.array_equal(np.isnan(u), np.isnan(u64)), name
    assert np.array_equal(np.isnan(u32), np.isnan(u64)), name
    rms = _rms(u, u64, fin)
    agree = _kagree(u, u32, psi, fin)
    rms_bar, agree_bar = 1e-3, 0.9999
    if envelope:
        rms_bar = max(rms_bar, 2 * _rms(u32, u64, fin))
        agree_bar = min(agree_bar, _kagree(u64, u32, psi, fin))
    print(f"  {name}: rms_vs_fp64={rms:.2e} (bar {rms_bar:.1e}) k_agree={agree:.6f} (bar {agree_bar:.6f}) "
          f"nan_px={int((~fin).sum())}", flush=True)
    assert rms <= rms_bar and agree >= agree_bar, name
    assert np.array_equal(cc, cc64), name


@pytest.mark.parametrize("inject", [(0, 0), (2, 7), (6, 49)])
def test_injected_nonfinite_x_restores_u_best(hip, inject):
    """x made non-finite in PCG iteration inject[1] of IRLS iteration inject[0]: PCG breaks, IRLS
    restores u_best (the centred iterate of the lowest residual so far) and stops.  The raw output
    equals the centred iterate of an un-injected solve stopped at that iteration; the drop-in
    output matches the oracles run with the same injection."""
    from core_amd import _lib, irls_unwrap_2d, solve_device
    from core_amd.unwrap2d import Workspace, image_view, pack
    from oracle import irls_fp64, irls_torch32, post
    psi, coh, _ = S.generate(S.Workload(96, 130, 31, ramp_max=40.0, fault=True, mask_frac=0.05))
    assert not np.isnan(psi[0]).all() and not np.isnan(psi[:, 0]).all()  # no crop: raw arenas align
    shapes = [psi.shape]
    ws = Workspace(shapes, hip)
    ph, co = pack([psi], ws, hip), pack([coh], ws, hip)
    L = _lib.load()
    print(f"inject {inject}", flush=True)
    _lib.check(L.irls2d_debug_inject(*inject), "irls2d_debug_inject")
    try:
        u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
        raw, info = solve_device(ph, co, shapes, workspace=ws, graph=True, **FIXED)  # graph ignored while armed
        raw = image_view(raw, ws, 0).cpu().numpy().copy()
        info = info.cpu().numpy()[0].copy()
    finally:
        _lib.check(L.irls2d_debug_inject(-1, -1), "irls2d_debug_inject")
    assert info[0] & 1 and info[1] == inject[0] + 1, info   # restored u_best, stopped at that iteration
    # u_best: the iterate with the lowest residual (FP64 trace), centred over the valid pixels
    tr = []
    _, _, nm, pf, wf = post.prologue(psi, coh)
    irls_fp64.solve(pf, wf, ~nm, trace=tr, inject=inject, **FIXED)
    best_it = int(np.argmin([t['residual'] for t in tr]))
    ub, _ = solve_device(ph, co, shapes, workspace=ws, **{**FIXED, 'max_iter': best_it})
    ub = image_view(ub, ws, 0).cpu().numpy()
    v = ~np.isnan(psi)
    expect = ub - np.float32(ub[v].astype(np.float64).sum() / v.sum())
    print(f"  best iteration {best_it}: max |raw - centred iterate| = {np.abs(raw - expect).max():.1e}", flush=True)
    np.testing.assert_allclose(raw, expect, rtol=0, atol=1e-6 * max(1.0, float(np.abs(expect).max())))
    ws.release()
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, inject=inject, **FIXED)
    u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, inject=inject, **FIXED)
    _compare(f"inject {inject}", psi, u, cc, u64, cc64, u32, envelope=True)
    # disarmed again: the plain solve is back
    u2, _ = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    assert not np.array_equal(u2, u, equal_nan=True)


def test_infinite_phase_pixels(hip):
    """+inf and -inf phases are valid pixels with NaN targets: the DCT start is all NaN and reset
    to 0, every PCG stops at r0, and the components holding an infinite pixel end NaN (their
    circular mean is NaN)."""
    from core_amd import irls_unwrap_2d
    from oracle import irls_fp64, irls_torch32
    psi, coh, _ = S.generate(S.Workload(40, 50, 32, ramp_max=20.0))
    psi[:, 24] = np.nan            # two components; the left one holds both infinities
    psi[5, 7] = np.inf
    psi[30, 10] = -np.inf
    print("inf phases 40x50", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
        u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
        u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    assert np.isnan(u64[:, :24]).all() and np.isfinite(u64[:, 25:]).all()
    _compare("inf phases", psi, u, cc, u64, cc64, u32)


def test_huge_finite_phases_use_the_host_median(hip):
    """Finite phases of ~1e10 rad put k = round((u - psi) / 2 pi) beyond the device histogram (and
    beyond int32): the median then comes from the float32 k values, as np.median takes it."""
    from core_amd import irls_unwrap_2d
    from oracle import irls_fp64, irls_torch32
    base, coh, _ = S.generate(S.Workload(40, 50, 33, ramp_max=20.0))
    psi = base.copy()
    for (i, j), off in zip([(3, 3), (20, 41), (35, 12), (11, 27)], [-3e12, 1e10, 5e9, -2e10]):
        psi[i, j] = np.float32(psi[i, j] + off)   # isolated pixels: k beyond int32
    print("huge finite phases 40x50", flush=True)
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    assert np.isfinite(u).all()
    # the reference forms the targets next to the huge pixels in float32 (multiples of 1024 rad
    # apart before wrapping), the FP64 oracle in float64: the float32 envelope is the bar
    _compare("huge phases", psi, u, cc, u64, cc64, u32, envelope=True)
    # more than half of the grid huge: the median shift itself is ~1.6e9 cycles
    psi2 = base.copy()
    psi2[:, 15:] = (psi2[:, 15:].astype(np.float64) + 1e10).astype(np.float32)
    u, cc = irls_unwrap_2d(psi2, coh, device='cuda', **FIXED)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi2, coh, **FIXED)
    u32, _ = irls_torch32.irls_unwrap_2d(psi2, coh, **FIXED)
    _compare("huge phases, majority", psi2, u, cc, u64, cc64, u32)


def test_batch_with_different_stopping_iterations(hip):
    """Reference defaults (tol 1e-3, cg_tol 1e-4) on a batch whose images stop after 2, 18 and 4
    IRLS iterations: each image matches its FP64 oracle and its own plain call bitwise."""
    from core_amd import irls_unwrap_2d, irls_unwrap_2d_batch, solve_device
    from core_amd.unwrap2d import Workspace, pack
    from oracle import irls_fp64, irls_torch32
    wls = [S.Workload(48, 64, 90, ramp_max=10.0, n_bumps=0, noise=0.0),
           S.Workload(70, 90, 91, ramp_max=30.0, fault=True, mask_frac=0.05),
           S.Workload(64, 80, 92, ramp_max=20.0, noise=0.6)]
    ims = [S.generate(wl)[:2] for wl in wls]
    print("batch with tol", flush=True)
    res = irls_unwrap_2d_batch([a for a, _ in ims], [b for _, b in ims], device='cuda')
    shapes = [a.shape for a, _ in ims]
    ws = Workspace(shapes, hip)
    _, info = solve_device(pack([a for a, _ in ims], ws, hip), pack([b for _, b in ims], ws, hip), shapes,
                           workspace=ws, max_iter=50, tol=1e-3, cg_max_iter=20, cg_tol=1e-4)
    iters = info.cpu().numpy()[:, 1].tolist()
    ws.release()
    print(f"  IRLS iterations per image: {iters}", flush=True)
    assert len(set(iters)) == 3 and all(0 < it < 50 for it in iters)
    for (psi, coh), (u, cc) in zip(ims, res):
        u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh)
        u32, _ = irls_torch32.irls_unwrap_2d(psi, coh)
        _compare(f"{psi.shape}", psi, u, cc, u64, cc64, u32)
        up, cp = irls_unwrap_2d(psi, coh, device='cuda')
        assert np.array_equal(u, up, equal_nan=True) and np.array_equal(cc, cp)

"""GPU parity of the gfx950 solver against the oracles (run on an MI355X: pytest -m gpu).

Bars (per case, valid pixels only):
  * RMS(phi_gpu - phi_fp64) <= 1e-3 rad after the identical host epilogue,
  * cycle count k = round((phi - psi) / 2 pi) agrees with the float32 reference-CPU path on
    >= 99.99% of pixels,
  * conncomp labels equal the oracle's.
"""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)
RMS_TOL = 1e-3
K_AGREE = 0.9999


def _case(H, W, seed, fault=False, mask=0.0, ramp=40.0):
    wl = S.Workload(H, W, seed, ramp_max=ramp, n_bumps=6, fault=fault, mask_frac=mask)
    psi, coh, _ = S.generate(wl)
    return psi, coh


def _metrics(u, ref, psi):
    v = ~np.isnan(psi)
    d = u[v].astype(np.float64) - ref[v].astype(np.float64)
    return float(np.sqrt(np.mean(d * d))) if d.size else 0.0


def _kagree(u, ref, psi):
    v = ~np.isnan(psi)
    k1 = np.round((u[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    k2 = np.round((ref[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    return float(np.mean(k1 == k2)) if k1.size else 1.0


CASES = [
    (64, 48, 0, False, 0.0),      # 5-smooth lengths, direct FFT
    (96, 130, 1, True, 0.05),     # 130 = 2*5*13 -> Bluestein rows; fault + masked blocks
    (127, 211, 2, False, 0.0),    # prime sizes -> Bluestein both ways
    (256, 256, 3, True, 0.05),
    (200, 333, 4, False, 0.0),
]


@pytest.mark.parametrize("H,W,seed,fault,mask", CASES)
def test_parity_vs_oracles(hip, H, W, seed, fault, mask):
    from core_amd import irls_unwrap_2d
    from oracle import irls_fp64, irls_torch32
    psi, coh = _case(H, W, seed, fault, mask)
    print(f"case {H}x{W} seed={seed}", flush=True)
    u_gpu, cc_gpu = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    u32, cc32 = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    assert u_gpu.dtype == np.float32 and cc_gpu.dtype == np.uint16
    assert np.array_equal(np.isnan(u_gpu), np.isnan(psi))
    rms = _metrics(u_gpu, u64, psi)
    ka = _kagree(u_gpu, u32, psi)
    print(f"  rms_vs_fp64={rms:.3e} k_agree_vs_fp32={ka:.6f} rms_fp32_vs_fp64={_metrics(u32, u64, psi):.3e}")
    assert rms <= RMS_TOL
    assert ka >= K_AGREE
    assert np.array_equal(cc_gpu, cc64)


def test_batched_equals_plain_bitwise(hip):
    from core_amd import irls_unwrap_2d, irls_unwrap_2d_batch
    cases = [_case(70, 90, 10), _case(33, 129, 11, True, 0.05), _case(128, 64, 12)]
    batch = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', **FIXED)
    for (psi, coh), (ub, cb) in zip(cases, batch):
        up, cp = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
        assert np.array_equal(ub, up, equal_nan=True)
        assert np.array_equal(cb, cp)


def test_result_independent_of_batch_mates(hip):
    """Every strategy choice is per image (column pass, slab split, elements per thread, long
    lines), so an image's solve is bitwise the same alone and in a batch mixing all of them: two
    slab splits, direct (Bluestein) and transposed columns, rows above 8192 points (12 values per
    thread) and below, and a long column."""
    import torch
    from core_amd import _lib, solve_device
    from core_amd.unwrap2d import Workspace, image_view, pack
    shapes = [(1024, 600), (2048, 130), (70, 90), (200, 333), (40, 10000), (13122, 15)]
    print("strategies", _lib.describe_columns(shapes), flush=True)
    ims = [_case(h, w, 100 + i, i % 2 == 0, 0.05 if i == 1 else 0.0) for i, (h, w) in enumerate(shapes)]

    def solve(idx):
        sh = [shapes[k] for k in idx]
        ws = Workspace(sh, hip)
        u, info = solve_device(pack([ims[k][0] for k in idx], ws, hip), pack([ims[k][1] for k in idx], ws, hip), sh,
                               workspace=ws, **FIXED)
        out = [image_view(u, ws, j).cpu().numpy().copy() for j in range(len(idx))]
        ws.release()
        torch.cuda.empty_cache()
        return out

    together = solve(list(range(len(shapes))))
    for k in range(len(shapes)):
        alone = solve([k])[0]
        assert np.array_equal(together[k], alone, equal_nan=True), shapes[k]
        assert np.isfinite(alone[~np.isnan(ims[k][0])]).all()
    reordered = solve([5, 2, 0, 4, 1, 3])
    for j, k in enumerate([5, 2, 0, 4, 1, 3]):
        assert np.array_equal(reordered[j], together[k], equal_nan=True), shapes[k]


def test_graph_replay_equals_eager(hip):
    import torch
    from core_amd import solve_device
    from core_amd.unwrap2d import Workspace, image_view, pack
    shapes = [(60, 75), (41, 128)]
    ims = [_case(h, w, 20 + i) for i, (h, w) in enumerate(shapes)]
    ws = Workspace.get(shapes, hip)
    ph = pack([im[0] for im in ims], ws, hip)
    co = pack([im[1] for im in ims], ws, hip)

    def images(u):  # the arena's row padding is never written by the solver
        return torch.cat([image_view(u, ws, k).reshape(-1) for k in range(len(shapes))])

    u_e, info_e = solve_device(ph, co, shapes, graph=False, workspace=ws, **FIXED)
    u_e = images(u_e).clone()
    info_e = info_e.clone()
    # eager twice: the solver is deterministic
    u_e2, _ = solve_device(ph, co, shapes, graph=False, workspace=ws, **FIXED)
    assert torch.equal(images(u_e2), u_e)
    out = torch.full((ws.total,), float('nan'), device=hip)
    def diff(a, b):
        d = (a - b).abs()
        msg = f"{int((d != 0).sum())} differing of {d.numel()}, max {float(d.max()):.3e}"
        for k in range(len(shapes)):
            v = image_view(out, ws, k)
            nan = torch.isnan(v).nonzero().cpu().numpy()
            if len(nan):
                msg += f" | img{k} nan rows {sorted(set(nan[:, 0].tolist()))[:20]} cols {sorted(set(nan[:, 1].tolist()))[:40]}"
        return msg

    for rep in range(4):
        u_g, info_g = solve_device(ph, co, shapes, graph=True, workspace=ws, out=out, **FIXED)
        torch.cuda.synchronize()
        assert torch.equal(images(u_g), u_e), f"replay {rep}: {diff(images(u_g), u_e)} {info_g.tolist()} {info_e.tolist()}"
        assert torch.equal(info_g, info_e)
        # eager work between replays
        _ = (ph * 2).sum().item()
        out.fill_(-1.0)
    # a second graph on another output buffer
    out2 = torch.empty(ws.total, device=hip)
    u_g2, _ = solve_device(ph, co, shapes, graph=True, workspace=ws, out=out2, **FIXED)
    torch.cuda.synchronize()
    assert torch.equal(images(u_g2), u_e)
    ws.release()


def test_default_params_early_exit(hip):
    """Reference defaults (tol=1e-3, cg_tol=1e-4) exercise the device-side early exits."""
    from core_amd import irls_unwrap_2d
    from oracle import irls_fp64
    psi, coh = _case(96, 80, 5)
    u_gpu, cc_gpu = irls_unwrap_2d(psi, coh, device='cuda')
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh)
    assert _metrics(u_gpu, u64, psi) <= RMS_TOL
    assert np.array_equal(cc_gpu, cc64)


def test_edge_cases(hip):
    from core_amd import irls_unwrap_2d
    from oracle import irls_torch32
    # all NaN
    a = np.full((8, 9), np.nan, np.float32)
    u, cc = irls_unwrap_2d(a, device='cuda')
    assert np.isnan(u).all() and (cc == 0).all() and cc.dtype == np.uint16
    # non-2D
    with pytest.raises(ValueError):
        irls_unwrap_2d(np.zeros((2, 3, 4), np.float32), device='cuda')
    with pytest.raises(ValueError):
        irls_unwrap_2d(np.zeros((4, 4), np.float32), device='cpu')
    # tiny and degenerate shapes, no weight, crop
    rng = np.random.default_rng(7)
    for shape in [(1, 1), (1, 17), (19, 1), (2, 2), (5, 3)]:
        psi = rng.uniform(-np.pi, np.pi, size=shape).astype(np.float32)
        u, cc = irls_unwrap_2d(psi, device='cuda', **FIXED)
        u32, cc32 = irls_torch32.irls_unwrap_2d(psi, **FIXED)
        assert np.allclose(u, u32, atol=1e-4, equal_nan=True), shape
        assert np.array_equal(cc, cc32)
    # constant phase -> b = 0 -> pAp = 0 break
    psi = np.full((16, 12), 0.5, np.float32)
    u, cc = irls_unwrap_2d(psi, device='cuda', **FIXED)
    assert np.allclose(u, 0.5, atol=1e-5)
    # NaN border -> crop/uncrop
    psi, coh = _case(40, 50, 6)
    psi[:5] = np.nan
    psi[:, -7:] = np.nan
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u32, cc32 = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    assert np.array_equal(np.isnan(u), np.isnan(psi))
    assert _metrics(u, u32, psi) < 1e-3
    assert np.array_equal(cc, cc32)


@pytest.mark.parametrize("H,W,seed", [(96, 130, 1), (64, 48, 0)])
def test_column_modes_agree(hip, H, W, seed):
    """Transposed, direct and tridiagonal column passes give the same solve (to float32 rounding of
    the reordered arithmetic) and all meet the FP64 bar; batched = plain holds in each mode."""
    from core_amd import _lib, irls_unwrap_2d, irls_unwrap_2d_batch
    from oracle import irls_fp64
    psi, coh = _case(H, W, seed, True, 0.05)
    u64, _ = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    outs = {}
    try:
        for mode in (_lib.COLUMN_TRANSPOSE, _lib.COLUMN_DIRECT, _lib.COLUMN_TRIDIAG):
            _lib.set_column_mode(mode)
            outs[mode] = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
            assert _metrics(outs[mode][0], u64, psi) <= RMS_TOL
            cases = [(psi, coh), _case(40, 70, seed + 5)]
            batch = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', **FIXED)
            for (p_, c_), (ub, cb) in zip(cases, batch):
                up, cp = irls_unwrap_2d(p_, c_, device='cuda', **FIXED)
                assert np.array_equal(ub, up, equal_nan=True) and np.array_equal(cb, cp)
        v = ~np.isnan(psi)
        ref = outs[_lib.COLUMN_TRANSPOSE]
        # direct: the same arithmetic reordered; tridiagonal: the same operator with other rounding
        # (float64 solves in place of the float32 column DCTs), amplified along the IRLS trajectory
        # of these fault-line grids (both within the FP64 bar above)
        for mode, max_bar, rms_bar in ((_lib.COLUMN_DIRECT, 1e-4, 1e-5), (_lib.COLUMN_TRIDIAG, 1e-2, 5e-4)):
            d = np.abs(ref[0][v] - outs[mode][0][v])
            print(f"  mode {mode} vs transpose: max {d.max():.1e} RMS {_metrics(ref[0], outs[mode][0], psi):.1e}",
                  flush=True)
            assert d.max() < max_bar and _metrics(ref[0], outs[mode][0], psi) < rms_bar
            assert np.array_equal(ref[1], outs[mode][1])
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)


def test_batch_split_by_memory_equals_one_group(hip):
    """A stack split into several device solves (tiny memory budget) gives the same results."""
    from core_amd import irls_unwrap_2d_batch
    cases = [_case(50, 60, 70 + i, i == 1, 0.05 if i == 2 else 0.0) for i in range(4)]
    a = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', **FIXED)
    b = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', memory_fraction=1e-9,
                             **FIXED)
    for (ua, ca), (ub, cb) in zip(a, b):
        assert np.array_equal(ua, ub, equal_nan=True) and np.array_equal(ca, cb)

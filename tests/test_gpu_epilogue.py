"""GPU tests of the device epilogue (core_amd/epilogue_gpu.py, irls2d_ep_*) against the host
epilogue (core_amd/postprocess.py, itself checked against the oracle's restatement of the
reference): conncomp labels identical, unwrapped phase within float32 sin/cos rounding."""
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _arena(arrs, shapes, dev):
    from core_amd.unwrap2d import Workspace, pack
    ws = Workspace(shapes, dev)
    return ws, pack(arrs, ws, dev)


@pytest.mark.parametrize("H,W,seed,nan_frac,blocks", [(37, 53, 0, 0.35, False), (300, 411, 1, 0.0, True),
                                                      (128, 96, 2, 0.02, True), (1, 200, 3, 0.1, False),
                                                      (250, 1, 4, 0.1, False)])
def test_device_epilogue_matches_host(hip, H, W, seed, nan_frac, blocks):
    from core_amd.epilogue_gpu import epilogue_device
    from core_amd.postprocess import epilogue
    r = np.random.default_rng(seed)
    psi = r.uniform(-np.pi, np.pi, size=(H, W)).astype(np.float32)
    psi[r.uniform(size=(H, W)) < nan_frac] = np.nan
    if blocks:
        for _ in range(12):
            y, x = r.integers(0, H), r.integers(0, W)
            psi[y:y + r.integers(1, 30), x:x + r.integers(1, 30)] = np.nan
        psi[:, W // 2] = np.nan                       # splits the grid into components
    nm = np.isnan(psi)
    u = (np.where(nm, 0, psi) + 2 * np.pi * r.integers(-3, 4, size=(H, W)) * (r.uniform(size=(H, W)) < 0.1)
         + 2 * np.pi * 2 + r.normal(0, 0.4, size=(H, W))).astype(np.float32)
    print(f"epilogue case {H}x{W} seed={seed}", flush=True)
    for size in (1, 4, 30):
        ws, ph_t = _arena([psi], [(H, W)], hip)
        _, u_t = _arena([u], [(H, W)], hip)
        a = epilogue_device(ph_t, u_t, ws.offsets[0], ws.pitches[0], H, W, size)
        b = epilogue(u.copy(), psi, nm, size)
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_allclose(a[0], b[0], atol=2e-5, rtol=0, equal_nan=True)


def test_device_epilogue_nonfinite_warns(hip):
    from core_amd.epilogue_gpu import epilogue_device
    psi = np.zeros((6, 7), np.float32)
    u = np.zeros((6, 7), np.float32)
    u[2, 3] = np.inf
    ws, ph_t = _arena([psi], [(6, 7)], hip)
    _, u_t = _arena([u], [(6, 7)], hip)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert epilogue_device(ph_t, u_t, ws.offsets[0], ws.pitches[0], 6, 7, 30) is None
    assert any(issubclass(x.category, RuntimeWarning) for x in w)


def test_device_epilogue_large_grid_timing(hip):
    """8192^2 with masked blocks: labels identical to the host epilogue; prints both timings."""
    import time
    import torch
    from core_amd import synthetic as S
    from core_amd.epilogue_gpu import epilogue_device
    from core_amd.postprocess import epilogue
    psi, _, true = S.generate(S.single_8192())
    nm = np.isnan(psi)
    u = np.where(nm, 0, true + 2 * np.pi * 3).astype(np.float32)
    ws, ph_t = _arena([psi], [psi.shape], hip)
    _, u_t = _arena([u], [psi.shape], hip)
    torch.cuda.synchronize()
    t0 = time.time()
    a = epilogue_device(ph_t, u_t, ws.offsets[0], ws.pitches[0], *psi.shape, 30)
    t1 = time.time()
    b = epilogue(u.copy(), psi, nm, 30)
    t2 = time.time()
    print(f"  8192^2 epilogue: device {t1 - t0:.2f} s (incl. D2H), host {t2 - t1:.2f} s", flush=True)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[0], b[0], atol=1e-4, rtol=0, equal_nan=True)


def test_batch_api_device_and_host_epilogue_agree(hip):
    from core_amd import irls_unwrap_2d_batch
    from core_amd import synthetic as S
    ims = [S.generate(S.Workload(70, 90, 80 + i, ramp_max=30.0, n_bumps=3, fault=i == 0, mask_frac=0.1))[:2]
           for i in range(3)]
    kw = dict(device='cuda', max_iter=5, tol=0.0, cg_max_iter=30, cg_tol=0.0)
    a = irls_unwrap_2d_batch([i[0] for i in ims], [i[1] for i in ims], epilogue='device', **kw)
    b = irls_unwrap_2d_batch([i[0] for i in ims], [i[1] for i in ims], epilogue='host', **kw)
    for (ua, ca), (ub, cb) in zip(a, b):
        np.testing.assert_array_equal(ca, cb)
        np.testing.assert_allclose(ua, ub, atol=2e-5, rtol=0, equal_nan=True)

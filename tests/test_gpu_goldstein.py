"""GPU tests of the Goldstein filter (irls2d_goldstein): equal NaN masks and values within float32
rounding of the reference's own two filters (fixtures) and of the FP64 oracle, for every patch
size pair, odd image sizes, NaN blocks, low and default thresholds, both normalisations; the
reference's argument handling; a timing line at 8192 x 8192."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = sorted(glob.glob(os.path.join(HERE, "goldstein_*.npz")))
TOL = 2e-5  # max |gpu - ref| relative to max |ref|


def _cmp(out, ref):
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = ~np.isnan(ref)
    err = float(np.abs(out[v] - ref[v]).max() / np.abs(ref[v]).max())
    return err


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(f) for f in FIX])
@pytest.mark.parametrize("mode,key", [("count", "out_torch"), ("weight", "out_numpy")])
def test_goldstein_matches_reference_fixtures(hip, path, mode, key):
    from core_amd.goldstein import goldstein
    d = np.load(path)
    py, px = (int(v) for v in d["psize"])
    print(f"goldstein fixture {os.path.basename(path)} {mode}", flush=True)
    out = goldstein(d["phase"], d["corr"], psize={'y': py, 'x': px}, threshold=float(d["threshold"]),
                    device='cuda', mode=mode)
    assert out.dtype == np.complex64
    err = _cmp(out, d[key])
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < TOL


def _field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx]
    psi = 0.17 * x - 0.09 * y + 0.001 * x * y + rng.normal(0, 0.8, (ny, nx))
    z = ((0.3 + rng.uniform(0, 1, (ny, nx))) * np.exp(1j * psi)).astype(np.complex64)
    c = rng.uniform(0, 1, (ny, nx)).astype(np.float32)
    z[ny // 3:ny // 3 + 9, nx // 4:nx // 4 + 30] = np.nan
    c[ny // 2:ny // 2 + 4, :11] = np.nan
    return z, c


@pytest.mark.parametrize("ny,nx,py,px,thr,mode", [
    (257, 333, 32, 32, 0.5, "count"), (257, 333, 32, 32, 0.5, "weight"),
    (200, 130, 16, 16, 0.2, "count"), (200, 130, 16, 16, 0.2, "weight"),
    (190, 260, 64, 64, 0.5, "count"), (130, 97, 16, 32, 0.5, "weight"),
    (160, 150, 32, 64, 0.3, "count"), (150, 70, 64, 16, 0.5, "count"),
])
def test_goldstein_matches_fp64_oracle(hip, ny, nx, py, px, thr, mode):
    from core_amd.goldstein import goldstein
    from oracle import goldstein as G
    z, c = _field(ny, nx, ny + nx + py)
    print(f"goldstein {ny}x{nx} psize {py}x{px} thr {thr} {mode}", flush=True)
    out = goldstein(z, c, psize={'y': py, 'x': px}, threshold=thr, device='cuda', mode=mode)
    ref = G.filter(z, c, py, px, thr, mode, np.float64)
    err = _cmp(out, ref)
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < TOL


def test_goldstein_argument_handling(hip):
    from core_amd import _lib
    from core_amd.goldstein import goldstein
    z, c = _field(64, 80, 5)
    a = goldstein(z, c, psize=16, device='cuda')
    b = goldstein(z[None], c[None], psize={'y': 16, 'x': 16}, device='cuda')
    assert b.shape == (1, 64, 80) and np.array_equal(a, b[0], equal_nan=True)
    # complex128 / float64 inputs are cast like the reference
    d = goldstein(z.astype(np.complex128), c.astype(np.float64), psize=16, device='cuda')
    assert d.dtype == np.complex64 and np.array_equal(a, d, equal_nan=True)
    assert goldstein(z, c, psize=None) is z
    with pytest.raises(_lib.SolverError):     # IRLS2D_EUNSUPPORTED: patch sizes 16, 32, 64
        goldstein(z, c, psize=24, device='cuda')
    with pytest.raises(ValueError):           # IRLS2D_EINVAL: smaller than half a patch
        goldstein(z[:6], c[:6], psize=16, device='cuda')
    with pytest.raises(ValueError):
        goldstein(z, c, psize=16, device='cpu')


def test_goldstein_timing_8192(hip):
    import torch
    from core_amd.goldstein import goldstein_device
    n = 8192
    g = torch.Generator(device=hip).manual_seed(0)
    ph = torch.polar(torch.ones(n, n, device=hip), torch.rand(n, n, device=hip, generator=g) * 6.28)
    cr = torch.rand(n, n, device=hip, generator=g)
    out = goldstein_device(ph, cr, psize=32)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        goldstein_device(ph, cr, psize=32, out=out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(f"  goldstein 8192x8192 psize 32: {ms:.2f} ms = {n * n / ms / 1e3:.0f} Mpx/s", flush=True)
    assert torch.isfinite(out).all()

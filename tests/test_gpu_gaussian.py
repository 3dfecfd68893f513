"""GPU tests of the NaN-aware Gaussian filter (irls2d_gaussian): equal NaN masks and values within
float32 rounding of the FP64 oracle (oracle/gaussian.py) for real and complex data, with and
without weights, isotropic and anisotropic sigmas, and gaussian_numpy's decimation path; the
reference's argument handling; a timing line at 8192 x 8192."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _cmp(out, ref):
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = ~np.isnan(ref)
    return float(np.abs(out[v] - ref[v]).max() / np.abs(ref[v]).max())


def _data(ny, nx, cplx, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx]
    a = np.sin(0.05 * x) + np.cos(0.03 * y) + rng.normal(0, 0.3, (ny, nx))
    if cplx:
        a = np.exp(1j * a) * (0.5 + rng.uniform(0, 1, (ny, nx)))
        a = a.astype(np.complex64)
        a.real[ny // 2, : nx // 3] = np.nan          # one part only
    else:
        a = a.astype(np.float32)
    a[ny // 4:ny // 4 + 7, nx // 5:nx // 5 + 20] = np.nan
    w = rng.uniform(0, 1, (ny, nx)).astype(np.float32)
    w[3:6, :] = np.nan
    return a, w


@pytest.mark.parametrize("ny,nx,cplx,weighted,sigma,thr", [
    (150, 170, False, False, 2.0, 0.5), (150, 170, True, False, 2.0, 0.5),
    (131, 97, False, True, (1.5, 3.0), 0.5), (131, 97, True, True, 0.7, 0.5),
    (200, 90, False, False, 10.0, 0.3), (64, 300, True, False, (4.0, 0.5), 0.8),
    (600, 530, True, True, 2.0, 0.5), (517, 261, False, False, (2.0, 16.0), 0.5),
])
def test_nanconv_matches_fp64_oracle(hip, ny, nx, cplx, weighted, sigma, thr):
    from core_amd.gaussian import nanconvolve2d_gaussian
    from oracle import gaussian as G
    a, w = _data(ny, nx, cplx, ny + nx)
    w = w if weighted else None
    print(f"gaussian {ny}x{nx} complex={cplx} weighted={weighted} sigma={sigma} thr={thr}", flush=True)
    out = nanconvolve2d_gaussian(a, w, sigma, 4.0, thr, device='cuda')
    assert out.dtype == (np.complex64 if cplx else np.float32)
    ref = G.nanconv(a.astype(np.complex128 if cplx else np.float64), w, sigma, 4.0, thr)
    if cplx:
        err = max(_cmp(out.real, ref.real), _cmp(out.imag, ref.imag))
    else:
        err = _cmp(out, ref)
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < TOL


def test_decimation_path_matches_oracle(hip):
    from core_amd.gaussian import gaussian
    from oracle import gaussian as G
    a, _ = _data(301, 283, False, 7)
    print("gaussian decimation 301x283 sigma 200 px at 15 m", flush=True)
    out = gaussian(a, None, 200.0, 4.0, 0.5, device='cuda', pixel_sizes=(15.0, 15.0))
    ref = G.gaussian(a.astype(np.float64), None, 200.0, 4.0, 0.5, pixel_sizes=(15.0, 15.0))
    err = _cmp(out, ref)
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < TOL


def test_gaussian_argument_handling(hip):
    from core_amd.gaussian import gaussian
    a, w = _data(40, 50, True, 3)
    b = gaussian(a, w, 2.0, device='cuda')
    c = gaussian(a[None], w, (2.0, 2.0), device='cuda')
    assert c.shape == (1, 40, 50) and np.array_equal(b, c[0], equal_nan=True)
    assert gaussian(a, w, None) is a
    with pytest.raises(ValueError):
        gaussian(a, w, 2.0, device='cpu')
    with pytest.raises(ValueError):
        gaussian(a, w, -1.0, device='cuda')


def test_gaussian_timing_8192(hip):
    import torch
    from core_amd.gaussian import gaussian_device
    n = 8192
    g = torch.Generator(device=hip).manual_seed(0)
    z = torch.polar(torch.ones(n, n, device=hip), torch.rand(n, n, device=hip, generator=g) * 6.28)
    out = gaussian_device(z, None, 2.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        gaussian_device(z, None, 2.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(f"  gaussian 8192x8192 complex sigma 2: {ms:.2f} ms = {n * n / ms / 1e3:.0f} Mpx/s", flush=True)
    assert torch.isfinite(out).all()

"""The Gaussian-filter oracle (oracle/gaussian.py) pinned: without NaN it is scipy.ndimage's
gaussian_filter with mode='nearest' (the same replicate-border normalised convolution when
sigma * truncate is an integer); with NaN and weights it equals a brute-force normalised sum over
the clamped window; the threshold and the weight limit of the reference's two cases; the
decimation parameters of gaussian_numpy's GPU branch.  Planted mistakes fail these checks."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import gaussian as G


@pytest.mark.parametrize("sy,sx", [(2.0, 2.0), (1.5, 3.0), (0.5, 1.0)])
def test_matches_scipy_without_nan(sy, sx):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(37, 53))
    out = G.nanconv(a, None, (sy, sx), 4.0, 0.5)
    # without NaN the denominator is the kernel sum, 1: out = scipy / (1 + 1e-10)
    ref = ndi.gaussian_filter(a, (sy, sx), mode="nearest", truncate=4.0) / (1.0 + 1e-10)
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def _brute(a, w, ky, kx, lim):
    ny, nx = a.shape
    ry, rx = len(ky) // 2, len(kx) // 2
    out = np.full(a.shape, np.nan)
    for y in range(ny):
        for x in range(nx):
            num = den = 0.0
            for i in range(-ry, ry + 1):
                for j in range(-rx, rx + 1):
                    yy, xx = min(max(y + i, 0), ny - 1), min(max(x + j, 0), nx - 1)
                    v = a[yy, xx]
                    ww = 0.0 if np.isnan(v) else (1.0 if w is None else (0.0 if np.isnan(w[yy, xx]) else w[yy, xx]))
                    k = ky[i + ry] * kx[j + rx]
                    num += k * (0.0 if np.isnan(v) else v) * ww
                    den += k * ww
            if den > lim:
                out[y, x] = num / (den + 1e-10)
    return out


@pytest.mark.parametrize("weighted", [False, True])
def test_nan_and_weights_against_brute_force(weighted):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(14, 17))
    a[2:6, 3:9] = np.nan
    a[10, :] = np.nan
    w = rng.uniform(0, 1, a.shape) if weighted else None
    if weighted:
        w[0, :4] = np.nan
    ky, kx = G.kernel_1d(1.0, 4.0), G.kernel_1d(1.5, 2.0)
    lim = 1e-10 if weighted else 0.5
    ref = _brute(a, w, ky, kx, lim)
    out = G._nanconv_real(a, w, ky, kx, 0.5)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = ~np.isnan(ref)
    assert np.allclose(out[v], ref[v], rtol=1e-12, atol=1e-12)


def test_complex_parts_filtered_separately():
    rng = np.random.default_rng(2)
    z = rng.normal(size=(20, 22)) + 1j * rng.normal(size=(20, 22))
    z.real[3:5, 3:5] = np.nan  # only the real part invalid there
    out = G.nanconv(z, None, 1.5)
    assert np.allclose(out.imag, G.nanconv(z.imag, None, 1.5), equal_nan=True)
    assert np.allclose(out.real, G.nanconv(z.real, None, 1.5), equal_nan=True)


def test_kernel_and_planted_mistakes():
    k = G.kernel_1d(2.0, 4.0)
    assert len(k) == 17 and np.isclose(k.sum(), 1.0) and np.allclose(k, k[::-1])
    assert np.isclose(k[9] / k[8], np.exp(-0.5 / 4.0))
    rng = np.random.default_rng(3)
    a = rng.normal(size=(30, 30))
    good = G.nanconv(a, None, 2.0)
    ref = ndi.gaussian_filter(a, 2.0, mode="nearest", truncate=4.0)
    wrap = ndi.gaussian_filter(a, 2.0, mode="wrap", truncate=4.0)       # wrong border
    wide = ndi.gaussian_filter(a, 2.2, mode="nearest", truncate=4.0)    # wrong sigma
    assert np.allclose(good, ref) and not np.allclose(good, wrap) and not np.allclose(good, wide)


def test_decimation_parameters():
    # 15 m pixels, resolution 67: scales 4; sigma 200 px -> wavelength 15.9 km >> 24 x 67 m
    use, scy, scx, s_aa, s_dec = G.decimation(200.0, (15.0, 15.0))
    assert use and scy == scx == 4
    assert np.isclose(s_aa[0], 67 / 5.3 / 15)
    wl = 200 * 5.3 * 15
    assert np.isclose(s_dec[0], np.sqrt(wl ** 2 - 67 ** 2) / 5.3 / 60)
    assert not G.decimation(2.0, (15.0, 15.0))[0]

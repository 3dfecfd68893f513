"""The Goldstein oracle (oracle/goldstein.py) pinned: against the reference's own two filters
(tests/golden/goldstein_*.npz, written by tests/golden/make_goldstein_fixtures.py), and by closed
forms that fix the per-patch arithmetic: the window, the alpha = 0 identity, a single spectral
line, the overlap-add, a brute-force DFT.  Planted mistakes in the oracle fail these checks."""
import glob
import os

import numpy as np
import pytest

from oracle import goldstein as G

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = sorted(glob.glob(os.path.join(HERE, "goldstein_*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(f) for f in FIX])
@pytest.mark.parametrize("mode,key", [("count", "out_torch"), ("weight", "out_numpy")])
def test_oracle_matches_reference_fixture(path, mode, key):
    d = np.load(path)
    py, px = (int(v) for v in d["psize"])
    out = G.filter(d["phase"], d["corr"], py, px, float(d["threshold"]), mode)
    ref = d[key]
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = ~np.isnan(ref)
    assert np.abs(out[v] - ref[v]).max() <= 2e-6 * np.abs(ref[v]).max()


def test_window_closed_form():
    w = G.triangle_window(32, 16, np.float64)
    # separable, symmetric, zero on the first/last row and column, 1 in the centre 2 x 2 block
    assert np.allclose(w, np.outer(w[:, 7], w[15, :]) / w[15, 7])
    assert np.array_equal(w, w[::-1, ::-1])
    assert (w[0] == 0).all() and (w[:, 0] == 0).all()
    assert w[15, 7] == 1.0 and w[16, 8] == 1.0
    assert np.isclose(w[3, 2], (3 / 15) * (2 / 7))


def _interior(ny, nx, p, mode):
    """Pixels covered only by patches that lie inside the image: the last patch along each axis
    reaches into the zero padding, where corr reads 0, so its alpha is not 0."""
    ys, xs = G.patch_grid(ny, nx, p, p, mode)
    bad = np.zeros((ys[-1] + p, xs[-1] + p), bool)
    for y0 in ys:
        for x0 in xs:
            if y0 + p > ny or x0 + p > nx:
                bad[y0:y0 + p, x0:x0 + p] = True
    return ~bad[:ny, :nx]


def test_alpha_zero_weight_mode_is_identity():
    """corr = 1 -> alpha = 0 -> each patch passes unchanged; weight normalisation returns the input
    wherever some window weight covers the pixel."""
    rng = np.random.default_rng(0)
    ny, nx = 48, 40
    x = (rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))).astype(np.complex64)
    out = G.filter(x, np.ones((ny, nx), np.float32), 16, 16, 0.5, "weight")
    m = _interior(ny, nx, 16, "weight") & ~np.isnan(out)
    assert m[1:32, 1:24].all()
    assert np.allclose(out[m], x[m], atol=1e-9, rtol=1e-9)


def test_alpha_zero_count_mode_closed_form():
    """alpha = 0: out = x * (sum of window values of the covering patches) / (their number)."""
    rng = np.random.default_rng(1)
    ny, nx, p = 40, 56, 16
    x = (rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))).astype(np.complex64)
    out = G.filter(x, np.ones((ny, nx), np.float32), p, p, 0.0, "count")
    w = G.triangle_window(p, p, np.float64)
    ys, xs = G.patch_grid(ny, nx, p, p, "count")
    ws = np.zeros((ys[-1] + p, xs[-1] + p)); n = np.zeros_like(ws)
    for y0 in ys:
        for x0 in xs:
            ws[y0:y0 + p, x0:x0 + p] += w
            n[y0:y0 + p, x0:x0 + p] += 1
    expect = x * (ws / n)[:ny, :nx]
    m = _interior(ny, nx, p, "count")
    assert m.sum() > 0.5 * m.size
    assert np.allclose(out[m], expect[m], atol=1e-9, rtol=1e-9)


def test_single_spectral_line_scales_by_its_magnitude_power():
    """A patch-periodic plane wave has one non-zero DFT bin of magnitude A P^2: every patch output
    is the input times (A P^2)^alpha, so the weight-normalised result is x (A P^2)^alpha."""
    p, ny, nx = 16, 48, 48
    y, x = np.mgrid[0:ny, 0:nx]
    A = 0.7
    z = (A * np.exp(2j * np.pi * (3 * y + 5 * x) / p)).astype(np.complex128)
    corr = np.full((ny, nx), 0.4, np.float32)
    out = G.filter(z, corr, p, p, 0.5, "weight")
    alpha = 1.0 - 0.4
    m = _interior(ny, nx, p, "weight") & ~np.isnan(out)
    assert m.sum() > 0.5 * m.size
    assert np.allclose(out[m], z[m] * (A * p * p) ** alpha, rtol=1e-6)


def test_patch_math_against_brute_force_dft():
    """One full patch (image = patch size, threshold 0): the output at interior pixels is the
    brute-force inverse DFT of the weighted spectrum times the window, over the count of covering
    patches (here the 3 x 3 grid of half-overlapping patches)."""
    rng = np.random.default_rng(2)
    p = 16
    x = (rng.normal(size=(p, p)) + 1j * rng.normal(size=(p, p)))
    corr = rng.uniform(0, 1, (p, p))
    w = G.triangle_window(p, p, np.float64)
    alpha = 1 - (w * corr).sum() / w.sum()
    k = np.arange(p)
    D = np.exp(-2j * np.pi * np.outer(k, k) / p)
    F = D @ x @ D
    F = F * np.maximum(np.abs(F), 1e-10) ** alpha
    patch = (np.conj(D) @ F @ np.conj(D)) / (p * p) * w
    out = G.filter(x, corr, p, p, 0.0, "count")
    # patch origins are 0 and 8 along each axis, so pixel (5, 6) lies only in the patch at the
    # origin, which is the whole image
    assert np.isclose(out[5, 6], patch[5, 6], rtol=1e-9, atol=1e-12)


def test_planted_mistakes_are_caught():
    """A sign error in the spectral weight or a dropped window fails the closed forms above."""
    rng = np.random.default_rng(3)
    p = 16
    x = (rng.normal(size=(p, p)) + 1j * rng.normal(size=(p, p)))
    corr = rng.uniform(0, 1, (p, p))
    w = G.triangle_window(p, p, np.float64)
    good = G.filter(x, corr, p, p, 0.0, "count")[5, 6]
    alpha = 1 - (w * corr).sum() / w.sum()
    k = np.arange(p)
    D = np.exp(-2j * np.pi * np.outer(k, k) / p)
    F = D @ x @ D
    wrong_sign = ((np.conj(D) @ (F * np.abs(F) ** (-alpha)) @ np.conj(D)) / p ** 2 * w)[5, 6]
    no_window = ((np.conj(D) @ (F * np.abs(F) ** alpha) @ np.conj(D)) / p ** 2)[5, 6]
    assert not np.isclose(good, wrong_sign) and not np.isclose(good, no_window)


def test_grids_of_the_two_reference_versions():
    # n - s divisible by s: same grid; otherwise the torch grid has one more patch
    assert len(G.patch_grid(48, 48, 16, 16, "count")[0]) == len(G.patch_grid(48, 48, 16, 16, "weight")[0]) == 6
    assert len(G.patch_grid(45, 37, 16, 16, "count")[0]) == 6
    assert len(G.patch_grid(45, 37, 16, 16, "weight")[0]) == 5
    with pytest.raises(ValueError):
        G.patch_grid(7, 40, 16, 16)

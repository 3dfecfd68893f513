"""Pins of the CPU oracles against mathematics that fixes each step (CPU only).

The reference ships no golden vectors or assertions for this path (its only test,
tests/test_unwrap_benchmark/test_unwrap_benchmark.py, never reaches the IRLS solver), so every
oracle function is checked here against an independent statement of the same quantity:
sparse-matrix brute force, library routines (scipy.fft, numpy.linalg), closed forms and
invariants.  The `planted_*` tests show that a plausible mistake in an oracle function (a
dropped term, a flipped sign, a wrong index) fails these checks.
"""
import numpy as np
import pytest
import scipy.fft as sfft
import scipy.sparse as sp
import torch

from core_amd import synthetic as S
from oracle import irls_fp64 as O, irls_torch32 as T32, post


def diff_ops(H, W):
    """Forward-difference operators with a zero last column/row (Neumann), as sparse matrices."""
    def d1(n):
        m = sp.lil_matrix((n, n))
        for k in range(n - 1):
            m[k, k] = -1.0
            m[k, k + 1] = 1.0
        return m.tocsr()
    Dx = sp.kron(sp.identity(H), d1(W)).tocsr()   # along columns (x)
    Dy = sp.kron(d1(H), sp.identity(W)).tocsr()   # along rows (y)
    return Dx, Dy


def rng_fields(H, W, seed=0):
    r = np.random.default_rng(seed)
    return r.normal(size=(H, W)), r.uniform(0, 2, size=(H, W)), r.uniform(0, 2, size=(H, W))


@pytest.mark.parametrize("H,W", [(1, 1), (1, 7), (6, 1), (5, 4), (9, 13)])
def test_laplacian_equals_sparse_DtWD(H, W):
    x, wx, wy = rng_fields(H, W)
    Dx, Dy = diff_ops(H, W)
    A = Dx.T @ sp.diags(wx.ravel()) @ Dx + Dy.T @ sp.diags(wy.ravel()) @ Dy
    got = O.laplacian(x, wx, wy)
    np.testing.assert_allclose(got.ravel(), A @ x.ravel(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("H,W", [(4, 6), (11, 8)])
def test_rhs_equals_sparse_DtWt(H, W):
    _, wx, wy = rng_fields(H, W, 1)
    r = np.random.default_rng(2)
    tx, ty = r.normal(size=(H, W)), r.normal(size=(H, W))
    Dx, Dy = diff_ops(H, W)
    want = Dx.T @ (wx * tx).ravel() + Dy.T @ (wy * ty).ravel()
    np.testing.assert_allclose(O.divergence_rhs(tx, ty, wx, wy).ravel(), want, rtol=1e-12, atol=1e-12)


def test_laplacian_symmetric_psd_and_kills_constants():
    x, wx, wy = rng_fields(12, 10, 3)
    y = np.random.default_rng(4).normal(size=x.shape)
    assert np.isclose(np.vdot(O.laplacian(x, wx, wy), y), np.vdot(x, O.laplacian(y, wx, wy)))
    assert np.vdot(x, O.laplacian(x, wx, wy)) >= 0
    np.testing.assert_allclose(O.laplacian(np.ones_like(x), wx, wy), 0, atol=1e-12)


@pytest.mark.parametrize("H,W", [(5, 7), (8, 8), (13, 3)])
def test_eigenvalues_are_dct_spectrum_of_unweighted_laplacian(H, W):
    lam = O.eigenvalues(H, W)
    one = np.ones((H, W))
    for i, j in [(0, 1), (1, 0), (2, 3), (H - 1, W - 1)]:
        if i >= H or j >= W:
            continue
        v = np.cos(np.pi * i * (2 * np.arange(H)[:, None] + 1) / (2 * H)) * \
            np.cos(np.pi * j * (2 * np.arange(W)[None, :] + 1) / (2 * W))
        np.testing.assert_allclose(O.laplacian(v, one, one), -lam[i, j] * v, atol=1e-12)
    # whole spectrum against a dense eigensolver
    Dx, Dy = diff_ops(H, W)
    A = (Dx.T @ Dx + Dy.T @ Dy).toarray()
    ev = np.sort(np.linalg.eigvalsh(A))
    lam2 = -O.eigenvalues(H, W)
    lam2[0, 0] = 0
    np.testing.assert_allclose(np.sort(lam2.ravel()), ev, atol=1e-10)


@pytest.mark.parametrize("N", [1, 2, 3, 8, 13, 64, 97, 100])
def test_torch_dct_restatement_matches_scipy(N):
    x = np.random.default_rng(N).normal(size=(3, N))
    X = T32.dct(torch.from_numpy(x)).numpy()
    np.testing.assert_allclose(X, sfft.dct(x, type=2, axis=-1), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(T32.idct(torch.from_numpy(X)).numpy(), x, rtol=1e-10, atol=1e-10)
    # closed form of the unnormalised DCT-II
    n = np.arange(N)
    direct = np.array([2 * np.sum(x * np.cos(np.pi * k * (2 * n + 1) / (2 * N)), axis=-1) for k in range(N)]).T
    np.testing.assert_allclose(X, direct, rtol=1e-9, atol=1e-9)


def test_preconditioner_inverts_unweighted_laplacian():
    H, W = 9, 14
    x = np.random.default_rng(5).normal(size=(H, W))
    one = np.ones((H, W))
    lam = O.eigenvalues(H, W)
    z = O.precondition(O.laplacian(x, one, one), lam)
    np.testing.assert_allclose(z, x - x.mean(), atol=1e-10)


def test_pcg_converges_to_least_squares_solution():
    H, W = 7, 6
    _, wx, wy = rng_fields(H, W, 6)
    wx[:, -1] = 0
    wy[-1, :] = 0
    Dx, Dy = diff_ops(H, W)
    A = (Dx.T @ sp.diags(wx.ravel()) @ Dx + Dy.T @ sp.diags(wy.ravel()) @ Dy).toarray()
    b = np.random.default_rng(7).normal(size=H * W)
    b -= b.mean()
    xs = np.linalg.lstsq(A, b, rcond=None)[0]
    x, n, reason = O.pcg(b.reshape(H, W), wx, wy, np.zeros((H, W)), O.eigenvalues(H, W), 200, 1e-12)
    d = x.ravel() - xs
    np.testing.assert_allclose(d - d.mean(), 0, atol=1e-8)
    assert reason in ('rnorm', 'pAp', 'rz')


def test_targets_and_edge_weights_closed_form():
    H, W = 6, 7
    r = np.random.default_rng(8)
    phi = r.uniform(-np.pi, np.pi, size=(H, W))
    tx, ty = O.targets(phi)
    d = phi[:, 1:] - phi[:, :-1]
    np.testing.assert_allclose(tx[:, :-1], d - 2 * np.pi * np.round(d / (2 * np.pi)), atol=1e-12)
    assert (tx[:, -1] == 0).all() and (ty[-1, :] == 0).all()
    w = r.uniform(0.01, 1, size=(H, W))
    valid = r.uniform(size=(H, W)) > 0.2
    wx, wy = O.edge_weights(w, valid)
    for i in range(H):
        for j in range(W - 1):
            assert wx[i, j] == pytest.approx((w[i, j] + w[i, j + 1]) / 2 * (valid[i, j] and valid[i, j + 1]))
    for i in range(H - 1):
        for j in range(W):
            assert wy[i, j] == pytest.approx((w[i, j] + w[i + 1, j]) / 2 * (valid[i, j] and valid[i + 1, j]))


def test_irls_recovers_residue_free_ramp_exactly():
    H, W = 40, 56
    y, x = np.mgrid[0:H, 0:W]
    true = 0.25 * x + 0.15 * y + 2.0 * np.sin(x / 9.0)
    psi = S.wrap_f64(true).astype(np.float32)
    u, cc = O.irls_unwrap_2d(psi, None, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
    d = u.astype(np.float64) - true
    assert np.ptp(d) < 1e-3
    k = np.round((u.astype(np.float64) - psi) / (2 * np.pi))
    k_true = np.round((true - psi) / (2 * np.pi))
    assert np.mean((k - k[0, 0]) == (k_true - k_true[0, 0])) == 1.0
    assert (cc == 1).all()


def test_irls_fixed_point_is_stationary_point_of_l1_objective():
    H, W = 24, 20
    psi, coh, _ = S.generate(S.Workload(H, W, 9, ramp_max=20.0))
    pro = post.prologue(psi, coh)
    _, _, nm, pf, wf = pro
    eps = 1e-2
    u = O.solve(pf, wf, ~nm, max_iter=60, tol=0, cg_max_iter=200, cg_tol=0, epsilon=eps)
    tx, ty = O.targets(pf)
    wx, wy = O.edge_weights(wf, ~nm)
    rx = np.zeros_like(u); ry = np.zeros_like(u)
    rx[:, :-1] = u[:, 1:] - u[:, :-1]
    ry[:-1, :] = u[1:, :] - u[:-1, :]
    rx -= tx; ry -= ty
    # grad F = D^T (w r / sqrt(r^2 + eps^2))
    gx = wx * rx / np.sqrt(rx ** 2 + eps ** 2)
    gy = wy * ry / np.sqrt(ry ** 2 + eps ** 2)
    g = O.divergence_rhs(gx, gy, np.ones_like(gx), np.ones_like(gy))
    scale = np.abs(O.divergence_rhs(np.abs(wx), np.abs(wy), np.ones_like(gx), np.ones_like(gy))).max()
    assert np.abs(g).max() < 2e-2 * max(scale, 1.0)


def test_fp32_restatement_tracks_fp64():
    psi, coh, _ = S.generate(S.Workload(64, 80, 11, ramp_max=30.0, fault=True, mask_frac=0.05))
    u64, cc64 = O.irls_unwrap_2d(psi, coh, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
    u32, cc32 = T32.irls_unwrap_2d(psi, coh, max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
    v = ~np.isnan(psi)
    assert np.sqrt(np.mean((u32[v].astype(np.float64) - u64[v]) ** 2)) < 1e-4
    assert np.array_equal(cc32, cc64)


def test_torch_restatement_in_float64_equals_fp64_oracle():
    """The torch restatement run in float64 is the FP64 oracle's trajectory (same operations,
    different libraries): it stands in for the numpy oracle where that is too slow (large grids,
    run on the HIP device by the GPU tests)."""
    for H, W, seed, fault in [(64, 80, 11, True), (37, 101, 12, False)]:
        psi, coh, _ = S.generate(S.Workload(H, W, seed, ramp_max=30.0, fault=fault, mask_frac=0.05))
        kw = dict(max_iter=10, tol=0, cg_max_iter=50, cg_tol=0)
        _, _, r64 = O.irls_unwrap_2d(psi, coh, return_raw=True, **kw)
        _, _, rt = T32.irls_unwrap_2d(psi, coh, return_raw=True, dtype=torch.float64, **kw)
        assert rt.dtype == np.float64
        v = ~np.isnan(psi[~np.all(np.isnan(psi), axis=1)][:, ~np.all(np.isnan(psi), axis=0)])
        assert np.abs(rt - r64)[v].max() < 1e-9 * max(1.0, np.abs(r64).max())


def test_epilogue_invariants():
    r = np.random.default_rng(12)
    H, W = 30, 40
    psi = r.uniform(-np.pi, np.pi, size=(H, W)).astype(np.float32)
    psi[:, 19:21] = np.nan  # two components
    psi[5:7, 5:7] = np.nan
    nm = np.isnan(psi)
    u = np.where(nm, 0, psi + 2 * np.pi * 7 + 0.3).astype(np.float32)
    out, cc = post.epilogue(u, psi, nm, conncomp_size=30)
    v = ~nm
    # after the median shift and circular-mean alignment the output rewraps onto psi
    k = (out[v].astype(np.float64) - psi[v]) / (2 * np.pi)
    np.testing.assert_allclose(k, np.round(k), atol=1e-5)
    assert set(np.unique(cc)) == {0, 1, 2}
    assert (cc[nm] == 0).all()
    sizes = [np.sum(cc == 1), np.sum(cc == 2)]
    assert sizes[0] >= sizes[1]


def test_label4_matches_brute_force_flood_fill():
    r = np.random.default_rng(13)
    valid = r.uniform(size=(17, 23)) > 0.45
    lab, n = post.label4(valid)
    # brute force: BFS in raster order
    ref = np.zeros(valid.shape, np.int32)
    nxt = 1
    for i in range(valid.shape[0]):
        for j in range(valid.shape[1]):
            if valid[i, j] and ref[i, j] == 0:
                stack = [(i, j)]
                ref[i, j] = nxt
                while stack:
                    a, b = stack.pop()
                    for da, db in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                        c, d = a + da, b + db
                        if 0 <= c < valid.shape[0] and 0 <= d < valid.shape[1] and valid[c, d] and ref[c, d] == 0:
                            ref[c, d] = nxt
                            stack.append((c, d))
                nxt += 1
    assert n == nxt
    assert np.array_equal(lab, ref)


# ---------------------------------------------------------------------------------------------
# Planted mistakes: each must be caught by one of the pins above
# ---------------------------------------------------------------------------------------------

def _bad_laplacian_drop_term(x, wx, wy):
    dx = np.zeros_like(x); dy = np.zeros_like(x)
    dx[:, :-1] = x[:, 1:] - x[:, :-1]
    dy[:-1, :] = x[1:, :] - x[:-1, :]
    dx *= wx; dy *= wy
    out = np.zeros_like(x)
    out[:, 1:] -= dx[:, :-1]
    out[:, :-1] += dx[:, :-1]
    out[1:, :] -= dy[:-1, :]          # the "+ dy_i" term is missing
    return -out


def _bad_laplacian_wrong_index(x, wx, wy):
    dx = np.zeros_like(x); dy = np.zeros_like(x)
    dx[:, :-1] = x[:, 1:] - x[:, :-1]
    dy[:-1, :] = x[1:, :] - x[:-1, :]
    dx *= wx; dy *= wy
    out = np.zeros_like(x)
    out[:, 1:] -= dx[:, 1:]           # shifted weight index
    out[:, :-1] += dx[:, :-1]
    out[1:, :] -= dy[:-1, :]
    out[:-1, :] += dy[:-1, :]
    return -out


@pytest.mark.parametrize("bad", [_bad_laplacian_drop_term, _bad_laplacian_wrong_index])
def test_planted_laplacian_mistakes_are_caught(monkeypatch, bad):
    monkeypatch.setattr(O, "laplacian", bad)
    with pytest.raises(AssertionError):
        test_laplacian_equals_sparse_DtWD(9, 13)


def test_planted_sign_flip_in_preconditioner_is_caught(monkeypatch):
    orig = O.precondition
    monkeypatch.setattr(O, "precondition", lambda r, lam: -orig(r, lam))
    with pytest.raises(AssertionError):
        test_preconditioner_inverts_unweighted_laplacian()


def test_planted_transposed_dct_axes_is_caught(monkeypatch):
    monkeypatch.setattr(O, "dct2", lambda x: sfft.dct(sfft.dct(x, type=2, axis=0), type=3, axis=1))
    with pytest.raises(AssertionError):
        test_preconditioner_inverts_unweighted_laplacian()


def test_planted_rhs_sign_is_caught(monkeypatch):
    orig = O.divergence_rhs
    monkeypatch.setattr(O, "divergence_rhs", lambda tx, ty, wx, wy: -orig(tx, ty, wx, wy))
    with pytest.raises(AssertionError):
        test_rhs_equals_sparse_DtWt(4, 6)


def test_injected_nonfinite_x_breaks_pcg_and_restores_best_iterate():
    """The oracle side of the library's test hook: +inf planted in p makes x non-finite; PCG breaks
    there ('x_nonfinite'), the IRLS loop restores u_best and stops, and u_best is the iterate of
    the lowest residual so far (utils_unwrap2d.py:1444-1449, :1551-1558, :1599-1604)."""
    psi, coh, _ = S.generate(S.Workload(40, 52, 21, ramp_max=20.0, fault=True, mask_frac=0.05))
    pro = post.prologue(psi, coh)
    _, _, nm, pf, wf = pro
    kw = dict(max_iter=6, tol=0, cg_max_iter=12, cg_tol=0, epsilon=1e-2)
    for inj in [(0, 0), (3, 5)]:
        tr = []
        u = O.solve(pf, wf, ~nm, trace=tr, inject=inj, **kw)
        assert [t['reason'] for t in tr][-1] == 'x_nonfinite' and tr[-1]['stop'] == 'nonfinite'
        assert len(tr) == inj[0] + 1 and np.isfinite(u).all()
        # u_best: the centred iterate with the lowest residual among iterations 0 .. inj[0]
        ref = []
        O.solve(pf, wf, ~nm, trace=ref, **{**kw, 'max_iter': inj[0]})
        best_it = int(np.argmin([t['residual'] for t in tr]))
        assert best_it <= inj[0]
        u_prefix = O.solve(pf, wf, ~nm, **{**kw, 'max_iter': best_it})
        vf = (~nm).astype(np.float64)
        expect = u_prefix - (u_prefix * vf).sum() / vf.sum()
        np.testing.assert_allclose(u, expect, rtol=0, atol=1e-12)

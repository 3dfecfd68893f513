"""GPU tests of the tridiagonal column pass (k_col_trc): after the row DCT-II, the column DCT pair
with the division by -lambda is replaced by the float64 solves (L_H + mu_k I) x = c_k (the same
operator, utils_unwrap2d.py:1390-1403).  Checked against scipy's FP64 DCT preconditioner on
edge shapes (one row, one or two columns, prime heights up to 12007, Bluestein rows, heights at
the chunk and band edges 512 / 513 / 1536 / 1537 / 3072) in batches with slab, transposed and
long images; against the DCT column pass; solver parity with the FP64 and float32 oracles;
batched = plain bitwise."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)

SHAPES = [(1, 5), (2, 7), (3, 3), (17, 1), (1, 1), (4, 3), (64, 48), (127, 211), (1030, 5000), (1201, 4999),
          (12007, 9), (5, 12007), (300, 2), (8192, 70), (12800, 8), (96, 130), (512, 100), (513, 65),
          (1536, 31), (1537, 40), (3072, 33), (33, 77)]


def _precond64(r):
    import scipy.fft as sf
    h, w = r.shape
    lam = 2 * np.cos(np.pi * np.arange(h) / h)[:, None] + 2 * np.cos(np.pi * np.arange(w) / w)[None, :] - 4
    lam[0, 0] = 1.0
    X = sf.dctn(r.astype(np.float64), type=2) / -lam
    X[0, 0] = 0.0
    return sf.idctn(X, type=2)


def _fields(shapes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for h, w in shapes:
        y, x = np.mgrid[0:h, 0:w]
        smooth = np.sin(2 * np.pi * x / max(w, 2) * 3) + np.cos(2 * np.pi * y / max(h, 2) * 2)
        out.append((rng.normal(size=(h, w)) + 5 * smooth).astype(np.float32))
    return out


def _precond_gpu(hip, rs):
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    shapes = [r.shape for r in rs]
    ws = Workspace(shapes, hip)
    z = precondition_device(pack(rs, ws, hip), shapes, workspace=ws).cpu().numpy()
    out = [image_view(z, ws, k).copy() for k in range(len(rs))]
    ws.release()
    return out


@pytest.mark.parametrize("mode", ["tridiag", "auto"])
def test_tridiag_preconditioner_matches_scipy_fp64(hip, mode):
    from core_amd import _lib
    rs = _fields(SHAPES, 21)
    try:
        _lib.set_column_mode(_lib.COLUMN_TRIDIAG if mode == "tridiag" else _lib.COLUMN_AUTO)
        strat = _lib.describe_columns(SHAPES)
        print(f"tridiag preconditioner, mode {mode}: {strat}", flush=True)
        if mode == "tridiag":
            short = [s for s, (h, _) in zip(strat, SHAPES) if h <= 12288]
            assert short == ["tridiag"] * len(short)
        else:
            assert strat[SHAPES.index((1030, 5000))] == "tridiag" and strat[SHAPES.index((8192, 70))] == "slab_register"
        zs = _precond_gpu(hip, rs)
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)
    for (h, w), r, z in zip(SHAPES, rs, zs):
        z64 = _precond64(r)
        scale = max(np.abs(z64).max(), 1e-30)
        err = np.abs(z.astype(np.float64) - z64).max() / scale
        print(f"  {h}x{w}: max rel err {err:.2e}", flush=True)
        if h * w == 1:
            assert np.all(z == 0)
        else:
            assert err < 2e-5, (h, w, err)


def test_tridiag_equals_dct_columns(hip):
    """The tridiagonal and the DCT column pass agree to float32 rounding on the stack's shapes."""
    from core_amd import _lib
    shapes = [(1030, 5000), (1201, 4999), (1499, 4013)]
    rs = _fields(shapes, 5)
    out = {}
    try:
        for mode in (_lib.COLUMN_TRIDIAG, _lib.COLUMN_TRANSPOSE):
            _lib.set_column_mode(mode)
            out[mode] = _precond_gpu(hip, rs)
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)
    for sh, a, b in zip(shapes, out[_lib.COLUMN_TRIDIAG], out[_lib.COLUMN_TRANSPOSE]):
        err = np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
        print(f"  {sh}: tridiag vs DCT columns max rel diff {err:.2e}", flush=True)
        assert err < 2e-5


@pytest.mark.parametrize("H,W,seed,fault,mask", [(103, 64, 3, False, 0.0), (127, 211, 4, True, 0.05),
                                                 (331, 257, 6, True, 0.0)])
def test_tridiag_solve_parity(hip, H, W, seed, fault, mask):
    """Solver with tridiagonal columns (automatic for these Bluestein heights) against the FP64 and
    float32 trajectories: RMS <= 1e-3 rad, cycle agreement >= 99.99%, equal labels; and against
    the DCT column pass."""
    from core_amd import _lib, irls_unwrap_2d
    from oracle import irls_fp64, irls_torch32
    wl = S.Workload(H, W, seed, ramp_max=40.0, n_bumps=6, fault=fault, mask_frac=mask)
    psi, coh, _ = S.generate(wl)
    print(f"tridiag solve {H}x{W} seed={seed}", flush=True)
    assert _lib.describe_columns([(H, W)]) == ["tridiag"]
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    v = ~np.isnan(psi)
    rms = float(np.sqrt(np.mean((u[v].astype(np.float64) - u64[v]) ** 2)))
    k1 = np.round((u[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    k2 = np.round((u32[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    print(f"  RMS vs FP64 {rms:.2e}, cycle agreement {np.mean(k1 == k2):.6f}", flush=True)
    assert np.array_equal(np.isnan(u), np.isnan(psi))
    assert rms <= 1e-3 and np.mean(k1 == k2) >= 0.9999
    assert np.array_equal(cc, cc64)
    try:
        _lib.set_column_mode(_lib.COLUMN_TRANSPOSE)
        ut, _ = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)
    assert np.abs(ut[v] - u[v]).max() < 1e-3


def test_tridiag_batched_equals_plain_bitwise(hip):
    from core_amd import _lib, irls_unwrap_2d, irls_unwrap_2d_batch
    shapes = [(127, 211), (256, 300), (96, 130), (1, 40), (50, 1), (103, 64)]
    cases = [S.generate(S.Workload(h, w, 40 + i, ramp_max=20.0, n_bumps=2))[:2] for i, (h, w) in enumerate(shapes)]
    strat = _lib.describe_columns(shapes)
    print(f"tridiag batch strategies {strat}", flush=True)
    assert "tridiag" in strat and "slab_register" in strat
    res = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', **FIXED)
    for (p_, c_), (ub, cb) in zip(cases, res):
        up, cp = irls_unwrap_2d(p_, c_, device='cuda', **FIXED)
        assert np.array_equal(ub, up, equal_nan=True) and np.array_equal(cb, cp)

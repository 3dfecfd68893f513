"""GPU tests of the slab column pass (the column DCT pair as three four-step passes on the
row-major field): the preconditioner against scipy in FP64 for column lengths of every split
shape, in batches mixed with transposed, Bluestein and long-line images; solver parity with the
FP64 oracle; agreement with the transposed column pass (RMS, cycle counts); batched = plain
bitwise."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)

# column lengths 4 .. 12288: power-of-two, 3- and 5-smooth, odd widths, one-column images;
# 127 x 211 (Bluestein columns), 12800 x 8 (long columns) and 12 x 12500 (long rows) share the batch
SHAPES = [(4, 3), (64, 48), (96, 130), (250, 33), (1000, 7), (480, 642), (8192, 70), (12288, 5), (300, 1),
          (127, 211), (12800, 8), (12, 12500)]


def _precond64(r):
    import scipy.fft as sf
    h, w = r.shape
    lam = 2 * np.cos(np.pi * np.arange(h) / h)[:, None] + 2 * np.cos(np.pi * np.arange(w) / w)[None, :] - 4
    lam[0, 0] = 1.0
    X = sf.dctn(r.astype(np.float64), type=2) / -lam
    X[0, 0] = 0.0
    return sf.idctn(X, type=2)


@pytest.mark.parametrize("mode", ["slab", "auto"])
def test_slab_preconditioner_matches_scipy_fp64(hip, mode):
    from core_amd import _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    rng = np.random.default_rng(11)
    rs = []
    for h, w in SHAPES:
        y, x = np.mgrid[0:h, 0:w]
        smooth = np.sin(2 * np.pi * x / max(w, 2) * 3) + np.cos(2 * np.pi * y / max(h, 2) * 2)
        rs.append((rng.normal(size=(h, w)) + 5 * smooth).astype(np.float32))
    try:
        _lib.set_column_mode(_lib.COLUMN_SLAB if mode == "slab" else _lib.COLUMN_AUTO)
        print(f"slab preconditioner ({mode}) {SHAPES}: {_lib.describe_columns(SHAPES)}", flush=True)
        ws = Workspace(SHAPES, hip)
        z = precondition_device(pack(rs, ws, hip), SHAPES, workspace=ws)
        zh = z.cpu().numpy()
        for k, ((h, w), r) in enumerate(zip(SHAPES, rs)):
            z64 = _precond64(r)
            err = np.abs(image_view(zh, ws, k).astype(np.float64) - z64).max() / np.abs(z64).max()
            print(f"  {h}x{w}: max rel err {err:.2e}", flush=True)
            assert err < 2e-5, (h, w, err)
        ws.release()
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)


# one power-of-two split per workspace: the register-stage kernels (N1 x N2 = 16x16 .. 128x64)
FAST = [[(256, 256)], [(512, 40)], [(1024, 33)], [(2048, 65), (2048, 3)], [(4096, 7)], [(8192, 70), (8192, 1)]]


@pytest.mark.parametrize("shapes", FAST, ids=lambda s: "x".join(map(str, s[0])))
def test_slab_fast_preconditioner_matches_scipy_fp64(hip, shapes):
    from core_amd import _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    rng = np.random.default_rng(12)
    rs = [rng.normal(size=s).astype(np.float32) for s in shapes]
    print(f"fast slab preconditioner {shapes}", flush=True)
    assert _lib.describe_columns(shapes) == ["slab_register"] * len(shapes)
    ws = Workspace(shapes, hip)
    z = precondition_device(pack(rs, ws, hip), shapes, workspace=ws)
    zh = z.cpu().numpy()
    for k, ((h, w), r) in enumerate(zip(shapes, rs)):
        z64 = _precond64(r)
        err = np.abs(image_view(zh, ws, k).astype(np.float64) - z64).max() / np.abs(z64).max()
        print(f"  {h}x{w}: max rel err {err:.2e}", flush=True)
        assert err < 2e-5, (h, w, err)
    ws.release()


@pytest.mark.parametrize("ept", [0, 24])
def test_long_rows_twelve_per_thread(hip, ept):
    """5-smooth rows above 8192 points: 12 values per thread on up to 1024 threads (the default)
    and 24 on 512 (a plan override), with compile-time (12288, 10240) and runtime-described
    (9000 = 2^3 3^2 5^3) Stockham schedules, against scipy FP64."""
    from core_amd import _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    shapes = [(8, 12288), (7, 10240), (9, 9000)]
    rng = np.random.default_rng(14)
    rs = [rng.normal(size=s).astype(np.float32) for s in shapes]
    try:
        for _, w in shapes:
            fl, rad, e = _lib.plan_describe(w)
            assert e == 12
            if ept:
                _lib.plan_override(w, fl, rad, ept)
        print(f"long rows, ept {ept or 'default'}: {shapes}", flush=True)
        ws = Workspace(shapes, hip)
        zh = precondition_device(pack(rs, ws, hip), shapes, workspace=ws).cpu().numpy()
        for k, ((h, w), r) in enumerate(zip(shapes, rs)):
            z64 = _precond64(r)
            err = np.abs(image_view(zh, ws, k).astype(np.float64) - z64).max() / np.abs(z64).max()
            print(f"  {h}x{w}: max rel err {err:.2e}", flush=True)
            assert err < 2e-5, (h, w, err)
        ws.release()
    finally:
        _lib.plan_reset()


def test_mixed_row_lengths_share_one_thread_shape(hip):
    """One launch, one elements-per-thread value: rows of 12288 (12 per thread, 1024 threads) in
    the same batch as 8192-point direct, Bluestein (5000, 211) and short rows, against scipy
    FP64."""
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    shapes = [(16, 8192), (12, 5000), (8, 12288), (10, 211), (6, 300)]
    rng = np.random.default_rng(15)
    rs = [rng.normal(size=s).astype(np.float32) for s in shapes]
    print(f"mixed row lengths {shapes}", flush=True)
    ws = Workspace(shapes, hip)
    zh = precondition_device(pack(rs, ws, hip), shapes, workspace=ws).cpu().numpy()
    for k, ((h, w), r) in enumerate(zip(shapes, rs)):
        z64 = _precond64(r)
        err = np.abs(image_view(zh, ws, k).astype(np.float64) - z64).max() / np.abs(z64).max()
        print(f"  {h}x{w}: max rel err {err:.2e}", flush=True)
        assert err < 2e-5, (h, w, err)
    ws.release()


def test_preconditioner_unaligned_buffers(hip):
    """Input and output arenas at a 4-byte (not 8-byte) offset: the row DCT-III then stores one
    float at a time instead of float2 pairs, with the same arithmetic, so z is bitwise equal to the
    aligned call's; widths even and odd, power-of-two, 5-smooth and Bluestein, slab and transposed
    columns."""
    import torch
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    shapes = [(64, 48), (127, 211), (300, 642), (256, 256), (96, 130)]
    rng = np.random.default_rng(13)
    rs = [rng.normal(size=s).astype(np.float32) for s in shapes]
    ws = Workspace(shapes, hip)
    r = pack(rs, ws, hip)
    print(f"preconditioner, aligned and 4-byte offset arenas {shapes}", flush=True)
    z = precondition_device(r, shapes, workspace=ws)
    rbig = torch.zeros(ws.total + 1, dtype=torch.float32, device=hip)
    zbig = torch.full((ws.total + 1,), float("nan"), dtype=torch.float32, device=hip)
    r1, z1 = rbig[1:], zbig[1:]
    r1.copy_(r)
    assert r1.data_ptr() % 8 == 4 and z1.data_ptr() % 8 == 4
    precondition_device(r1, shapes, out=z1, workspace=ws)
    zh, z1h = z.cpu().numpy(), z1.cpu().numpy()
    for k, ((h, w), rk) in enumerate(zip(shapes, rs)):
        a, b = image_view(zh, ws, k), image_view(z1h, ws, k)
        assert np.array_equal(a, b), (h, w)
        z64 = _precond64(rk)
        assert np.abs(a.astype(np.float64) - z64).max() / np.abs(z64).max() < 2e-5, (h, w)
    ws.release()


@pytest.mark.parametrize("H,W,seed,fault,mask", [(96, 130, 1, True, 0.05), (250, 33, 2, False, 0.0),
                                                 (480, 642, 3, True, 0.0), (300, 1, 4, False, 0.0),
                                                 (512, 300, 5, True, 0.05), (1024, 96, 6, False, 0.0)])
def test_slab_solve_parity(hip, H, W, seed, fault, mask):
    """Slab and transposed column passes: both within the FP64 bar, the same labels, and the same
    solution to float32 rounding of the reordered arithmetic."""
    from core_amd import _lib, irls_unwrap_2d
    from oracle import irls_fp64
    wl = S.Workload(H, W, seed, ramp_max=min(40.0, 0.5 * W), n_bumps=6, fault=fault, mask_frac=mask)
    psi, coh, _ = S.generate(wl)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    v = ~np.isnan(psi)
    outs = {}
    try:
        for mode in (_lib.COLUMN_TRANSPOSE, _lib.COLUMN_SLAB):
            _lib.set_column_mode(mode)
            print(f"solve {H}x{W} column mode {mode}", flush=True)
            outs[mode] = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
            d = outs[mode][0][v].astype(np.float64) - u64[v]
            assert float(np.sqrt(np.mean(d * d))) <= 1e-3, (mode, H, W)
            assert np.array_equal(outs[mode][1], cc64)
        # the four-step order rounds differently; on faulted grids float32 differences grow over
        # the 500 PCG steps, so compare in RMS and by cycle count
        d = outs[1][0][v].astype(np.float64) - outs[3][0][v]
        assert float(np.sqrt(np.mean(d * d))) < 1e-4
        k1 = np.round((outs[1][0][v].astype(np.float64) - psi[v]) / (2 * np.pi))
        k3 = np.round((outs[3][0][v].astype(np.float64) - psi[v]) / (2 * np.pi))
        assert np.array_equal(k1, k3)
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)


def test_slab_batched_equals_plain_bitwise(hip):
    from core_amd import _lib, irls_unwrap_2d, irls_unwrap_2d_batch
    cases = []
    for i, (h, w) in enumerate([(96, 130), (64, 48), (127, 211), (250, 33)]):
        psi, coh, _ = S.generate(S.Workload(h, w, 30 + i, ramp_max=min(30.0, 0.5 * w), n_bumps=4, fault=i == 0))
        cases.append((psi, coh))
    try:
        _lib.set_column_mode(_lib.COLUMN_SLAB)
        print("slab batch of 4", flush=True)
        batch = irls_unwrap_2d_batch([c[0] for c in cases], [c[1] for c in cases], device='cuda', **FIXED)
        for (p_, c_), (ub, cb) in zip(cases, batch):
            up, cp = irls_unwrap_2d(p_, c_, device='cuda', **FIXED)
            assert np.array_equal(ub, up, equal_nan=True) and np.array_equal(cb, cp)
    finally:
        _lib.set_column_mode(_lib.COLUMN_AUTO)


def test_full_size_preconditioner_matches_scipy_fp64(hip):
    """At 8192 x 8192 (the bench shape: register-stage slab columns, 128 x 64 split) the whole
    transform chain (row DCT-II, three slab passes, row DCT-III) against scipy in FP64."""
    from core_amd import _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    n = 8192
    shapes = [(n, n)]
    assert _lib.describe_columns(shapes) == ["slab_register"]
    rng = np.random.default_rng(8)
    r = rng.standard_normal((n, n)).astype(np.float32)
    r += np.sin(np.arange(n, dtype=np.float32) / 700.0)[None, :]
    print("preconditioner 8192^2 (slab_register)", flush=True)
    ws = Workspace(shapes, hip)
    z = precondition_device(pack([r], ws, hip), shapes, workspace=ws)
    zg = image_view(z.cpu().numpy(), ws, 0).astype(np.float64)
    ws.release()
    import scipy.fft as sf
    lam = 2 * np.cos(np.pi * np.arange(n) / n)[:, None] + 2 * np.cos(np.pi * np.arange(n) / n)[None, :] - 4
    lam[0, 0] = 1.0
    X = sf.dctn(r.astype(np.float64), type=2, workers=-1)
    X /= -lam
    X[0, 0] = 0.0
    z64 = sf.idctn(X, type=2, workers=-1)
    err = float(np.abs(zg - z64).max() / np.abs(z64).max())
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < 2e-5

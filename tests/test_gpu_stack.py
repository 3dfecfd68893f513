"""GPU test of the stack array core: all pairs of a (pair, y, x) stack in one batched solve equal
the per-pair drop-in calls bitwise, including an all-NaN pair and the 2-D form; tiled unwrapping
with overlap margins equals the drop-in call on each extended tile."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu


def test_unwrap_stack_equals_per_pair_calls(hip):
    from core_amd import irls_unwrap_2d
    from core_amd.stack import unwrap_stack
    pairs = [S.generate(S.Workload(48, 64, 60 + i, ramp_max=30.0, n_bumps=3, fault=i == 1, mask_frac=0.05))[:2]
             for i in range(3)]
    ph = np.stack([p[0] for p in pairs])
    co = np.stack([p[1] for p in pairs])
    ph[2] = np.nan                                  # an all-NaN pair
    print("unwrap_stack 3 x 48 x 64", flush=True)
    unw, cc = unwrap_stack(ph, co, device='cuda')
    assert unw.shape == ph.shape and unw.dtype == np.float32 and cc.dtype == np.uint16
    for k in range(3):
        u1, c1 = irls_unwrap_2d(ph[k], co[k], device='cuda', max_iter=50, tol=1e-2, cg_max_iter=10, cg_tol=1e-3)
        assert np.array_equal(unw[k], u1, equal_nan=True)
        assert np.array_equal(cc[k], c1)
    assert np.isnan(unw[2]).all() and (cc[2] == 0).all()
    u2, c2 = unwrap_stack(ph[0], co[0], device='cuda')
    assert np.array_equal(u2, unw[0], equal_nan=True) and np.array_equal(c2, cc[0])


def test_unwrap_chunked_equals_per_tile_calls(hip):
    """Tiled unwrapping with overlap (all tiles of all pairs in one batched solve) equals the
    drop-in call on each extended tile, trimmed to its chunk, including an all-NaN tile."""
    from core_amd import irls_unwrap_2d
    from core_amd.stack import chunk_windows, unwrap_chunked
    pairs = [S.generate(S.Workload(100, 90, 80 + i, ramp_max=30.0, n_bumps=3, fault=i == 0, mask_frac=0.05))[:2]
             for i in range(2)]
    ph = np.stack([p[0] for p in pairs])
    co = np.stack([p[1] for p in pairs])
    ph[1, :48, :48] = np.nan                       # the first extended tile of pair 1 is all NaN
    kw = dict(max_iter=50, tol=1e-2, cg_max_iter=10, cg_tol=1e-3, epsilon=1e-2, conncomp_size=30)
    print("unwrap_chunked 2 x 100 x 90, chunks (40, 40), overlap 8", flush=True)
    out = unwrap_chunked(ph, co, chunks=(40, 40), overlap=8, device='cuda', **kw)
    wins = chunk_windows(100, 90, (40, 40), 8)
    for p in range(2):
        for (y0, y1, x0, x1), (ey0, ey1, ex0, ex1) in wins:
            u1, _ = irls_unwrap_2d(ph[p, ey0:ey1, ex0:ex1], co[p, ey0:ey1, ex0:ex1], device='cuda', **kw)
            core = u1[y0 - ey0:y1 - ey0, x0 - ex0:x1 - ex0]
            assert np.array_equal(out[p, y0:y1, x0:x1], core, equal_nan=True), (p, y0, x0)
    assert np.isnan(out[1, :40, :40]).all()

"""GPU test of the stack array core: all pairs of a (pair, y, x) stack in one batched solve equal
the per-pair drop-in calls bitwise, including an all-NaN pair and the 2-D form."""
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

"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py from the FP64
oracle).  CPU: the oracle still reproduces them.  GPU: the device solver matches their final
outputs to the parity bar."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "*.npz")) if not os.path.basename(f).startswith("goldstein_"))
KW = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)


def _load(p):
    with np.load(p, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_fixture(path):
    from oracle import irls_fp64 as O, post
    f = _load(path)
    crop, ph_c, nm, pf, wf = post.prologue(f["psi"], f["coh"])
    assert tuple(crop) == tuple(f["crop"])
    tx, ty = O.targets(pf)
    wx, wy = O.edge_weights(wf, ~nm)
    for a, b in ((tx, "tx"), (ty, "ty"), (wx, "wx"), (wy, "wy")):
        np.testing.assert_allclose(a, f[b], rtol=0, atol=1e-12)
    lam = O.eigenvalues(*pf.shape)
    np.testing.assert_allclose(O.precondition(f["rho"], lam), f["u0"], atol=1e-9)
    v = ~nm
    u = f["u0"] - (f["u0"] * v).sum() / v.sum()
    np.testing.assert_allclose(O.laplacian(u, f["wxi1"], f["wyi1"]), f["Au1"], atol=1e-9)
    np.testing.assert_allclose(O.divergence_rhs(tx, ty, f["wxi1"], f["wyi1"]), f["b1"], atol=1e-9)
    unw, cc = O.irls_unwrap_2d(f["psi"], f["coh"], **KW)
    np.testing.assert_allclose(unw, f["unwrapped"], atol=1e-9, equal_nan=True)
    np.testing.assert_array_equal(cc, f["conncomp"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_matches_fixture(hip, path):
    from core_amd import irls_unwrap_2d
    f = _load(path)
    print(f"fixture {os.path.basename(path)}", flush=True)
    unw, cc = irls_unwrap_2d(f["psi"], f["coh"], device="cuda", **KW)
    v = ~np.isnan(f["psi"])
    assert np.array_equal(np.isnan(unw), ~v)
    rms = float(np.sqrt(np.mean((unw[v].astype(np.float64) - f["unwrapped"][v]) ** 2)))
    assert rms <= 1e-3, rms
    np.testing.assert_array_equal(cc, f["conncomp"])

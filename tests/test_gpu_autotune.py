"""GPU tests of the plan autotuner: every candidate plan (radix order, Bluestein length, elements
per thread) computes the same preconditioner as scipy in FP64, and a tuned, saved and re-applied
plan set keeps the solver's parity with the FP64 oracle."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)


def _lib_describe(n):
    from core_amd import _lib
    return _lib.plan_describe(n)


def _precond64(r):
    import scipy.fft as sf
    h, w = r.shape
    lam = 2 * np.cos(np.pi * np.arange(h) / h)[:, None] + 2 * np.cos(np.pi * np.arange(w) / w)[None, :] - 4
    lam[0, 0] = 1.0
    X = sf.dctn(r.astype(np.float64), type=2) / -lam
    X[0, 0] = 0.0
    return sf.idctn(X, type=2)


@pytest.mark.parametrize("n", [48, 211, 1500, 2048, 4999])
def test_every_candidate_plan_is_exact(hip, n):
    from core_amd import autotune as A, _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    shapes = [(24, n)]
    rng = np.random.default_rng(n)
    r = rng.normal(size=shapes[0]).astype(np.float32)
    z64 = _precond64(r)
    print(f"candidates for length {n}", flush=True)
    try:
        for fl, rad, ept in A.candidates(n):
            _lib.plan_override(n, fl, rad, ept)
            ws = Workspace(shapes, hip)
            z = precondition_device(pack([r], ws, hip), shapes, workspace=ws)
            zg = image_view(z.cpu().numpy(), ws, 0).astype(np.float64)
            ws.release()
            err = np.abs(zg - z64).max() / np.abs(z64).max()
            assert err < 2e-5, (n, fl, rad, ept, err)
    finally:
        A.reset()


def test_tuned_plans_keep_parity(hip, tmp_path):
    from core_amd import autotune as A, irls_unwrap_2d
    from oracle import irls_fp64
    shapes = [(96, 130), (127, 211)]
    path = str(tmp_path / "plans.json")
    try:
        res = A.tune(shapes, reps=2, path=path)
        print(f"tuned {res['before_ms']} -> {res['after_ms']} ms: {res['plans']}", flush=True)
        assert set(res["plans"]) <= {96, 130, 127, 211}
        assert res["after_ms"] <= res["before_ms"] * 1.05
        # a saved set of non-default plans, re-applied in a fresh state
        A._merge(path, A._device_key(hip), {127: {"fft_len": 256, "radices": [4, 4, 4, 4], "ept": 24},
                                             130: {"fft_len": 300, "radices": [3, 5, 5, 4], "ept": 16}})
        A.reset()
        assert A.apply(path) >= 2
        assert _lib_describe(127) == (256, [4, 4, 4, 4], 24)
        for (h, w), seed in zip(shapes, (1, 2)):
            wl = S.Workload(h, w, seed, ramp_max=40.0, n_bumps=6, fault=seed == 1, mask_frac=0.05 if seed == 1 else 0)
            psi, coh, _ = S.generate(wl)
            u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
            u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
            v = ~np.isnan(psi)
            rms = float(np.sqrt(np.mean((u[v].astype(np.float64) - u64[v]) ** 2)))
            assert rms <= 1e-3, (h, w, rms)
            assert np.array_equal(cc, cc64)
    finally:
        A.reset()


def test_committed_plans_are_exact(hip):
    """Every plan in core_amd/tuned_plans.json for this GPU (what bench.py applies): the
    preconditioner of a 24-row image of that width against scipy FP64."""
    import json
    from core_amd import autotune as A, _lib
    from core_amd.unwrap2d import Workspace, image_view, pack, precondition_device
    with open(A.DEFAULT_FILE) as f:
        plans = json.load(f).get(A._device_key(hip), {})
    print(f"{len(plans)} committed plans", flush=True)
    try:
        assert A.apply() == len(plans)
        rng = np.random.default_rng(5)
        for n, p in sorted(plans.items(), key=lambda kv: int(kv[0])):
            n = int(n)
            assert _lib_describe(n) == (p["fft_len"], p["radices"], p["ept"])
            shapes = [(24, n)]
            r = rng.normal(size=shapes[0]).astype(np.float32)
            ws = Workspace(shapes, hip)
            z = precondition_device(pack([r], ws, hip), shapes, workspace=ws)
            zg = image_view(z.cpu().numpy(), ws, 0).astype(np.float64)
            ws.release()
            z64 = _precond64(r)
            err = np.abs(zg - z64).max() / np.abs(z64).max()
            assert err < 2e-5, (n, p, err)
    finally:
        A.reset()

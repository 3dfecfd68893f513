"""Per-workload GPU parity: each benchmark workload's own generator through the drop-in call
(`irls_unwrap_2d` / `irls_unwrap_2d_batch`: crop, H2D, solve, device epilogue, D2H), fixed
10 IRLS x 50 PCG iterations.

Oracles:
  * the CPU workload (1024 x 1024) in full against oracle.irls_fp64 (numpy FP64) and
    oracle.irls_torch32 (the reference's float32 torch operations, on the CPU);
  * larger grids against the same torch restatement run on the HIP device: in float64 it is the
    FP64 trajectory (equal to the numpy oracle to 1e-9 relative,
    tests/test_oracle_pins.py::test_torch_restatement_in_float64_equals_fp64_oracle), in float32
    it is the reference's own float32 operations on a GPU device (the reference runs exactly these
    torch calls for device='cuda').  Both use the host epilogue of oracle.post.

Bars, per case, valid pixels only:
  * RMS(phi_gpu - phi_fp64) <= 1e-3 rad;
  * cycle count k = round((phi - psi) / 2 pi) agrees with the float32 reference path on
    >= 99.99% of pixels;
  * conncomp equal to the FP64 oracle's, NaN mask equal to psi's;
  * rewrap: the congruent output psi + 2 pi k rewraps onto psi within 1e-5 rad (in FP64).
"""
import numpy as np
import pytest
import torch

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)
RMS_TOL = 1e-3
K_AGREE = 0.9999
REWRAP_TOL = 1e-5


def _metrics(u, u64, u32, psi):
    v = ~np.isnan(psi)
    d = u[v].astype(np.float64) - u64[v].astype(np.float64)
    rms = float(np.sqrt(np.mean(d * d)))
    k = np.round((u[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    k32 = np.round((u32[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    agree = float(np.mean(k == k32))
    cong = psi[v].astype(np.float64) + 2 * np.pi * k
    rewrap = float(np.abs(np.angle(np.exp(1j * (cong - psi[v])))).max())
    return rms, agree, rewrap


def _check(name, psi, u, cc, u64, cc64, u32):
    assert u.dtype == np.float32 and cc.dtype == np.uint16
    assert np.array_equal(np.isnan(u), np.isnan(psi)), name
    rms, agree, rewrap = _metrics(u, u64, u32, psi)
    print(f"  {name}: rms_vs_fp64={rms:.3e} k_agree_vs_fp32={agree:.6f} rewrap={rewrap:.1e} "
          f"components={int(cc.max())}", flush=True)
    assert rms <= RMS_TOL, name
    assert agree >= K_AGREE, name
    assert rewrap <= REWRAP_TOL, name
    assert np.array_equal(cc, cc64), name


def _oracles_on_device(psi, coh, dev):
    """(u, conncomp) of the torch restatement in float64 and float32 on the HIP device."""
    from oracle import irls_torch32
    u64, cc64 = irls_torch32.irls_unwrap_2d(psi, coh, dtype=torch.float64, device=dev, **FIXED)
    torch.cuda.empty_cache()
    u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, dtype=torch.float32, device=dev, **FIXED)
    torch.cuda.empty_cache()
    return u64, cc64, u32


@pytest.mark.timeout(600)
def test_cpu_workload_1024_in_full(hip):
    """The 1024 x 1024 CPU workload (ramp 0-40 rad, 6 bumps, coherence U[0.3, 1], seed 0) against
    both CPU oracles."""
    from core_amd import irls_unwrap_2d
    from oracle import irls_fp64, irls_torch32
    psi, coh, _ = S.generate(S.single_1024())
    print("1024x1024 seed 0", flush=True)
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64 = irls_fp64.irls_unwrap_2d(psi, coh, **FIXED)
    u32, _ = irls_torch32.irls_unwrap_2d(psi, coh, **FIXED)
    _check("1024^2", psi, u, cc, u64, cc64, u32)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,wl", [
    ("8192-workload generator at 2048^2", S.scaled(S.single_8192(), 2048, 2048)),
    ("40000-workload generator at 2048^2", S.scaled(S.single_40000(), 2048, 2048)),
])
def test_large_workload_generators_at_2048(hip, name, wl):
    """The generators of the 8192^2 (ramp scaled, 3 pi fault, 5% masked blocks) and 40000^2
    (20 bumps, fault, masks) workloads at 2048 x 2048."""
    from core_amd import irls_unwrap_2d
    psi, coh, _ = S.generate(wl)
    print(name, flush=True)
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64, u32 = _oracles_on_device(psi, coh, hip)
    _check(name, psi, u, cc, u64, cc64, u32)


@pytest.mark.timeout(900)
def test_plain_8192_workload_in_full(hip):
    """The headline 8192 x 8192 workload (ramp 0-300 rad, 6 bumps, 3 pi fault, 5% masked 16 x 16
    blocks, seed 1) in full: the drop-in call against the FP64 and float32 trajectories, including
    the cycle agreement along the 8192-pixel fault line."""
    from core_amd import irls_unwrap_2d
    psi, coh, _ = S.generate(S.single_8192())
    print("8192x8192 seed 1", flush=True)
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
    u64, cc64, u32 = _oracles_on_device(psi, coh, hip)
    _check("8192^2", psi, u, cc, u64, cc64, u32)


@pytest.mark.timeout(900)
def test_stack_48_batched(hip):
    """The 48-burst stack (H ~ U{1000..1500}, W ~ U{4000..6000}, seeds 1000 + i, ramp 0-80 rad,
    coherence U[0.2, 1]) in one batched call: every image finite with psi's NaN mask and bitwise
    equal to its own plain call (a slab, a tridiagonal and a transposed column pass among them); the
    three smallest bursts against the FP64 and float32 trajectories."""
    from core_amd import irls_unwrap_2d, irls_unwrap_2d_batch
    wls = S.stack_48()
    ims = [S.generate(wl)[:2] for wl in wls]
    print("stack 48 batched", flush=True)
    res = irls_unwrap_2d_batch([a for a, _ in ims], [b for _, b in ims], device='cuda', **FIXED)
    for i, ((psi, coh), (u, cc)) in enumerate(zip(ims, res)):
        assert np.array_equal(np.isnan(u), np.isnan(psi)), i
        assert np.isfinite(u[~np.isnan(psi)]).all(), i
    torch.cuda.empty_cache()
    from core_amd import _lib
    strat = _lib.describe_columns([(wl.H, wl.W) for wl in wls])
    picks = sorted({strat.index(k) for k in ("slab_register", "tridiag", "direct", "transpose") if k in strat} | {0, 47})
    for i in picks:  # plain calls of bursts of every column strategy present: bitwise the same solve
        psi, coh = ims[i]
        up, cp = irls_unwrap_2d(psi, coh, device='cuda', **FIXED)
        print(f"  burst {i} ({strat[i]}) batched vs plain: max |diff| = "
              f"{np.nanmax(np.abs(res[i][0] - up)):.1e}", flush=True)
        assert np.array_equal(res[i][0], up, equal_nan=True) and np.array_equal(res[i][1], cp), i
    order = np.argsort([wl.H * wl.W for wl in wls])[:3]
    for i in order:
        psi, coh = ims[i]
        u64, cc64, u32 = _oracles_on_device(psi, coh, hip)
        _check(f"burst {i} {wls[i].H}x{wls[i].W}", psi, res[i][0], res[i][1], u64, cc64, u32)

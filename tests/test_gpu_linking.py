"""GPU test of the default unwrap2d path (conncomp=False): the drop-in solve on the device, then
component linking (core_amd.linking) against the brute-force linking restatement
(oracle/linking.py) applied to the same solver output, on grids split into several components by
NaN bands.  (The solve itself is pinned against the oracles by the parity tests; linking an
FP64 solve instead would compare two different inputs: where a component's circular-mean
alignment sits near +-pi, the reference's own float32 and FP64 paths differ there by a whole
cycle.)  The linked grid differs from the solver output by a constant multiple of 2 pi on each
component and equals it elsewhere."""
import numpy as np
import pytest

from core_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)


# NaN bands (rows r0:r1, cols c0:c1) cutting each grid into 3-4 components, plus holes
LAYOUTS = {
    "A": (90, 120, 200, [(None, None, 38, 42), (None, None, 80, 83), (50, 53, 83, None), (5, 9, 5, 9),
                          (85, None, 0, 6), (86, None, 6, 8)]),
    "B": (96, 130, 201, [(40, 43, None, None), (None, 40, 60, 63), (70, 74, 10, 14)]),
    "C": (256, 300, 203, [(None, None, 100, 103), (120, 124, 103, None), (124, None, 200, 204)]),
}


def _grid(name):
    H, W, seed, bands = LAYOUTS[name]
    psi, coh, _ = S.generate(S.Workload(H, W, seed, ramp_max=40.0, n_bumps=4))
    for r0, r1, c0, c1 in bands:
        psi[r0:r1, c0:c1] = np.nan
    return psi, coh


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_unwrap_then_link_matches_oracle(hip, name):
    from core_amd import irls_unwrap_2d, link_components
    from oracle import linking
    psi, coh = _grid(name)
    print(f"linking layout {name} {psi.shape}", flush=True)
    u, cc = irls_unwrap_2d(psi, coh, device='cuda', conncomp_size=20, **FIXED)
    ncomp = int(cc.max())
    a = link_components(u, cc, conncomp_size=20)
    info = {}
    b = linking.link(u, cc, conncomp_size=20, info=info)
    print(f"  components={ncomp} optima={info.get('n_opt')} k={info.get('k')}", flush=True)
    assert 2 <= ncomp <= 6   # the oracle enumerates the offsets
    assert np.array_equal(np.isnan(a), np.isnan(b))
    if info.get('n_opt', 1) == 1:
        assert np.array_equal(a, b, equal_nan=True)
    # per component a constant 2 pi multiple; label-0 pixels unchanged
    for lab in range(1, ncomp + 1):
        d = (a[cc == lab].astype(np.float64) - u[cc == lab]) / (2 * np.pi)
        assert np.allclose(d, np.round(d[0]), atol=1e-3), lab
    z = (cc == 0) & ~np.isnan(u)
    assert np.array_equal(a[z], u[z])

"""Writes the small golden fixtures of tests/golden/*.npz from the FP64 oracle.

The reference holds no fixtures for this path, so these are produced by the FP64 restatement
(oracle/irls_fp64.py, itself pinned by tests/test_oracle_pins.py) on seeded synthetic inputs:
inputs, the intermediate quantities of the reference's algorithm (targets, edge weights, first
preconditioner application, IRLS weights and right-hand side of the first IRLS step) and the final
outputs at fixed 10 IRLS x 50 PCG iterations.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from core_amd import synthetic as S  # noqa: E402
from oracle import irls_fp64 as O, post  # noqa: E402

CASES = {
    "smooth_48x64": (S.Workload(48, 64, 0, ramp_max=40.0), None),
    "fault_mask_40x50": (S.Workload(40, 50, 6, ramp_max=30.0, fault=True, mask_frac=0.1), "border"),
    "prime_31x37": (S.Workload(31, 37, 7, ramp_max=20.0), None),
}
KW = dict(max_iter=10, tol=0.0, cg_max_iter=50, cg_tol=0.0)


def build(name):
    wl, extra = CASES[name]
    psi, coh, _ = S.generate(wl)
    if extra == "border":
        psi[:3] = np.nan
        psi[:, -4:] = np.nan
    crop, ph_c, nm, pf, wf = post.prologue(psi, coh)
    valid = ~nm
    tx, ty = O.targets(pf)
    wx, wy = O.edge_weights(wf, valid)
    lam = O.eigenvalues(*pf.shape)
    rho = O.divergence_rhs(tx, ty, wx, wy)
    u0 = O.precondition(rho, lam)
    # first IRLS step quantities at the centred DCT start
    u = u0 - (u0 * valid).sum() / valid.sum()
    rx = np.zeros_like(u); ry = np.zeros_like(u)
    rx[:, :-1] = u[:, 1:] - u[:, :-1]
    ry[:-1, :] = u[1:, :] - u[:-1, :]
    rx -= tx; ry -= ty
    wxi = np.clip(wx / np.sqrt(rx * rx + 1e-4), 1e-6, 1e6)
    wyi = np.clip(wy / np.sqrt(ry * ry + 1e-4), 1e-6, 1e6)
    b1 = O.divergence_rhs(tx, ty, wxi, wyi)
    Au = O.laplacian(u, wxi, wyi)
    unw, cc, raw = O.irls_unwrap_2d(psi, coh, return_raw=True, **KW)
    return dict(psi=psi, coh=coh, crop=np.array(crop), tx=tx, ty=ty, wx=wx, wy=wy, rho=rho, u0=u0,
                wxi1=wxi, wyi1=wyi, b1=b1, Au1=Au, raw=raw, unwrapped=unw, conncomp=cc)


def main():
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **build(name))
        print("wrote", name)


if __name__ == "__main__":
    main()

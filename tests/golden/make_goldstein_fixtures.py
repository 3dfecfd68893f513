"""Writes tests/golden/goldstein_*.npz: inputs and outputs of the reference's own Goldstein filters.

The closed forms in tests/test_goldstein_cpu.py pin the per-patch arithmetic of the oracle
(oracle/goldstein.py): the window, alpha, the spectral weighting, the overlap-add.  What they do
not fix is how each of the reference's two implementations sets up its patch grid and combines
overlapping patches (numpy: weight-normalised, floor grid; torch: count-normalised, ceil grid) —
that is pinned here by the reference functions themselves, insardev/insardev/utils_goldstein.py
`goldstein_numpy` (:12-154) and `goldstein_pytorch` (:157-330, on the CPU device), loaded from
that single file (it needs only numpy, scipy and torch; the package itself is not imported).

    python tests/golden/make_goldstein_fixtures.py
"""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/insardev/insardev/utils_goldstein.py"

# name: (ny, nx, psize_y, psize_x, threshold, seed)
CASES = {
    "goldstein_70x90_p32": (70, 90, 32, 32, 0.5, 1),
    "goldstein_45x37_p16_t025": (45, 37, 16, 16, 0.25, 2),
    "goldstein_64x48_p16x32": (64, 48, 16, 32, 0.5, 3),
}


def inputs(ny, nx, seed):
    """Complex interferogram-like field (ramp + noise, varying amplitude) with NaN blocks and NaN
    correlation pixels."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx]
    psi = 0.21 * x + 0.13 * y + 0.002 * x * y + rng.normal(0, 0.6, (ny, nx))
    amp = 0.5 + rng.uniform(0, 1, (ny, nx))
    phase = (amp * np.exp(1j * psi)).astype(np.complex64)
    corr = rng.uniform(0.05, 0.95, (ny, nx)).astype(np.float32)
    phase[5:12, 7:20] = np.nan
    phase[ny - 6:, nx - 9:] = np.nan
    corr[20:23, 3:9] = np.nan
    return phase, corr


def main():
    spec = importlib.util.spec_from_file_location("ref_goldstein", REF)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    import torch
    for name, (ny, nx, py, px, thr, seed) in CASES.items():
        phase, corr = inputs(ny, nx, seed)
        out_np = ref.goldstein_numpy(phase.copy(), corr.copy(), py, px, threshold=thr)
        out_pt = ref.goldstein_pytorch(phase.copy(), corr.copy(), py, px, torch.device("cpu"), threshold=thr)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), phase=phase, corr=corr,
                            psize=np.array([py, px]), threshold=np.array(thr),
                            out_numpy=out_np, out_torch=out_pt)
        print(name, out_np.shape, int(np.isnan(out_np).sum()), int(np.isnan(out_pt).sum()))


if __name__ == "__main__":
    main()

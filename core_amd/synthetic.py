"""Seeded synthetic interferograms for the unwrap workloads.

One generator serves every workload so that tests, the benchmark and the
fixtures all see the same bytes for the same (shape, seed, options):

    rng  = numpy.random.default_rng(seed)            (PCG64)
    true = ramp_max * x / (W - 1)
    for each bump: A ~ U[-15, 15], s ~ U[40, 150], cy ~ U[0, H), cx ~ U[0, W)
        true += A * exp(-((y - cy)^2 + (x - cx)^2) / (2 s^2))
    if fault: theta ~ U[0, pi), fy ~ U[0, H), fx ~ U[0, W)
        true += 3 pi * [(x - fx) sin(theta) - (y - fy) cos(theta) > 0]
    psi  = wrap(true + N(0, noise)) into (-pi, pi]
    coh  ~ U[c_lo, c_hi];  py ~ U{0..H-128}, px ~ U{0..W-128}
    coh[py:py+128, px:px+128] ~ U[0.05, 0.15]
    if mask_frac: round(mask_frac * (H//16) * (W//16)) distinct 16x16 blocks on
        the 16-aligned grid get psi = NaN and coh = 0.

All arithmetic is float64 and the results are cast to float32.  Rows are
produced in bands so that 40000 x 40000 grids never hold a full float64 field;
numpy's generators consume their stream sequentially, so banded draws give the
same values as one whole-field draw.

The reference's own benchmark script builds its inputs in the same spirit
(tests/test_unwrap_benchmark/test_unwrap_benchmark.py:236-303: ramp + bumps +
noise + NaN gaps, wrapped), but it pins no values, so these definitions are the
contract for this package's workloads.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from concurrent.futures import ThreadPoolExecutor

import numpy as np

TWO_PI = 2.0 * np.pi


@dataclass(frozen=True)
class Workload:
    """Parameters of one synthetic interferogram."""
    H: int
    W: int
    seed: int
    ramp_max: float = 40.0
    n_bumps: int = 6
    fault: bool = False
    mask_frac: float = 0.0
    coh_lo: float = 0.3
    coh_hi: float = 1.0
    noise: float = 0.3


def wrap_f64(x: np.ndarray) -> np.ndarray:
    """Wrap float64 angles into (-pi, pi]."""
    w = np.mod(x + np.pi, TWO_PI) - np.pi
    # mod maps to [-pi, pi); move the -pi end to +pi
    w[w <= -np.pi] += TWO_PI
    return w


def generate(wl: Workload, band_rows: int = 2048, workers: int | None = None, progress=None):
    """Return (psi float32[H,W] with NaN holes, coh float32[H,W], true float32[H,W]).

    `progress(rows_done, H)` is called after each band of the phase field when given."""
    H, W = int(wl.H), int(wl.W)
    rng = np.random.default_rng(wl.seed)
    bumps = []
    for _ in range(wl.n_bumps):
        A = rng.uniform(-15.0, 15.0)
        s = rng.uniform(40.0, 150.0)
        cy = rng.uniform(0.0, H)
        cx = rng.uniform(0.0, W)
        bumps.append((A, s, cy, cx))
    fault = None
    if wl.fault:
        theta = rng.uniform(0.0, np.pi)
        fy = rng.uniform(0.0, H)
        fx = rng.uniform(0.0, W)
        fault = (theta, fy, fx)

    psi = np.empty((H, W), np.float32)
    true32 = np.empty((H, W), np.float32)
    x = np.arange(W, dtype=np.float64)[None, :]
    ramp = wl.ramp_max * x / max(W - 1, 1)

    def deterministic(r0, r1):
        y = np.arange(r0, r1, dtype=np.float64)[:, None]
        t = np.broadcast_to(ramp, (r1 - r0, W)).copy()
        for (A, s, cy, cx) in bumps:
            # exp(-d^2 / 2s^2) is exactly 0.0 in float64 beyond d = 38.6 s (argument < -745), so
            # evaluating it only inside that window leaves every value bit-identical
            R = 39.0 * s + 1.0
            i0, i1 = max(r0, int(np.floor(cy - R))), min(r1, int(np.ceil(cy + R)) + 1)
            j0, j1 = max(0, int(np.floor(cx - R))), min(W, int(np.ceil(cx + R)) + 1)
            if i0 >= i1 or j0 >= j1:
                continue
            yy = y[i0 - r0:i1 - r0]
            xx = x[:, j0:j1]
            t[i0 - r0:i1 - r0, j0:j1] += A * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2.0 * s * s))
        if fault is not None:
            theta, fy, fx = fault
            t += 3.0 * np.pi * (((x - fx) * np.sin(theta) - (y - fy) * np.cos(theta)) > 0)
        true32[r0:r1] = t
        return t

    def finish(r0, r1, t, noise):
        t += noise
        psi[r0:r1] = wrap_f64(t)

    # The random draws stay in stream order on this thread; the deterministic field and the
    # wrap run on worker threads (numpy releases the GIL), a bounded number of bands ahead.
    bands = [(r0, min(H, r0 + band_rows)) for r0 in range(0, H, band_rows)]
    nw = max(1, min(workers or 8, len(bands)))
    with ThreadPoolExecutor(nw) as ex:
        det = {}
        fin = []
        for k, (r0, r1) in enumerate(bands):
            for kk in range(k, min(len(bands), k + nw)):
                if kk not in det:
                    det[kk] = ex.submit(deterministic, *bands[kk])
            t = det.pop(k).result()
            noise = rng.normal(0.0, wl.noise, size=(r1 - r0, W))
            fin.append(ex.submit(finish, r0, r1, t, noise))
            while len(fin) > nw:
                fin.pop(0).result()
            if progress is not None:
                progress(r1, H)
        for f in fin:
            f.result()

    coh = np.empty((H, W), np.float32)
    for r0 in range(0, H, band_rows):
        r1 = min(H, r0 + band_rows)
        coh[r0:r1] = rng.uniform(wl.coh_lo, wl.coh_hi, size=(r1 - r0, W))
    ph = min(128, H)
    pw = min(128, W)
    py = int(rng.integers(0, H - ph + 1))
    px = int(rng.integers(0, W - pw + 1))
    coh[py:py + ph, px:px + pw] = rng.uniform(0.05, 0.15, size=(ph, pw))

    if wl.mask_frac > 0:
        by, bx = H // 16, W // 16
        nblk = by * bx
        k = int(round(wl.mask_frac * nblk))
        if k > 0:
            sel = rng.choice(nblk, size=k, replace=False)
            for b in sel:
                i0 = (int(b) // bx) * 16
                j0 = (int(b) % bx) * 16
                psi[i0:i0 + 16, j0:j0 + 16] = np.nan
                coh[i0:i0 + 16, j0:j0 + 16] = 0.0
    return psi, coh, true32


# ---------------------------------------------------------------------------
# Named workloads.  Sizes, seeds and options are the benchmark definitions.
# ---------------------------------------------------------------------------

def single_1024(seed: int = 0) -> Workload:
    """1024 x 1024, ramp 0-40 rad, 6 bumps, coherence U[0.3, 1]; CPU-sized case."""
    return Workload(1024, 1024, seed, ramp_max=40.0, n_bumps=6)


def single_8192(seed: int = 1) -> Workload:
    """8192 x 8192, ramp 0-300 rad, 6 bumps, 3 pi fault line, 5% masked 16x16 blocks."""
    return Workload(8192, 8192, seed, ramp_max=300.0, n_bumps=6, fault=True, mask_frac=0.05)


def stack_shapes(n: int = 48, seed: int = 2):
    """Uneven burst shapes H ~ U{1000..1500}, W ~ U{4000..6000}."""
    rng = np.random.default_rng(seed)
    Hs = rng.integers(1000, 1501, size=n)
    Ws = rng.integers(4000, 6001, size=n)
    return [(int(h), int(w)) for h, w in zip(Hs, Ws)]


def stack_48(n: int = 48):
    """48 interferograms of uneven shapes, image i uses seed 1000 + i, ramp 0-80 rad, coherence U[0.2, 1]."""
    return [Workload(h, w, 1000 + i, ramp_max=80.0, n_bumps=6, coh_lo=0.2, coh_hi=1.0)
            for i, (h, w) in enumerate(stack_shapes(n))]


def adjoint_batch(n: int = 16, size: int = 2048):
    """16 x 2048 x 2048 with the 1024-case generator, seeds 3 + i."""
    return [Workload(size, size, 3 + i, ramp_max=40.0, n_bumps=6) for i in range(n)]


def single_40000(seed: int = 4) -> Workload:
    """40000 x 40000, ramp 0-1500 rad, 20 bumps, fault line, 5% masked blocks."""
    return Workload(40000, 40000, seed, ramp_max=1500.0, n_bumps=20, fault=True, mask_frac=0.05)


def scaled(wl: Workload, H: int, W: int) -> Workload:
    """The same generator options at another grid size (for parity at oracle-friendly sizes)."""
    return Workload(H, W, wl.seed, wl.ramp_max * W / max(wl.W, 1), wl.n_bumps, wl.fault,
                    wl.mask_frac, wl.coh_lo, wl.coh_hi, wl.noise)

"""Autotuner of the DCT transform plans (radix schedule, Bluestein length, elements per thread).

For every LDS-resident DCT length of a set of image shapes it times candidate plans with the
library's preconditioner entry (irls2d_precondition: one row DCT-II, the column DCT-II/-III pair
and one row DCT-III, exactly the transforms of a PCG iteration) on a probe image whose rows have
that length, keeps the fastest, and then settles the elements per thread (shared by all images
of a launch) on the whole batch.  Results are kept
per GPU name in a JSON file and can be re-applied at start-up with `apply()`:

    from core_amd import autotune
    plans = autotune.tune(shapes)            # measure, apply, save
    autotune.apply()                         # later processes: re-apply the saved plans

Candidates per length N (plan = fft_len, radix order, elements per thread):
  * direct FFT (N 5-smooth): the power-of-two part as radix-16, radix-8 or radix-4 passes, with
    the radix-3/5 passes after or before it;
  * Bluestein (prime factors > 5): the three 5-smooth lengths M >= 2N - 1 of lowest
    M * (passes + 1) and the lengths with compile-time schedules in that range, each with its
    default order;
  * 12 (on up to 1024 threads), 16 or 24 elements per thread where the FFT length allows them.
"""
from __future__ import annotations

import json
import os
import time

import torch

from . import _lib

DEFAULT_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_plans.json")
MAX_FFT_LEN = 12288


def _factor235(m: int):
    e = {}
    for p in (2, 3, 5):
        while m % p == 0:
            e[p] = e.get(p, 0) + 1
            m //= p
    return e, m


def _pow2_split(a: int, big: int):
    """Radices for 2^a using passes of `big` (16, 8 or 4) first, then the remainder."""
    lg = {16: 4, 8: 3, 4: 2}[big]
    out = [big] * (a // lg)
    rest = a % lg
    if rest:
        out.append(1 << rest)
    return out


def _smooth(m: int) -> bool:
    return _factor235(m)[1] == 1


def _default_radices(m: int):
    out = []
    for r in (16, 8, 4, 2, 3, 5):
        while m % r == 0:
            out.append(r)
            m //= r
    return out


def _has_ct(m: int) -> bool:
    """Lengths with a compile-time pass schedule in the library: 2^a * {1, 3, 5}, 1024 .. 12288."""
    if m < 1024 or m > MAX_FFT_LEN:
        return False
    while m % 2 == 0:
        m //= 2
    return m in (1, 3, 5)


def bluestein_candidates(n: int, k: int = 3):
    """The k 5-smooth lengths M >= 2n - 1 of lowest M * (passes + 1) (within 30%), plus every
    length in that range with a compile-time schedule."""
    lo = 2 * n - 1
    hi = min(MAX_FFT_LEN, lo + lo * 3 // 10 + 8)
    cands = [(m * (len(_default_radices(m)) + 1), m) for m in range(lo, hi + 1) if _smooth(m)]
    best = [m for _, m in sorted(cands)[:k]]
    return best + [m for m in range(lo, hi + 1) if _has_ct(m) and m not in best]


def candidates(n: int):
    """Candidate plans (fft_len, radices, ept) for DCT length n."""
    out = []
    if _smooth(n):
        lens = [n]
    else:
        lens = bluestein_candidates(n)
    for m in lens:
        e, _ = _factor235(m)
        odd = [3] * e.get(3, 0) + [5] * e.get(5, 0)
        scheds = []
        for big in (16, 8, 4):
            p2 = _pow2_split(e.get(2, 0), big)
            scheds.append(p2 + odd)
            if odd and p2:
                scheds.append(odd + p2)
            if m != n:  # Bluestein: one order per length is enough
                break
        epts = [16, 12, 24] if m <= 8192 else [12, 24]
        seen = set()
        for sch in scheds:
            if len(sch) > 16:
                continue
            for ept in epts:
                key = (m, tuple(sch), ept)
                if key not in seen:
                    seen.add(key)
                    out.append((m, list(sch), ept))
    return out


class _Bench:
    """Workspace and vectors of one shape set; the workspace grows when a plan needs more tables."""

    def __init__(self, shapes, dev):
        from .unwrap2d import Workspace
        self.shapes, self.dev = shapes, dev
        self.ws = Workspace(shapes, dev)
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        self.r = torch.randn(self.ws.total, device=dev, generator=g)
        self.z = torch.empty_like(self.r)

    def time(self, reps):
        from .unwrap2d import Workspace
        if _lib.workspace_size(self.shapes) > self.ws.nbytes:
            self.ws.release()
            self.ws = Workspace(self.shapes, self.dev)
        return _time_precondition(self.shapes, self.r, self.z, self.ws, reps)

    def close(self):
        self.ws.release()


def _time_precondition(shapes, r, z, ws, reps):
    L = _lib.load()
    n, H, W = _lib.shapes_arrays(shapes)
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        _lib.check(L.irls2d_precondition(n, H, W, r.data_ptr(), z.data_ptr(), ws.buf.data_ptr(), ws.nbytes,
                                         stream), "irls2d_precondition")

    run()  # uploads the tables of the current plans
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(3):
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / reps)
    return best


def _device_key(device):
    p = torch.cuda.get_device_properties(device)
    return getattr(p, "gcnArchName", None) or p.name


def _probe_shapes(n: int):
    """Probe batch whose row transforms (length n) dominate: 64-row images, about 4 Mpx."""
    k = max(1, 4_000_000 // (64 * n))
    return [(64, n)] * k


def tune(shapes, device="cuda", reps=5, path=DEFAULT_FILE, save=True, verbose=False):
    """Measure candidate plans for the LDS-resident lengths of `shapes`, apply the fastest and
    (optionally) merge them into the JSON file under this GPU's name.

    Each length is tuned on a probe image whose rows have that length (its row transforms are
    what changes between candidates); the elements per thread, which one launch shares across
    all images, are then settled on the whole batch.  Returns {"plans": {length: {"fft_len",
    "radices", "ept", "ms"}}, "before_ms", "after_ms"} with the batch's preconditioner times."""
    dev = torch.device(device) if not isinstance(device, torch.device) else device
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    shapes = [(int(h), int(w)) for h, w in shapes]
    weight = {}
    for h, w in shapes:
        for n in (h, w):
            if 1 < n <= MAX_FFT_LEN:
                weight[n] = weight.get(n, 0) + h * w
    lengths = sorted(weight, key=lambda n: -weight[n])
    t0 = time.time()
    with torch.cuda.device(dev):
        batch = _Bench(shapes, dev)
        base = batch.time(reps)
        chosen = {}
        for i, n in enumerate(lengths):
            probe = _Bench(_probe_shapes(n), dev)
            default = _lib.plan_describe(n)
            _lib.plan_override(n, 0)
            best = default
            best_ms = def_ms = probe.time(reps)
            for fl, rad, ept in candidates(n):
                if (fl, rad, ept) == tuple(default):
                    continue
                _lib.plan_override(n, fl, rad, ept)
                ms = probe.time(reps)
                if verbose:
                    print(f"  N={n} fft_len={fl} radices={rad} ept={ept}: {ms:.4f} ms (default {def_ms:.4f})",
                          flush=True)
                if ms < best_ms and ms < 0.98 * def_ms:  # only a clear gain replaces the default
                    best_ms, best = ms, (fl, rad, ept)
            probe.close()
            _lib.plan_override(n, *best)
            chosen[n] = {"fft_len": best[0], "radices": best[1], "ept": best[2], "ms": round(best_ms or 0.0, 4)}
            if verbose or (i + 1) % 10 == 0 or i + 1 == len(lengths):
                print(f"autotune: {i + 1}/{len(lengths)} lengths ({time.time() - t0:.0f} s)", flush=True)
        # one elements-per-thread value per launch: keep the per-length choices, or force one
        variants = {"per_length": {n: p["ept"] for n, p in chosen.items()},
                    "all_16": {n: (16 if p["fft_len"] <= 8192 else 12) for n, p in chosen.items()},
                    "all_12": {n: 12 for n in chosen},
                    "all_24": {n: 24 for n in chosen}}
        best_v, best_ms = None, None
        for name, epts in variants.items():
            for n, p in chosen.items():
                _lib.plan_override(n, p["fft_len"], p["radices"], epts[n])
            ms = batch.time(reps)
            if verbose:
                print(f"  batch ept {name}: {ms:.4f} ms", flush=True)
            if best_ms is None or ms < best_ms:
                best_ms, best_v = ms, name
        for n, p in chosen.items():
            p["ept"] = variants[best_v][n]
            _lib.plan_override(n, p["fft_len"], p["radices"], p["ept"])
        # column strategy (transpose + column kernel, column kernel on the row-major field, slab
        # passes): measured for the record; the library chooses it per image
        column_ms = {}
        for name, mode in (("transpose", _lib.COLUMN_TRANSPOSE), ("direct", _lib.COLUMN_DIRECT),
                           ("slab", _lib.COLUMN_SLAB), ("tridiag", _lib.COLUMN_TRIDIAG), ("auto", _lib.COLUMN_AUTO)):
            _lib.set_column_mode(mode)
            column_ms[name] = round(batch.time(reps), 4)
        after = batch.time(reps)
        if after > 0.99 * base:  # no gain on the real batch: keep the built-in plans
            for n in chosen:
                _lib.plan_override(n, 0)
            chosen = {}
            after = batch.time(reps)
        batch.close()
    result = {"plans": chosen, "before_ms": round(base, 4), "after_ms": round(after, 4), "column_ms": column_ms}
    if save:
        _merge(path, _device_key(dev), chosen)
    return result


def _merge(path, key, chosen):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    entry = data.setdefault(key, {})
    for n, p in chosen.items():
        entry[str(n)] = {"fft_len": p["fft_len"], "radices": p["radices"], "ept": p["ept"]}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(path=DEFAULT_FILE, device=None) -> int:
    """Apply the saved plans of this GPU (by its architecture name); returns how many were set."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        data = json.load(f)
    dev = device or torch.device("cuda", torch.cuda.current_device())
    plans = data.get(_device_key(dev), {})
    for n, p in plans.items():
        _lib.plan_override(int(n), p["fft_len"], p["radices"], p["ept"])
    return len(plans)


def reset():
    """Back to the built-in plans."""
    _lib.plan_reset()


if __name__ == "__main__":  # python -m core_amd.autotune [--out FILE]: tune the benchmark shapes
    import argparse
    from . import synthetic as S
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--out", default=DEFAULT_FILE, help="JSON file the chosen plans are merged into")
    args = ap.parse_args()
    sets = {
        "8192": [(8192, 8192)],
        "stack48": S.stack_shapes(),
        "2048x16": [(2048, 2048)] * 16,
    }
    for name, shp in sets.items():
        t0 = time.time()
        res = tune(shp, reps=args.reps, verbose=args.verbose, path=args.out)
        print(json.dumps({"set": name, "before_ms": res["before_ms"], "after_ms": res["after_ms"],
                          "lengths": len(res["plans"]), "column_ms": res["column_ms"],
                          "s": round(time.time() - t0, 1)}), flush=True)

#!/usr/bin/env python3
"""Benchmark of the MI355X DCT+IRLS L1 unwrap (one JSON line on stdout).

Workload (value): one plain 8192 x 8192 float32 interferogram (ramp 0-300 rad, 6 bumps, 3 pi
fault line, 5% masked 16x16 blocks, seed 1), fixed 10 IRLS x 50 PCG iterations, inputs resident
in HBM.  A step is one complete solve (set-up, DCT start, 10 IRLS steps of 50 PCG iterations),
replayed from a captured hipGraph.  value = pixels / step time, in Mpx/s.

Also reported:
  roofline      the kernel with the largest total time in an event-timed solve: its
                algorithmic bytes per launch / its mean launch duration vs 8 TB/s HBM3E;
                traffic = measured HBM bytes per launch of that kernel (rocprofv3 FETCH_SIZE and
                WRITE_SIZE, copy-calibrated) from the committed profile, when present.
  solve_model   whole-solve algorithmic bytes N [100 + 10 (109 + 50 * 76)] / step time: the
                reference algorithm's field traffic (76 B/px per PCG step with the x update
                and the column pass as separate 24 and 8 B/px passes), an effective rate.
  batched       48 uneven SBAS-burst interferograms (H ~ U{1000..1500}, W ~ U{4000..6000},
                seeds 1000 + i) solved in one batched graph launch vs 48 sequential plain calls
                (--only-batched: this workload alone, with per-kernel durations).
  adjoint       16 x 2048 x 2048 forward solve and implicit adjoint (dL/dphase, dL/dweight).
  filters       the upstream Goldstein (psize 32) and Gaussian (sigma 2) filters on an 8192 x 8192
                complex64 interferogram, with their numpy restatements timed on a 2048 x 2048
                crop (--no-filters).
  max_scale     (--max-scale / --only-max-scale) 40000 x 40000 solve with two-pass transforms,
                per-kernel GB/s, and cycle-correctness against the generator's noiseless phase.
  kernel_roofline  every PCG kernel's mean launch time and fraction of HBM peak, side by side.
  end_to_end    the drop-in irls_unwrap_2d on the 8192 x 8192 host arrays (crop, H2D, solve,
                device epilogue, D2H); `value` times the solver on inputs already in HBM.
  cpu_baseline  the float32 torch-CPU restatement of the reference solver (oracle.irls_torch32)
                on a bounded sample of the same workload, and on the 1024 x 1024 CPU workload in
                full, on this host's cores (--no-cpu skips this leg and `accuracy`).
  accuracy      per workload: RMS against the FP64 trajectory, cycle agreement with the float32
                reference path, conncomp equality (the oracles run in the cpu_baseline leg).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpx/s unwrapped (10 IRLS iters) plain vs batched at ≤1e-3 rad; % HBM peak"
HBM_PEAK_GBS = 8000.0
N_IRLS, N_CG = 10, 50
# Algorithmic bytes per pixel of one launch of each PCG kernel (reads + writes of float32 fields).
# Images whose row transform is a Cooley-Tukey length (8192^2, 40000^2) take the PCG step's
# x += alpha p in the next step's k_cg_pap, which reads p anyway; Bluestein rows keep it in the
# row DCT-II pass (ALG_BPX_STACK).
ALG_BPX = {
    "k_cg_pap": 32,           # z, p_old, wxi, wyi, x -> p_new, Ap, x
    "k_row_dct2_update": 16,  # r, Ap -> r, T
    "k_col": 8,               # Tt -> Tt
    "k_col_direct": 8,        # T -> T, column pairs read and written in place
    "k_row_dct3": 12,         # T, r -> z
    "k_transpose": 8,         # T -> Tt / Tt -> T
    # column pass on slabs of the row-major field: T (4 B/px real) <-> S (one complex per pixel
    # pair, 4 B/px)
    "k_slab1_f": 8,           # T -> S
    "k_slab2_f": 8,           # S -> S
    "k_slab3_f": 8,           # S -> T
    "k_slab1": 8,
    "k_slab2": 8,
    "k_slab3": 8,
}
# The same for the two-pass transforms of sides above 12288 (the 40000 x 40000 workload); the
# scratch line holds one complex value per pixel pair, i.e. 4 B per pixel each way.
ALG_BPX_LONG = {
    "k_long_p1_dct2_update": 16,  # r, Ap -> r, scratch (x updated in k_cg_pap)
    "k_long_p2_dct2_rows": 8,     # scratch -> T
    "k_long_p1_dct2_cols": 8,     # Tt -> scratch
    "k_long_p2_dct2_cols": 8,     # scratch -> Tt
    "k_long_p1_dct3_scale": 8,    # Tt -> scratch
    "k_long_p2_dct3_cols": 8,     # scratch -> Tt
    "k_long_p1_dct3_rows": 8,     # T -> scratch
    "k_long_p2_dct3_rows": 12,    # scratch, r -> z
    "k_cg_pap": 32,
    "k_transpose": 8,
    # column pass of tall images as tridiagonal solves: bands of 1536 rows
    "k_col_trc_band": 4,          # T -> per-band transfers (a few values per column and band)
    "k_col_trc": 8,               # T -> T
}
# the 48-burst stack: most widths are Bluestein lengths, whose images keep the update in the row pass
ALG_BPX_STACK = dict(ALG_BPX, k_cg_pap=24, k_row_dct2_update=28)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_8192.json")


def model_bytes_per_px(n_irls=N_IRLS, n_cg=N_CG):
    return 100 + n_irls * (109 + 76 * n_cg)


def upload(ws, arrays, dev):
    from core_amd.unwrap2d import pack
    return pack(arrays, ws, dev)


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps


def plain_8192(args, dev):
    import torch
    from core_amd import solve_device, synthetic as S, _lib
    from core_amd.unwrap2d import Workspace
    wl = S.single_8192()
    psi, coh, _ = S.generate(wl)
    shapes = [(wl.H, wl.W)]
    ws = Workspace(shapes, dev)
    ph = upload(ws, [psi], dev)
    co = upload(ws, [coh], dev)
    out = torch.empty(ws.total, dtype=torch.float32, device=dev)
    kw = dict(max_iter=N_IRLS, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0, out=out, workspace=ws)
    ms, wall_ms = time_steps(lambda: solve_device(ph, co, shapes, graph=True, **kw), args.steps, args.warmup)
    info = solve_device(ph, co, shapes, graph=True, **kw)[1].cpu().tolist()[0]
    # event-timed eager solve: per-kernel mean durations on the launch stream
    _lib.kernel_stats_reset()
    solve_device(ph, co, shapes, graph=False, time_kernels=True, **kw)
    stats = _lib.kernel_stats()
    ws.release()
    del ws, ph, co, out
    torch.cuda.empty_cache()
    return psi, coh, ms, wall_ms, info, stats


def kernel_fracs(stats, npx, table=ALG_BPX):
    """Every kernel with algorithmic bytes: its mean launch duration and fraction of HBM peak."""
    out = {}
    for k, (tot, n) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        if k in table:
            us = tot / n * 1e3
            gbs = table[k] * npx / (us / 1e6) / 1e9
            out[k] = {"mean_us": round(us, 2), "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def roofline(stats, npx):
    cand = {k: v for k, v in stats.items() if k in ALG_BPX}
    dom = max(cand, key=lambda k: cand[k][0])
    tot_ms, n = cand[dom]
    avg_s = tot_ms / n / 1e3
    alg = ALG_BPX[dom] * npx
    achieved = alg / avg_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if dom in t.get("kernels", {}):
            traffic = t["kernels"][dom]["bytes_per_launch"]
    return dom, {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": dom, "alg_bytes_per_launch": alg, "mean_launch_us": round(avg_s * 1e6, 2),
        "launches": n,
    }


def batched_stack(args, dev, kernels=False):
    import torch
    from core_amd import solve_device, synthetic as S, _lib
    from core_amd.unwrap2d import Workspace
    wls = S.stack_48()
    ims = [S.generate(wl)[:2] for wl in wls]
    shapes = [(wl.H, wl.W) for wl in wls]
    npx = sum(h * w for h, w in shapes)
    n_irls = getattr(args, "stack_irls", None) or N_IRLS
    kw = dict(max_iter=n_irls, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0)
    ws = Workspace(shapes, dev)
    ph = upload(ws, [im[0] for im in ims], dev)
    co = upload(ws, [im[1] for im in ims], dev)
    out = torch.empty(ws.total, dtype=torch.float32, device=dev)
    print(f"batched stack: {npx} px generated and uploaded", file=sys.stderr, flush=True)
    ms_b, _ = time_steps(lambda: solve_device(ph, co, shapes, graph=True, out=out, workspace=ws, **kw),
                         args.batched_steps, 1)
    print(f"batched stack: graph steps done ({ms_b:.1f} ms)", file=sys.stderr, flush=True)
    kstats = None
    if kernels:
        # per-kernel durations of one event-timed eager batched solve
        _lib.kernel_stats_reset()
        solve_device(ph, co, shapes, graph=False, time_kernels=True, out=out, workspace=ws, **kw)
        print("batched stack: event-timed eager solve done", file=sys.stderr, flush=True)
        kstats = {}
        for k, (tot, n) in sorted(_lib.kernel_stats().items(), key=lambda kv: -kv[1][0]):
            e = {"total_ms": round(tot, 1), "launches": n, "mean_us": round(tot / n * 1e3, 1)}
            if k in ALG_BPX_STACK:
                e["gbs"] = round(ALG_BPX_STACK[k] * npx / (tot / n / 1e3) / 1e9, 1)
            kstats[k] = e
    ws.release()
    del ws, ph, co, out
    torch.cuda.empty_cache()
    if getattr(args, "no_sequential", False):
        return {"workload": f"48 SBAS bursts H~U{{1000..1500}} W~U{{4000..6000}}, seeds 1000+i, {n_irls} IRLS x 50 PCG",
                "pixels": npx, "batched_ms": round(ms_b, 2), "batched_mpx_s": round(npx / ms_b / 1e3, 2),
                **({"kernels": kstats} if kstats is not None else {})}
    # 48 sequential plain calls (one image per call, eager launches), inputs already in HBM
    singles = []
    for (psi, coh), sh in zip(ims, shapes):
        w1 = Workspace([sh], dev)
        singles.append((w1, upload(w1, [psi], dev), upload(w1, [coh], dev),
                        torch.empty(w1.total, dtype=torch.float32, device=dev)))

    def seq():
        for w1, p1, c1, o1 in singles:
            solve_device(p1, c1, w1.shapes, out=o1, workspace=w1, **kw)

    ms_s, _ = time_steps(seq, 1, 1)
    for w1, *_ in singles:
        w1.release()
    singles.clear()
    torch.cuda.empty_cache()
    res = {
        "workload": "48 SBAS bursts H~U{1000..1500} W~U{4000..6000}, seeds 1000+i, 10 IRLS x 50 PCG",
        "pixels": npx,
        "batched_ms": round(ms_b, 2), "batched_mpx_s": round(npx / ms_b / 1e3, 2),
        "sequential_plain_ms": round(ms_s, 2), "sequential_plain_mpx_s": round(npx / ms_s / 1e3, 2),
        "speedup": round(ms_s / ms_b, 3),
        "solve_model_frac_hbm": round(npx * model_bytes_per_px() / (ms_b / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    if kstats is not None:
        res["kernels"] = kstats
    return res


def adjoint_batch(args, dev):
    """16 x 2048 x 2048 differentiable unwrap: forward solve and implicit adjoint (50 PCG)."""
    import ctypes
    import numpy as np
    import torch
    from core_amd import solve_device, synthetic as S, _lib
    from core_amd.unwrap2d import Workspace
    wls = S.adjoint_batch()
    ims = [S.generate(wl)[:2] for wl in wls]
    shapes = [(wl.H, wl.W) for wl in wls]
    npx = sum(h * w for h, w in shapes)
    ws = Workspace(shapes, dev)
    ph = upload(ws, [im[0] for im in ims], dev)
    co = upload(ws, [im[1] for im in ims], dev)
    out = torch.empty(ws.total, dtype=torch.float32, device=dev)
    kw = dict(max_iter=N_IRLS, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0, out=out, workspace=ws)
    ms_f, _ = time_steps(lambda: solve_device(ph, co, shapes, graph=True, **kw), 1, 1)
    g = torch.randn(ws.total, device=dev)
    gph = torch.empty_like(g)
    gw = torch.empty_like(g)
    n, H, W = _lib.shapes_arrays(shapes)
    L = _lib.load()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def adj():
        _lib.check(L.irls2d_adjoint(n, H, W, ph.data_ptr(), co.data_ptr(), out.data_ptr(), g.data_ptr(),
                                    gph.data_ptr(), gw.data_ptr(), 1e-2, N_CG, 0.0, ws.buf.data_ptr(), ws.nbytes,
                                    ws.info.data_ptr(), stream, 0), "irls2d_adjoint")

    ms_a, _ = time_steps(adj, 1, 1)
    ok = bool(torch.isfinite(gph).all() and torch.isfinite(gw).all())
    # accuracy of the 50-iteration adjoint: dL/dphase = valid * C^T g holds exactly at the solution
    # of H lam = C^T g (its error is the adjoint CG residual); dL/dweight against a 1000-iteration
    # adjoint of the same fixed point
    from core_amd.unwrap2d import image_view
    ident = 0.0
    for k in range(len(shapes)):
        v = ~torch.isnan(image_view(ph, ws, k))
        gk = torch.where(v, image_view(g, ws, k), torch.zeros_like(image_view(g, ws, k)))
        ct = torch.where(v, gk - gk.sum() / v.sum(), torch.zeros_like(gk))
        ident = max(ident, float((image_view(gph, ws, k) - ct).abs().max() / ct.abs().max()))
    gw50 = gw.clone()
    _lib.check(L.irls2d_adjoint(n, H, W, ph.data_ptr(), co.data_ptr(), out.data_ptr(), g.data_ptr(),
                                gph.data_ptr(), gw.data_ptr(), 1e-2, 1000, 0.0, ws.buf.data_ptr(), ws.nbytes,
                                ws.info.data_ptr(), stream, 0), "irls2d_adjoint")
    wrel = max(float((image_view(gw50, ws, k) - image_view(gw, ws, k)).abs().max()
                     / image_view(gw, ws, k).abs().max().clamp_min(1e-30)) for k in range(len(shapes)))
    ws.release()
    del ws, ph, co, out, g, gph, gw, gw50
    torch.cuda.empty_cache()
    return {
        "workload": "16 x 2048x2048, 1024-case generator seeds 3+i; forward 10 IRLS x 50 PCG; adjoint 50 PCG",
        "pixels": npx, "forward_ms": round(ms_f, 2), "forward_mpx_s": round(npx / ms_f / 1e3, 2),
        "adjoint_ms": round(ms_a, 2), "adjoint_mpx_s": round(npx / ms_a / 1e3, 2), "grads_finite": ok,
        "phase_grad_identity_rel_err": float(f"{ident:.3g}"),
        "weight_grad_rel_err_vs_1000_pcg": float(f"{wrel:.3g}"),
    }


def max_scale(args, dev):
    """40000 x 40000 plain solve (sides above the LDS-resident DCT lengths: two-pass transforms),
    10 IRLS x 50 PCG from a captured graph, plus an event-timed eager solve for per-kernel GB/s.
    No oracle runs at this size; the result is checked against the generator's noiseless phase:
    after removing one constant, the share of valid pixels within pi of it (cycle-correct)."""
    import numpy as np
    import torch
    from core_amd import solve_device, synthetic as S, _lib
    from core_amd.unwrap2d import Workspace
    wl = S.single_40000()
    t0 = time.time()

    def progress(done, total):
        if done % 4096 < 1024 or done == total:
            print(f"max_scale: generated {done}/{total} rows ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)

    psi, coh, true = S.generate(wl, band_rows=1024, progress=progress)
    gen_s = time.time() - t0
    shapes = [(wl.H, wl.W)]
    ws = Workspace(shapes, dev)
    ph = upload(ws, [psi], dev)
    co = upload(ws, [coh], dev)
    tr = upload(ws, [true], dev)
    valid = upload(ws, [(~np.isnan(psi)).astype(np.float32)], dev) > 0.5
    del psi, coh, true
    out = torch.empty(ws.total, dtype=torch.float32, device=dev)
    n_irls = getattr(args, "max_scale_irls", None) or N_IRLS
    kw = dict(max_iter=n_irls, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0, out=out, workspace=ws)
    print(f"max_scale: inputs in HBM, workspace {ws.nbytes / 2**30:.1f} GiB; graph solve", file=sys.stderr, flush=True)
    ms, _ = time_steps(lambda: solve_device(ph, co, shapes, graph=True, **kw), 1, 1)
    info = ws.info.cpu().tolist()[:4]
    d = (out - tr)[valid]
    sub = d[torch.randint(0, d.numel(), (4_000_000,), device=dev)]
    c = sub.median()
    dd = d - c
    cyc = float((dd.abs() < np.pi).float().mean())
    rms = float(dd.pow(2).mean().sqrt())
    finite = bool(torch.isfinite(out[valid]).all())
    del d, dd, sub
    print(f"max_scale: {ms / 1e3:.1f} s per solve; event-timed eager solve", file=sys.stderr, flush=True)
    _lib.kernel_stats_reset()
    solve_device(ph, co, shapes, graph=False, time_kernels=True, **kw)
    stats = _lib.kernel_stats()
    npx = wl.H * wl.W
    kernels = {}
    for k, (tot, n) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        e = {"total_ms": round(tot, 1), "launches": n, "mean_us": round(tot / n * 1e3, 1)}
        if k in ALG_BPX_LONG:
            e["gbs"] = round(ALG_BPX_LONG[k] * npx / (tot / n / 1e3) / 1e9, 1)
        kernels[k] = e
    peak_gib = torch.cuda.max_memory_allocated(dev) / 2**30
    ws.release()
    del ws, ph, co, tr, valid, out
    torch.cuda.empty_cache()
    return {
        "workload": "40000x40000, ramp 0-1500 rad, 20 bumps, 3 pi fault, 5% masked blocks, seed 4; "
                    f"{n_irls} IRLS x {N_CG} PCG; rows: four-step two-pass DCT (200 x 200); "
                    "columns: tridiagonal solves",
        "pixels": npx, "ms_per_solve": round(ms, 1), "mpx_s": round(npx / ms / 1e3, 2),
        "solve_model_frac_hbm": round(npx * model_bytes_per_px(n_irls) / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "info": info, "finite": finite,
        "vs_truth": {"cycle_correct_frac": round(cyc, 6), "rms_rad": round(rms, 4),
                     "note": "the 3 pi fault wraps to a pi step, so any unwrapper leaves one side of it "
                             "2 pi from the noiseless phase; the FP64 oracle scores 0.73 on a faulted "
                             "512x512 grid of this generator and 1.0 without the fault"},
        "generate_s": round(gen_s, 1), "torch_peak_gib": round(peak_gib, 1), "kernels": kernels,
    }


def goldstein_bench(args, dev, with_cpu=True):
    """Goldstein filter (the unwrapper's upstream step, utils_goldstein.py) on an 8192 x 8192
    complex64 interferogram: GPU time per call (inputs in HBM), the GB/s of its algorithmic bytes
    (phase 8 + corr 4 + out 8 B/px), and the float32 numpy restatement of
    the reference's CPU path on a 2048 x 2048 crop (one thread, as its scipy.fft workers=1)."""
    import numpy as np
    import torch
    from core_amd.goldstein import goldstein_device
    n = 8192
    g = torch.Generator(device=dev).manual_seed(5)
    ph = torch.polar(0.5 + torch.rand(n, n, device=dev, generator=g), torch.rand(n, n, device=dev, generator=g) * 6.2832)
    cr = torch.rand(n, n, device=dev, generator=g)
    out = goldstein_device(ph, cr, psize=32)
    ms, _ = time_steps(lambda: goldstein_device(ph, cr, psize=32, out=out), 10, 2)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    e[0].record()
    goldstein_device(ph, cr, psize=32, out=out)
    e[1].record()
    torch.cuda.synchronize()
    res = {
        "workload": "8192x8192 complex64 (random phase, amplitude 0.5-1.5), corr U(0,1), psize 32, threshold 0.5, "
                    "valid-patch-count normalisation (the reference's GPU path)",
        "ms": round(ms, 3), "mpx_s": round(n * n / ms / 1e3, 1),
        "alg_bytes_per_px": 20,
        "alg_gbs": round(20 * n * n / (ms / 1e3) / 1e9, 1),
        "single_call_ms": round(e[0].elapsed_time(e[1]), 3),
    }
    del ph, cr, out
    torch.cuda.empty_cache()
    if with_cpu:
        import time
        from oracle import goldstein as G
        rng = np.random.default_rng(5)
        m = 2048
        z = ((0.5 + rng.uniform(0, 1, (m, m))) * np.exp(1j * rng.uniform(0, 6.2832, (m, m)))).astype(np.complex64)
        c = rng.uniform(0, 1, (m, m)).astype(np.float32)
        t0 = time.time()
        G.filter(z, c, 32, 32, 0.5, "count", np.float32)
        dt = time.time() - t0
        res["cpu_baseline"] = {"value": round(m * m / dt / 1e6, 3), "unit": "Mpx/s", "cores": 1, "kind": "port",
                               "sample": f"2048x2048 crop, float32 numpy restatement (batched FFT2), {dt:.2f} s"}
    return res


def gaussian_bench(args, dev, with_cpu=True):
    """NaN-aware Gaussian filter (utils_gaussian.py, the GPU path) on an 8192 x 8192 complex64
    field, sigma 2 px (17 taps per axis): GPU time per call, GB/s of its algorithmic bytes (data
    8 + out 8 B/px), and the float64 numpy restatement on a 2048 x 2048 crop (one thread; the
    reference's CPU path is OpenCV's sepFilter2D, not available here)."""
    import numpy as np
    import torch
    from core_amd.gaussian import gaussian_device
    n = 8192
    g = torch.Generator(device=dev).manual_seed(6)
    z = torch.polar(0.5 + torch.rand(n, n, device=dev, generator=g), torch.rand(n, n, device=dev, generator=g) * 6.2832)
    out = gaussian_device(z, None, 2.0)
    ms, _ = time_steps(lambda: gaussian_device(z, None, 2.0), 10, 2)
    res = {
        "workload": "8192x8192 complex64 (random phase, amplitude 0.5-1.5), sigma 2 px, truncate 4, no weight",
        "ms": round(ms, 3), "mpx_s": round(n * n / ms / 1e3, 1),
        "alg_bytes_per_px": 16, "alg_gbs": round(16 * n * n / (ms / 1e3) / 1e9, 1),
    }
    del z, out
    torch.cuda.empty_cache()
    if with_cpu:
        import time
        from oracle import gaussian as G
        rng = np.random.default_rng(6)
        m = 2048
        z = (0.5 + rng.uniform(0, 1, (m, m))) * np.exp(1j * rng.uniform(0, 6.2832, (m, m)))
        t0 = time.time()
        G.nanconv(z, None, 2.0)
        dt = time.time() - t0
        res["cpu_baseline"] = {"value": round(m * m / dt / 1e6, 3), "unit": "Mpx/s", "cores": 1, "kind": "port",
                               "sample": f"2048x2048 crop, float64 numpy restatement, {dt:.2f} s"}
    return res


def end_to_end_8192(args, dev, psi, coh):
    """The drop-in call on the 8192 x 8192 host arrays: crop, H2D, solve (eager), device epilogue,
    D2H, as a caller of the reference's irls_unwrap_2d sees it.  Returns (timing dict, output)."""
    import torch
    from core_amd import irls_unwrap_2d
    kw = dict(device='cuda', max_iter=N_IRLS, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0)
    out = irls_unwrap_2d(psi, coh, **kw)  # warm-up: plans, workspace
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = irls_unwrap_2d(psi, coh, **kw)
        times.append(time.perf_counter() - t0)
    ms = min(times) * 1e3
    npx = psi.size
    return {
        "workload": "drop-in irls_unwrap_2d on the 8192x8192 numpy inputs: crop, H2D, solve (eager launches), "
                    "device epilogue (k-median, 4-connected labels, circular-mean alignment), D2H",
        "ms": round(ms, 1), "mpx_s": round(npx / ms / 1e3, 2),
    }, out


def _accuracy(u, cc, u64, cc64, u32, psi):
    """RMS vs the FP64 trajectory, cycle agreement with the float32 reference path, conncomp
    equality and the rewrap of the congruent output psi + 2 pi k (FP64), over valid pixels."""
    import numpy as np
    v = ~np.isnan(psi)
    d = u[v].astype(np.float64) - u64[v].astype(np.float64)
    k = np.round((u[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    k32 = np.round((u32[v].astype(np.float64) - psi[v]) / (2 * np.pi))
    cong = psi[v].astype(np.float64) + 2 * np.pi * k
    rms = float(np.sqrt(np.mean(d * d)))
    agree = float(np.mean(k == k32))
    return {"rms_vs_fp64_rad": float(f"{rms:.3e}"), "k_agree_vs_fp32": round(agree, 6),
            "conncomp_equal": bool(np.array_equal(cc, cc64)),
            "rewrap_max_rad": float(f"{float(np.abs(np.angle(np.exp(1j * (cong - psi[v])))).max()):.1e}"),
            "pass": bool(rms <= 1e-3 and agree >= 0.9999 and np.array_equal(cc, cc64))}


def cpu_baseline_leg(psi, coh, e2e_out, dev, rows=2048, cols=2048):
    """The reference path, where the oracle runs (nothing here is timed as the GPU's value):

    cpu_baseline  the float32 torch restatement of the reference solver (oracle.irls_torch32), on
                  this host's cores, on the top-left rows x cols of the 8192 x 8192 workload (the
                  sample of the metric's workload) and on the 1024 x 1024 CPU workload in full;
    accuracy      per workload, the GPU's drop-in output against the FP64 trajectory (RMS) and the
                  reference's float32 path (cycle agreement): 1024^2 against the CPU oracles (numpy
                  FP64, torch float32), 8192^2, three stack bursts and the 40000^2 generator at 2048^2
                  against the same torch restatement run on the HIP device in float64 and float32."""
    import numpy as np
    import torch
    from core_amd import irls_unwrap_2d, irls_unwrap_2d_batch, synthetic as S
    from oracle import irls_fp64, irls_torch32, post
    ncores = len(os.sched_getaffinity(0))
    threads = min(16, ncores)
    torch.set_num_threads(threads)
    fixed = dict(max_iter=N_IRLS, tol=0.0, cg_max_iter=N_CG, cg_tol=0.0)

    def timed_cpu(ph, wt):
        t0 = time.perf_counter()
        _, ph_c, nm, pf, wf = post.prologue(ph, wt)
        u = irls_torch32.solve(pf, wf, ~nm, N_IRLS, 0.0, N_CG, 0.0, 1e-2)
        out = post.epilogue(u, ph_c, nm, 30)
        return time.perf_counter() - t0, out

    dt, _ = timed_cpu(np.ascontiguousarray(psi[:rows, :cols]), np.ascontiguousarray(coh[:rows, :cols]))
    base = {
        "value": round(rows * cols / dt / 1e6, 4), "unit": "Mpx/s", "cores": threads, "kind": "port",
        "sample": f"top-left {rows}x{cols} crop of the 8192x8192 workload, 10 IRLS x 50 PCG, float32 "
                  f"torch-CPU restatement incl. host epilogue, {dt:.1f} s on {threads} of {ncores} host threads",
    }
    # the 1024 x 1024 CPU workload in full
    p0, c0, _ = S.generate(S.single_1024())
    dt0, ep32 = timed_cpu(p0, c0)
    base["cpu_workload_1024"] = {"value": round(p0.size / dt0 / 1e6, 4), "unit": "Mpx/s", "seconds": round(dt0, 2),
                                 "cores": threads, "sample": "1024x1024 CPU workload in full, 10 IRLS x 50 PCG"}
    acc = {}
    u, cc = irls_unwrap_2d(p0, c0, device='cuda', **fixed)
    u64, cc64 = irls_fp64.irls_unwrap_2d(p0, c0, **fixed)
    acc["cpu_workload_1024"] = {"oracles": "numpy FP64 and torch float32 on the CPU",
                                **_accuracy(u, cc, u64, cc64, ep32[0], p0)}

    def on_device(ph, wt):
        r64 = irls_torch32.irls_unwrap_2d(ph, wt, dtype=torch.float64, device=dev, **fixed)
        torch.cuda.empty_cache()
        r32 = irls_torch32.irls_unwrap_2d(ph, wt, dtype=torch.float32, device=dev, **fixed)
        torch.cuda.empty_cache()
        return r64, r32

    r64, r32 = on_device(psi, coh)
    acc["plain_8192"] = {"oracles": "torch restatement on the HIP device, float64 and float32",
                         **_accuracy(e2e_out[0], e2e_out[1], r64[0], r64[1], r32[0], psi)}
    del r64, r32
    wls = S.stack_48()
    small = sorted(range(len(wls)), key=lambda i: wls[i].H * wls[i].W)[:3]
    ims = [S.generate(wls[i])[:2] for i in small]
    res = irls_unwrap_2d_batch([a for a, _ in ims], [b for _, b in ims], device='cuda', **fixed)
    bursts = []
    for i, (ph, wt), (u, cc) in zip(small, ims, res):
        r64, r32 = on_device(ph, wt)
        bursts.append({"burst": i, "shape": [wls[i].H, wls[i].W], **_accuracy(u, cc, r64[0], r64[1], r32[0], ph)})
    acc["stack_48_smallest3"] = {"oracles": "torch restatement on the HIP device, float64 and float32",
                                 "note": "solved as one batch; an image's solve is bitwise independent of its "
                                         "batch-mates", "bursts": bursts}
    p4, c4, _ = S.generate(S.scaled(S.single_40000(), 2048, 2048))
    u, cc = irls_unwrap_2d(p4, c4, device='cuda', **fixed)
    r64, r32 = on_device(p4, c4)
    acc["max_scale_generator_2048"] = {"oracles": "torch restatement on the HIP device, float64 and float32",
                                       **_accuracy(u, cc, r64[0], r64[1], r32[0], p4)}
    return base, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batched-steps", type=int, default=2)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-filters", action="store_true")
    ap.add_argument("--max-scale", action="store_true",
                    help="also solve the 40000 x 40000 workload (about 4 minutes, ~140 GB of HBM)")
    ap.add_argument("--only-max-scale", action="store_true", help="run only the 40000 x 40000 workload")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the 48-burst batched stack, with per-kernel durations")
    ap.add_argument("--stack-irls", type=int, default=None,
                    help="IRLS steps of the batched stack (default 10; fewer for counter runs under a profiler)")
    ap.add_argument("--max-scale-irls", type=int, default=None,
                    help="IRLS steps of the 40000 x 40000 solve (default 10; fewer for counter runs under a profiler)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the 48 sequential plain calls the batched stack is compared with")
    ap.add_argument("--column-mode", choices=["auto", "transpose", "direct", "slab", "tridiag"], default="auto",
                    help="column-pass strategy for every solve (irls2d_set_column_mode; default automatic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the per-GPU shard is what one GPU shows; see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP GPU", file=sys.stderr)
        sys.exit(1)
    import core_amd._lib as L
    L.load()
    dev = torch.device("cuda", 0)
    from core_amd import autotune
    n_tuned = autotune.apply()  # plans measured by `python -m core_amd.autotune` on this GPU type
    L.set_column_mode({"auto": L.COLUMN_AUTO, "transpose": L.COLUMN_TRANSPOSE, "direct": L.COLUMN_DIRECT,
                       "slab": L.COLUMN_SLAB, "tridiag": L.COLUMN_TRIDIAG}[args.column_mode])
    if args.only_max_scale:
        print(json.dumps({"max_scale": max_scale(args, dev)}), flush=True)
        return
    if args.only_batched:
        print(json.dumps({"batched": batched_stack(args, dev, kernels=True)}), flush=True)
        return
    npx = 8192 * 8192
    psi, coh, ms, wall_ms, info, stats = plain_8192(args, dev)
    dom, roof = roofline(stats, npx)
    e2e, e2e_out = end_to_end_8192(args, dev, psi, coh)
    result = {
        "metric": METRIC,
        "value": round(npx / ms / 1e3, 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain 8192x8192 float32, ramp 0-300 rad, 6 bumps, 3pi fault, 5% masked 16x16 blocks, seed 1",
            "irls_iters": N_IRLS, "pcg_iters": N_CG, "pixels": npx, "graph": True, "autotuned_plans": n_tuned,
            "solver_info": {"irls_iters": info[1], "pcg_iters": info[2], "valid_px": info[3]},
        },
        "roofline": roof,
        "kernel_roofline": kernel_fracs(stats, npx),
        "end_to_end": e2e,
        "solve_model": {
            "bytes_per_px": model_bytes_per_px(),
            "achieved_gbs": round(npx * model_bytes_per_px() / (ms / 1e3) / 1e9, 1),
            "frac": round(npx * model_bytes_per_px() / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])},
    }
    if not args.no_batched:
        result["batched"] = batched_stack(args, dev)
    if not args.no_adjoint:
        result["adjoint"] = adjoint_batch(args, dev)
    if not args.no_filters:
        result["filters"] = {"goldstein": goldstein_bench(args, dev, with_cpu=not args.no_cpu),
                             "gaussian": gaussian_bench(args, dev, with_cpu=not args.no_cpu)}
    if args.max_scale:
        result["max_scale"] = max_scale(args, dev)
    if not args.no_cpu:
        result["cpu_baseline"], result["accuracy"] = cpu_baseline_leg(psi, coh, e2e_out, dev)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput benchmark: deconvolved volumes/s of phase-space RL on one GPU.

Workload (default): 101 x 101 lenslets (1515 x 1515 px), 51 planes
(-100..100 um step 4), 177 active views, 8 RL iterations per volume, synthetic
PSF / 400-sphere truth / Poisson measurement at 1000 counts peak (seeded).
One "step" = one full volume: init (recon:72-73) + 8 iterations of 177 view
steps + sample_filter.  Inputs (PSF tables, image) are resident in HBM
before the timed region starts.

Prints one JSON line.  ``roofline`` describes the dominant kernel, timed
live with HIP events on the launch stream, priced with the algorithmic bytes
and flops of the path (3 volume passes of 2 B per voxel and view step for the
16-bit paths, 4 B for FP32; 2 n1 n2 sum_z K_z^2 flops per projection);
``cpu_baseline`` is the oracle's C/OpenMP FP32 restatement timed on a bounded
sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15}  # dense MFMA peaks, flop/s
HBM_PEAK = 8.0e12


def algorithmic_flops_per_view(wl) -> float:
    """2 * n1 * n2 * sum_z K_z^2 with K_z = 30 r_z + 1 px (one projection
    direction, one view)."""
    return 2.0 * wl.n1 * wl.n2 * sum((30 * r + 1) ** 2 for r in wl.radii)


def algorithmic_bytes_per_view(wl, precision: str, kernel: str) -> float:
    """Algorithmic HBM bytes of one launch: a view step moves the volume three
    times at the operand width (2 B per voxel for BF16/FP16, 4 B for FP32):
    the forward reads it once, the backprojection + update reads and writes it
    (2 of the 3 passes)."""
    vox = wl.n1 * wl.n2 * wl.Z * 225
    esz = 4 if precision == "fp32" else 2
    return float((2 if kernel == "backward_update" else 1) * esz * vox)


def design_bytes_per_view(wl, precision: str, kernel: str) -> float:
    """Bytes this design's launch must move: the backprojection + update reads
    and writes the FP32 volume (the update stays FP32: for BF16 as the two
    16-bit planes of its split form, whose upper plane is the next forward's
    operand; for FP16 as FP32 X plus the FP16 shadow it writes), and reads the
    view's taps, its 1/B1 table and the ratio image; the forward reads the
    16-bit operand plane (or X) and the taps."""
    vox = wl.n1 * wl.n2 * wl.Z * 225
    esz = 4 if precision == "fp32" else 2
    taps = 225 * sum((2 * r + 1) ** 2 for r in wl.radii)
    if kernel == "backward_update":
        shadow = 2 * vox if precision == "fp16" else 0
        return 8.0 * vox + shadow + esz * taps + 4.0 * taps + 4.0 * wl.n1 * wl.n2
    return float(esz * vox + esz * taps)


def roofline_for(wl, precision: str, kernel: str, avg_s: float) -> dict:
    """Price one launch against the MFMA peak and the HBM peak; the bound is
    the one that takes longer at peak (arithmetic intensity vs the machine
    balance), and `frac` is that time over the measured launch time."""
    flops = algorithmic_flops_per_view(wl)
    nbytes = algorithmic_bytes_per_view(wl, precision, kernel)
    dbytes = design_bytes_per_view(wl, precision, kernel)
    t_mfma, t_hbm = flops / PEAK[precision], nbytes / HBM_PEAK
    common = {"kernel": kernel, "avg_launch_ms": avg_s * 1e3, "traffic": None,
              "algorithmic_per_launch": f"{flops:.4e} flop (2*n1*n2*sum_z(30r_z+1)^2), "
                                        f"{nbytes:.4e} B ({2 if kernel == 'backward_update' else 1} "
                                        f"volume pass(es) x {4 if precision == 'fp32' else 2} B/voxel)",
              "time_at_peak_ms": {"mfma": t_mfma * 1e3, "hbm": t_hbm * 1e3},
              "mfma_TFLOPs": flops / avg_s / 1e12, "mfma_frac": flops / avg_s / PEAK[precision],
              "design_bytes_per_launch": dbytes,
              "design_GBs": dbytes / avg_s / 1e9, "design_frac": dbytes / avg_s / HBM_PEAK}
    if t_mfma >= t_hbm:
        return {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": PEAK[precision] / 1e12,
                "unit": "TFLOP/s", "frac": t_mfma / avg_s, **common}
    return {"bound": "hbm", "achieved": nbytes / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": t_hbm / avg_s, **common}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(wl, tabs, LF, nviews_sample: int):
    """The oracle's FP32 restatements (oracle/cpu_baseline.py) on this host's
    cores, on bounded samples -- the strided form on numpy's BLAS (sgemm),
    the literal full-resolution form in C/OpenMP (oracle/cpu_rl.c):

    * the workload itself: `nviews_sample` strided view steps and one filtered
      plane at its shape, extrapolated to a volume (views x iterations);
    * 33 x 33 lenslets x 11 planes: one whole strided iteration (177 view
      steps + sample_filter) timed, x 8 iterations;
    * 33 x 33 x 11, the reference CPU branch as written: one view step of
      full-resolution conv2(..., 'same') (recon:191-208) timed, x 177 views
      x 8 iterations (+ the strided iteration's filter time)."""
    from mesolf_amd import synth, tables as T
    from oracle import cpu_baseline as CB
    from oracle import mesolf_oracle as O
    from tests.helpers import oracle_compact_from_abi

    cores = CB.threads()
    per_view = sum((2 * r + 1) ** 2 * 225 for r in wl.radii)
    sub = dict(tabs)
    sub["uv"] = tabs["uv"][:nviews_sample]
    sub["weights"] = tabs["weights"][:nviews_sample]
    sub["fwd"] = tabs["fwd"][:per_view * nviews_sample]
    st = CB.StridedTables(oracle_compact_from_abi(wl, sub))
    Xl = O.to_lenslet(O.initial_volume(LF.astype(np.float64), wl.Z), 15).astype(np.float32)
    t0 = time.perf_counter()
    CB.view_steps_strided(Xl, LF, st)
    t_view = (time.perf_counter() - t0) / nviews_sample
    X2 = np.ones((wl.I1, wl.I2, 2), np.float32)
    t0 = time.perf_counter()
    CB.filter_planes(X2)
    t_plane = (time.perf_counter() - t0) / 2
    t_volume = wl.iters * (len(tabs["uv"]) * t_view + wl.Z * t_plane)

    # 33 x 33 x 11: one whole strided iteration
    w1 = synth.CFG1
    tabs1 = T.synthetic_compact(w1)
    cp1 = oracle_compact_from_abi(w1, tabs1)
    st1 = CB.StridedTables(cp1)
    lf1 = np.random.default_rng(1).poisson(100.0, (w1.I1, w1.I2)).astype(np.float32)
    X1 = O.initial_volume(lf1.astype(np.float64), w1.Z)
    t0 = time.perf_counter()
    Xl1 = CB.view_steps_strided(O.to_lenslet(X1, 15).astype(np.float32), lf1, st1)
    t1_views = time.perf_counter() - t0
    t0 = time.perf_counter()
    CB.filter_planes(O.from_lenslet(Xl1.astype(np.float64), 15))
    t1_filter = time.perf_counter() - t0
    t1_iter = t1_views + t1_filter
    # ... and one literal view step
    u, v = cp1.views[0]
    Hv, Htv = CB.view_slices_from_synth(w1, u, v)
    t0 = time.perf_counter()
    CB.view_step_literal(X1.astype(np.float32), lf1, Hv, Htv, u, v, cp1.weights[0])
    t1_lit_view = time.perf_counter() - t0
    t1_lit_volume = w1.iters * (len(cp1.views) * t1_lit_view + t1_filter)
    return {
        "value": 1.0 / t_volume,
        "unit": "volumes/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP32 strided form on numpy BLAS ({cores} threads): {nviews_sample} view "
                   f"steps ({t_view:.3f} s each) + 1 filtered plane ({t_plane:.3f} s) of this workload, "
                   f"x{len(tabs['uv'])} views x{wl.iters} iterations (+{wl.Z} planes per iteration): "
                   f"{t_volume:.0f} s per volume"),
        "cfg1_33x33x11": {
            "strided_iteration_s": round(t1_iter, 3),
            "strided_volume_s_x8": round(w1.iters * t1_iter, 2),
            "literal_view_step_s": round(t1_lit_view, 3),
            "literal_volume_s_x177x8": round(t1_lit_volume, 1),
            "note": ("strided: one whole iteration timed (177 view steps + sample_filter), x8; literal: "
                     "the reference CPU branch as written (full-resolution conv2 'same' of "
                     f"{w1.I1}x{w1.I2} images with {w1.IMG}x{w1.IMG} kernels, recon:191-208), one view "
                     "step timed, x177 views x8 iterations plus the strided filter time"),
        },
    }


def oracle_check(wl, tabs, LF, psf, precisions, nviews: int = 1):
    """Relative L2 error, against the FP64 oracle, of each GPU path's volume
    after the first `nviews` view steps from the initial volume (filter off)."""
    import mesolf_amd
    from oracle import mesolf_oracle as O
    from tests.helpers import oracle_compact_from_abi

    per_view = sum((2 * r + 1) ** 2 * 225 for r in wl.radii)
    sub = dict(tabs)
    sub["uv"] = tabs["uv"][:nviews]
    sub["weights"] = tabs["weights"][:nviews]
    sub["fwd"] = tabs["fwd"][:per_view * nviews]
    cp = oracle_compact_from_abi(wl, sub)
    t0 = time.perf_counter()
    ref = O.rl_compact(LF.astype(np.float64), cp, 1, views_limit=nviews, filter_planes=False)
    t_ref = time.perf_counter() - t0
    out = {"view_steps": nviews, "oracle_s": round(t_ref, 1)}
    nr = np.linalg.norm(ref.ravel())
    for prec in precisions:
        p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
        p.set_image(LF)
        p.init_volume()
        p.run(1, views_limit=nviews, filter=False)
        x = p.get_volume(f64=True)
        out[f"rel_l2_{prec}"] = float(np.linalg.norm((x - ref).ravel()) / nr)
        p.close()
        del x
    return out


def _movie_patches(psf, grids, frames, rng, groups=4):
    """Movie on the largest grid, crop offsets per patch."""
    from mesolf_amd import serving

    n1m = max(g[0] for g in grids)
    n2m = max(g[1] for g in grids)
    t = time.perf_counter()
    movie = serving.make_patch(psf, n1m, n2m, frames, rng, groups=groups)
    offs = [(int(rng.integers(0, n1m - g[0] + 1)), int(rng.integers(0, n2m - g[1] + 1))) for g in grids]
    log(f"movie {n1m}x{n2m}, {groups} activity groups, {frames} frames in {time.perf_counter() - t:.1f}s")
    return movie, offs


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, int(np.ceil(q * len(xs))) - 1))] if xs else None


def bench_batched(args):
    """12 uneven patches x consecutive frames, 3 warm-start RL iterations per
    frame, all patches of a frame in one batch (shared launches, one graph
    replay per iteration)."""
    import mesolf_amd
    from mesolf_amd import serving, synth, tables

    wl = synth.CFG2
    rng = np.random.default_rng(4)
    grids = serving.patch_grids(rng, args.patches, 61, 121)
    tabs = tables.synthetic_compact(wl)
    ctx = mesolf_amd.Context(0)
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    del tabs
    movie, offs = _movie_patches(psf, grids, args.frames, rng)
    nrng = np.random.default_rng(5)
    timed = min(args.frames, args.timed_frames)
    frames = [[serving.crop(movie.frame(f, nrng), g[0], g[1], *o) for g, o in zip(grids, offs)]
              for f in range(timed + 1)]
    log(f"{len(grids)} patches {grids}; {timed + 1} frames prepared")

    b = mesolf_amd.Batch(psf, capacity=len(grids), precision=args.precision, graph=True)
    ps = [mesolf_amd.Patch(psf, *g, precision=args.precision) for g in grids]
    # frame 0: cold start (recon:72-73), kept for the FP32 comparison
    for p, im in zip(ps, frames[0]):
        p.set_image(im)
        p.init_volume()
        b.attach(p)
    tuned = None if args.no_autotune else b.autotune(views=8)
    log(f"batch tile shape {tuned}")
    t0 = time.perf_counter()
    b.run(args.rl_iters)
    ctx.synchronize()
    t_first = time.perf_counter() - t0
    first = [p.get_volume() for p in ps]
    lat = []
    t0 = time.perf_counter()
    for f in range(1, timed + 1):
        tf = time.perf_counter()
        for p, im in zip(ps, frames[f]):
            p.set_image(im)  # warm start: the volume of the previous frame stays
        b.run(args.rl_iters)
        ctx.synchronize()
        lat.append((time.perf_counter() - tf) * 1e3)
        log(f"frame {f}/{timed}: {lat[-1]:.0f} ms")
    elapsed = time.perf_counter() - t0
    value = len(grids) * timed / elapsed
    b.close()
    for p in ps:
        p.close()

    # plain FP32 per-call path on frame 0 (one patch at a time), and the error
    errs, t32 = [], 0.0
    for g, im, ref in zip(grids, frames[0], first):
        p = mesolf_amd.Patch(psf, *g, precision="fp32")
        p.set_image(im)
        p.init_volume()
        ctx.synchronize()
        t1 = time.perf_counter()
        p.run(args.rl_iters)
        ctx.synchronize()
        t32 += time.perf_counter() - t1
        x = p.get_volume()
        errs.append(float(np.linalg.norm((ref.astype(np.float64) - x).ravel()) /
                          np.linalg.norm(x.astype(np.float64).ravel())))
        p.close()
    fp32_rate = len(grids) / t32
    out = {
        "metric": "deconvolved patch-frames/s, 12 uneven patches (61-121 lenslets/side) x 51 planes, "
                  "3 warm-start RL iterations per frame, batched",
        "value": value, "unit": "patch-frames/s", "n_gpus": 1, "steps": timed, "warmup": 1,
        "ms_per_step": elapsed / timed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic movie (calcium-like traces per sphere group, Poisson noise), "
                "patches are crops of one 121x121 movie",
        "config": {"workload": "cfg4_movie_12_patches", "patches": grids, "planes": 51,
                   "frames_timed": timed, "rl_iterations_per_frame": args.rl_iters,
                   "graph": True, "precision": args.precision, "autotuned": tuned},
        "frame_latency_ms": {"p50": _pct(lat, 0.5), "p99": _pct(lat, 0.99), "max": max(lat)},
        "first_frame_cold_ms": t_first * 1e3,
        "fp32_per_call": {"patch_frames_per_s": fp32_rate, "speedup": value / fp32_rate,
                          "rel_l2_frame0_vs_fp32": {"max": max(errs), "mean": float(np.mean(errs))}},
    }
    print(json.dumps(out), flush=True)
    psf.close()
    ctx.close()


def bench_serving(args):
    """Poisson arrivals of frame x patch requests over 36 patches (81-121
    lenslets/side) into the continuous-batching server with an HBM budget."""
    import mesolf_amd
    from mesolf_amd import serving, synth, tables

    wl = synth.CFG2
    rng = np.random.default_rng(6)
    grids = serving.patch_grids(rng, args.patches, 81, 121)
    tabs = tables.synthetic_compact(wl)
    ctx = mesolf_amd.Context(0)
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    del tabs
    nfr = 4
    movie, offs = _movie_patches(psf, grids, nfr, rng)
    nrng = np.random.default_rng(7)
    imgs = [[serving.crop(movie.frame(f, nrng), g[0], g[1], *o) for f in range(nfr)]
            for g, o in zip(grids, offs)]
    budget = int(args.hbm_budget_gb * (1 << 30))
    srv = mesolf_amd.Server(psf, precision=args.precision, max_slots=args.slots,
                            iterations=args.rl_iters, hbm_budget_bytes=budget)
    # capacity: every patch queued at once, cold round then warm round
    t0 = time.perf_counter()
    ts = [srv.submit(i, imgs[i][0], want_volume=(i < 2)) for i in range(len(grids))]
    served0 = {}
    for i, t in enumerate(ts):
        _, vol = srv.wait(t)
        if vol is not None:
            served0[i] = vol
    t_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    for t in [srv.submit(i, imgs[i][1]) for i in range(len(grids))]:
        srv.wait(t)
    t_warm = time.perf_counter() - t0
    capacity = len(grids) / t_warm
    log(f"capacity: cold round {t_cold:.1f}s, warm round {t_warm:.1f}s -> {capacity:.2f} req/s")
    # open-loop Poisson arrivals at load x capacity
    lam = args.load * capacity
    arr = np.cumsum(np.random.default_rng(8).exponential(1.0 / lam, int(lam * args.duration * 2) + 8))
    arr = arr[arr < args.duration]
    counters = [2] * len(grids)
    pick = np.random.default_rng(9).integers(0, len(grids), len(arr))
    tickets = []
    t0 = time.perf_counter()
    for a, i in zip(arr, pick):
        dt = a - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        tickets.append(srv.submit(int(i), imgs[i][counters[i] % nfr]))
        counters[i] += 1
    lat = [srv.wait(t)[0] for t in tickets]
    span = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # error of two served (cold, first-frame) results against the FP32 per-call path
    errs = []
    for i, vol in served0.items():
        p = mesolf_amd.Patch(psf, *grids[i], precision="fp32")
        p.set_image(imgs[i][0])
        p.init_volume()
        p.run(args.rl_iters)
        x = p.get_volume()
        p.close()
        errs.append(float(np.linalg.norm((vol.astype(np.float64) - x).ravel()) /
                          np.linalg.norm(x.astype(np.float64).ravel())))
    nominal = len(grids) * serving.FPS
    out = {
        "metric": "p99 latency of frame x patch deconvolution requests (Poisson arrivals), "
                  "36 patches (81-121 lenslets/side) x 51 planes, 3 warm-start RL iterations",
        "value": _pct(lat, 0.99), "unit": "ms", "n_gpus": 1, "steps": len(lat), "warmup": 2 * len(grids),
        "ms_per_step": span / max(1, len(lat)) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic movie crops (calcium-like traces, Poisson noise)",
        "config": {"workload": "cfg5_serving_36_patches", "patches": grids, "planes": 51,
                   "slots": args.slots, "hbm_budget_gb": args.hbm_budget_gb,
                   "rl_iterations": args.rl_iters, "offered_load": args.load,
                   "arrival_rate_per_s": lam, "duration_s": args.duration},
        "latency_ms": {"p50": _pct(lat, 0.5), "p99": _pct(lat, 0.99), "mean": float(np.mean(lat)),
                       "max": max(lat)},
        "sustained_volumes_per_s": len(lat) / span,
        "capacity_volumes_per_s": capacity,
        "nominal_arrivals_per_s": nominal,
        "nominal_over_capacity": nominal / capacity,
        "server": {k: st[k] for k in ("completed", "evictions", "cold_starts", "iterations_run",
                                      "peak_resident_bytes", "mean_slots")},
        "rel_l2_vs_fp32_per_call": errs,
    }
    print(json.dumps(out), flush=True)
    psf.close()
    ctx.close()


def bench_mex(args):
    """The MEX operator at the reference demo's shape (mex:276; 396 x 396
    bitmaps, 144 per call, kw = 241): host arrays in and out, as the MATLAB
    gateway calls it."""
    import mesolf_amd

    W = H = 396
    n, kw = 144, 241
    rng = np.random.default_rng(11)
    bitmap = (rng.random((W, H, n)) < 0.005).astype(np.float32) * rng.random((W, H, n)).astype(np.float32)
    K = rng.random((kw, kw, 15, 15)).astype(np.float32)
    mesolf_amd.mex_lfm_convolution_vec(bitmap[:, :, :4], K)  # warm-up (operand table, code)
    times = []
    for _ in range(max(1, args.steps)):
        t = time.perf_counter()
        mesolf_amd.mex_lfm_convolution_vec(bitmap, K)
        times.append(time.perf_counter() - t)
    dt = min(times)
    mac = W * H * n * kw * kw
    print(json.dumps({
        "metric": "MEX operator throughput, 396x396 bitmaps x 144, kw = 241 (host arrays in/out)",
        "value": mac / dt / 1e12, "unit": "TMAC/s", "n_gpus": 1, "steps": len(times), "warmup": 1,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (MFMA, f32 accumulate)", "data": "synthetic sparse bitmaps, random kernels",
        "config": {"workload": "mex_demo_396x396x144_kw241"},
        "roofline": {"bound": "mfma", "achieved": 2 * mac / dt / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "unit": "TFLOP/s", "frac": 2 * mac / dt / PEAK["fp32"],
                     "note": "call time incl. 2 x 90 MB host transfers"},
    }), flush=True)


def bench_seedlib(args):
    """The seed-library projection loop of
    gen_seed_library_local_with_ring_bg_same_size.m:263-318 at the demo's
    patch-1 shape: 51 planes x 4 arrays (sample, sample mask, ring, ring mask)
    of 396 x 396 slices, N = 88-144 slices per plane, kw = 241, each slice's
    projection added to its segment.  All four arrays, the kernels
    (Ht(:,:,:,:,z), 241 x 241 x 15 x 15 x 51) and the outputs are resident in
    HBM; one mlf_lfm_convolution_segments_device call per array."""
    import torch

    import mesolf_amd

    W = H = 396
    Z, kw, nseg = 51, 241, 144
    # plane z is covered by the first n_z segments: 144 in the middle planes,
    # 88 at the ends (the demo log's range of N per call)
    n_z = [int(round(88 + 56 * (1 - abs(z - 25) / 25))) for z in range(Z)]
    planes, seg_ptr = [], [0]
    for s_ in range(nseg):
        pl = [z for z in range(Z) if s_ < n_z[z]]
        planes += pl
        seg_ptr.append(len(planes))
    ns = len(planes)
    ctx = mesolf_amd.Context(0)
    g = torch.Generator(device="cuda").manual_seed(7)
    # sparse segment stacks: a disc of ~40 px radius around each segment's centre
    yy, xx = torch.meshgrid(torch.arange(W, device="cuda"), torch.arange(H, device="cuda"), indexing="ij")
    arrays = []
    for a in range(4):
        x = torch.zeros((ns, H, W), device="cuda", dtype=torch.float32)  # [slice][y][x] = [W, H, n] F order
        cx = torch.randint(60, W - 60, (nseg,), device="cuda", generator=g)
        cy = torch.randint(60, H - 60, (nseg,), device="cuda", generator=g)
        for s_ in range(nseg):
            disc = (((xx - cx[s_]) ** 2 + (yy - cy[s_]) ** 2) < 1600).T.float()
            vals = torch.rand((seg_ptr[s_ + 1] - seg_ptr[s_], H, W), device="cuda", generator=g)
            x[seg_ptr[s_]:seg_ptr[s_ + 1]] = vals * disc if a % 2 == 0 else (vals * disc > 0).float()
        arrays.append(x)
    K = torch.rand((Z, 15, 15, kw, kw), device="cuda", generator=g, dtype=torch.float32)  # [kw,kw,15,15,Z] F
    outs = [torch.empty((nseg, H, W), device="cuda", dtype=torch.float32) for _ in range(4)]
    torch.cuda.synchronize()

    def run():
        for a in range(4):
            mesolf_amd.lfm_convolution_segments_device(ctx, arrays[a].data_ptr(), W, H, ns, planes, seg_ptr,
                                                       K.data_ptr(), kw, Z, outs[a].data_ptr())

    # warm-up on a 2-segment prefix (code objects, operand tables)
    mesolf_amd.lfm_convolution_segments_device(ctx, arrays[0].data_ptr(), W, H, seg_ptr[2], planes[:seg_ptr[2]],
                                               seg_ptr[:3], K.data_ptr(), kw, Z, outs[0].data_ptr())
    log(f"seed library: {ns} slices per array, N per plane {min(n_z)}-{max(n_z)}")
    times = []
    for i in range(max(1, args.steps)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"seed library step {i + 1} {times[-1]:.3f} s")
    dt = min(times)
    mac = 4.0 * ns * W * H * kw * kw
    print(json.dumps({
        "metric": "seed-library MEX projection loop, 396x396 slices, 51 planes x 4 arrays, N = 88-144 per "
                  "plane, kw = 241 (gen_seed_library_local_with_ring_bg_same_size.m:263-318)",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": len(times), "warmup": 1, "ms_per_step": dt * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (MFMA, f32 accumulate)", "data": "synthetic sparse segment stacks, random kernels",
        "config": {"workload": "seedlib_396x396_51x4_kw241", "slices_per_array": ns, "segments": nseg,
                   "n_per_plane": [min(n_z), max(n_z)]},
        "roofline": {"bound": "mfma", "achieved": 2 * mac / dt / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "unit": "TFLOP/s", "frac": 2 * mac / dt / PEAK["fp32"], "traffic": None},
        "reference_published_s": {"patch1_N88_144": 272.26, "patch2_N1_128": 73.19,
                                  "note": "TITAN V, reference MEX; this run's N per plane is synthetic"},
    }), flush=True)
    ctx.close()


def bench_prep(args):
    """Host preparation of the recon drop-in at the benchmark patch's dense
    PSF size: H, Ht [525, 525, 15, 15, 51] single (12.6 GB each), through
    mlf_psf_create_dense (recon:13-40: normalisation, phase-space tables,
    weights, upload) and through the drop-in twice (the second call reuses the
    tables of identical H/Ht)."""
    import mesolf_amd
    from mesolf_amd import synth

    wl = synth.CFG2
    S, Z = wl.IMG, wl.Z
    c = S // 2
    x = np.arange(S)
    d2 = (x[:, None] - c) ** 2 + (x[None, :] - c) ** 2
    view = (1.0 + 0.01 * np.add.outer(np.arange(15), np.arange(15))).astype(np.float32)
    t = time.perf_counter()
    H = np.empty((S, S, 15, 15, Z), np.float32, order="F")
    Ht = np.empty((S, S, 15, 15, Z), np.float32, order="F")
    for z, r in enumerate(wl.radii):
        m = (d2 <= (15 * r + 7) ** 2).astype(np.float32)
        H[:, :, :, :, z] = m[:, :, None, None] * view[None, None, :, :]
        mt = (d2 <= (15 * r + 14) ** 2).astype(np.float32)
        Ht[:, :, :, :, z] = mt[:, :, None, None] * view[None, None, :, :]
    log(f"dense H, Ht {H.nbytes / 1e9:.1f} GB each built in {time.perf_counter() - t:.1f}s")
    ctx = mesolf_amd.Context(0)
    t = time.perf_counter()
    psf = mesolf_amd.PSF.from_dense(ctx, H, Ht)
    ctx.synchronize()
    t_prep = time.perf_counter() - t
    rf, rb = psf.radii()
    log(f"mlf_psf_create_dense {t_prep:.2f} s, {psf.nviews} views, max radius {rf.max()}/{rb.max()}")
    psf.close()
    LF = np.random.default_rng(3).poisson(50.0, (wl.I1, wl.I2)).astype(np.float32)
    calls = []
    for i in range(2):
        t = time.perf_counter()
        mesolf_amd.recon_theory_phase_space_LFM_with_filter(LF, H, Ht, 1, True, 1)
        calls.append(time.perf_counter() - t)
        log(f"recon drop-in call {i + 1} (1 iteration, FP32): {calls[-1]:.2f} s")
    print(json.dumps({
        "metric": "recon drop-in host preparation, dense H/Ht 525x525x15x15x51 single",
        "value": t_prep, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": t_prep * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 in, f64 sums",
        "data": "synthetic separable PSF supports of the benchmark's radii",
        "config": {"workload": "prep_cfg2_dense_525x525x15x15x51", "host_threads": os.cpu_count()},
        "recon_dropin_1_iteration_s": {"first_call": calls[0], "second_call_tables_reused": calls[1]},
    }), flush=True)
    ctx.close()


def bench_hbm_copy(args):
    """A plain device-to-device copy of 4 GiB (torch): the practical HBM
    ceiling for the rooflines, and the reference kernel of known traffic
    (4 GiB read + 4 GiB written per launch) for the HBM byte counters."""
    import torch

    n = 1 << 30  # floats
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(max(1, args.warmup)):
        b.copy_(a)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(max(1, args.steps)):
        e0.record(st)
        b.copy_(a)
        e1.record(st)
        e1.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    dt = min(times)
    nbytes = 8.0 * n
    print(json.dumps({
        "metric": "device copy bandwidth, 4 GiB fp32 (read + write)", "value": nbytes / dt / 1e9,
        "unit": "GB/s", "n_gpus": 1, "steps": len(times), "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "constant", "config": {"workload": "hbm_copy_4GiB"},
        "roofline": {"bound": "hbm", "achieved": nbytes / dt / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": nbytes / dt / HBM_PEAK, "traffic": nbytes},
    }), flush=True)


def bench_summary(args):
    """Summary image + rank-1 background (summary_image_generation_module.m,
    'pixelwise' detrend, defaults 100 / 20 / 3 / 5) of a 1515 x 1515-px movie
    of 400 frames resident in HBM, through mlf_summary_image_device."""
    import ctypes as C

    import torch

    import mesolf_amd
    from mesolf_amd import lib as L

    h = w = 1515
    T = 400
    npix = h * w
    ctx = mesolf_amd.Context(0)
    g = torch.Generator(device="cuda").manual_seed(5)
    movie = torch.rand((T, npix), device="cuda", generator=g, dtype=torch.float32) * 100 + 10
    img = torch.empty(npix, device="cuda", dtype=torch.float64)
    bs = torch.empty(npix, device="cuda", dtype=torch.float64)
    bt = torch.empty(T, device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    lib = L.load_library()

    def call():
        L._check(lib.mlf_summary_image_device(ctx.handle, C.c_void_p(movie.data_ptr()), h, w, T, 100, 20,
                                              3, 5, C.c_void_p(img.data_ptr()), C.c_void_p(bs.data_ptr()),
                                              C.c_void_p(bt.data_ptr())))

    for _ in range(max(1, args.warmup)):
        call()
    times = []
    for _ in range(max(1, args.steps)):
        t = time.perf_counter()
        call()  # returns when the device work is done
        times.append(time.perf_counter() - t)
    dt = min(times)
    nbytes = 4.0 * npix * T * (4 + 2 * 5)  # max, 5 x (s'Y, Y t), detrend read + write, smooth read

    # the producers in front of it: rectification of the raw stack (img_rectify_mod,
    # 14.8-px lenslet pitch) and the cut of its 2 x 2 carpet patches
    xc, yc, dx = 8.3, 7.9, 14.8
    hq, wq = mesolf_amd.rectify_size(h, w, xc, yc, 15, dx)
    rect = torch.empty(hq * wq * T, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()

    def rectify():
        L._check(lib.mlf_rectify_device(ctx.handle, C.c_void_p(movie.data_ptr()), h, w, T, xc, yc, 15, dx,
                                        C.c_void_p(rect.data_ptr())))

    rectify()
    t = time.perf_counter()
    rectify()
    t_rect = time.perf_counter() - t
    patches = mesolf_amd.carpet_patch_info(hq, wq, [2, 2])
    pbuf = torch.empty(max(p["size"][0] * p["size"][1] for p in patches) * T, device="cuda",
                       dtype=torch.float32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for pi in patches:
        mesolf_amd.crop_movie_device(ctx, rect.data_ptr(), hq, wq, T, pi, pbuf.data_ptr())
    ctx.synchronize()
    t_crop = time.perf_counter() - t
    crop_px = sum(p["size"][0] * p["size"][1] for p in patches) * T
    producers = {
        "rectify": {"ms": round(t_rect * 1e3, 3), "in": [h, w, T], "out": [hq, wq, T],
                    "GB_per_s": round(4.0 * (h * w + hq * wq) * T / t_rect / 1e9, 1)},
        "crop_2x2_carpet": {"ms": round(t_crop * 1e3, 3), "patches": len(patches),
                            "GB_per_s": round(8.0 * crop_px / t_crop / 1e9, 1)},
        "note": "call times (synchronous entry points); bytes = samples read + written x 4",
    }
    del rect, pbuf
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import mesolf_oracle as O
        sub = 20000
        sample = movie[:, :sub].double().cpu().numpy().T.copy()
        t = time.perf_counter()
        O.summary_pixelwise(sample, 100, 20, 3, 5)
        tc = (time.perf_counter() - t) * npix / sub
        cpu = {"value": T / tc, "unit": "frames/s", "cores": os.cpu_count(), "kind": "port",
               "sample": f"FP64 numpy oracle on {sub} of {npix} pixels x {T} frames, scaled by the pixel count"}
    print(json.dumps({
        "metric": "summary image + rank-1 background, 1515x1515 px x 400 frames (pixelwise detrend)",
        "value": T / dt, "unit": "frames/s", "n_gpus": 1, "steps": len(times), "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 samples, f32 detrended movie)", "data": "synthetic uniform movie",
        "config": {"workload": "summary_1515x1515x400"},
        "roofline": {"bound": "hbm", "achieved": nbytes / dt / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": nbytes / dt / HBM_PEAK, "traffic": None,
                     "note": "algorithmic bytes 56 per sample; call time incl. scratch allocation"},
        "cpu_baseline": cpu,
        "producers": producers,
    }), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--workload", default="cfg2",
                    choices=["cfg1", "cfg2", "cfg4", "cfg5", "mex", "summary", "seedlib", "prep", "hbm_copy"],
                    help="cfg4: batched movie of 12 uneven patches; cfg5: serving loop; "
                         "mex: the MEX operator at the demo shape")
    ap.add_argument("--patches", type=int, default=0, help="cfg4/cfg5: patch count (12 / 36)")
    ap.add_argument("--frames", type=int, default=64, help="cfg4: movie length")
    ap.add_argument("--timed-frames", type=int, default=6, help="cfg4: warm frames timed")
    ap.add_argument("--rl-iters", type=int, default=3, help="cfg4/cfg5: RL iterations per frame")
    ap.add_argument("--slots", type=int, default=8, help="cfg5: patches in flight")
    ap.add_argument("--hbm-budget-gb", type=float, default=48.0, help="cfg5: resident patch memory")
    ap.add_argument("--load", type=float, default=0.8, help="cfg5: offered load / capacity")
    ap.add_argument("--duration", type=float, default=40.0, help="cfg5: arrival window, s")
    ap.add_argument("--views-limit", type=int, default=0,
                    help="profiling only: visit the first N views (result flagged partial)")
    ap.add_argument("--iters", type=int, default=0, help="profiling only: RL iterations per volume")
    ap.add_argument("--cpu-views", type=int, default=2)
    ap.add_argument("--oracle-views", type=int, default=1,
                    help="view steps checked against the FP64 oracle (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP32-path volume used for the speed ratio and the error")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                          "(sharding across GPUs is designed but not implemented)"}))
        sys.exit(2)

    if args.workload == "summary":
        return bench_summary(args)
    if args.workload == "mex":
        return bench_mex(args)
    if args.workload == "seedlib":
        return bench_seedlib(args)
    if args.workload == "prep":
        return bench_prep(args)
    if args.workload == "hbm_copy":
        return bench_hbm_copy(args)
    if args.workload == "cfg4":
        args.patches = args.patches or 12
        return bench_batched(args)
    if args.workload == "cfg5":
        args.patches = args.patches or 36
        return bench_serving(args)

    import mesolf_amd
    from mesolf_amd import synth, tables

    wl = {"cfg1": synth.CFG1, "cfg2": synth.CFG2}[args.workload]
    iters = args.iters or wl.iters
    partial = bool(args.views_limit) or iters != wl.iters

    t = time.perf_counter()
    tabs = tables.synthetic_compact(wl)
    log(f"PSF tables {tabs['fwd'].nbytes / 1e9:.2f} GB in {time.perf_counter() - t:.1f}s")
    rng = np.random.default_rng(wl.seed)
    truth = synth.sphere_truth(wl, rng)

    ctx = mesolf_amd.Context(0)
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    patch = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision="fp32")
    patch.set_volume(truth)
    t = time.perf_counter()
    clean = patch.project_all()
    log(f"measurement: forward of 177 views in {time.perf_counter() - t:.1f}s")
    LF = synth.poisson_measurement(clean, rng)
    del truth, clean

    work = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=args.precision) \
        if args.precision != "fp32" else patch
    work.set_image(LF)
    tuned = None
    if not args.no_autotune and args.precision != "fp32":
        work.init_volume()
        tuned = work.autotune(views=8)
        log(f"autotuned tile shape {tuned}")

    def volume():
        work.init_volume()
        work.run(iters, views_limit=args.views_limit, filter=True)

    for w in range(args.warmup):
        volume()
        ctx.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")

    # per-kernel HIP events on every 8th view step of the timed region (each
    # event pair is a marker in the stream; timing every step costs ~2 %)
    work.set_timing(True, stride=8)
    work.kernel_times()  # reset
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        volume()
        ctx.synchronize()
        log(f"step {s + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    ktimes = work.kernel_times()
    work.set_timing(False)

    nviews = args.views_limit or psf.nviews
    vol_per_s = args.steps / elapsed
    flops_view = algorithmic_flops_per_view(wl)
    dominant = max(ktimes, key=lambda c: ktimes[c][0])
    dom_ms, dom_n = ktimes[dominant]
    avg_s = dom_ms / max(dom_n, 1) * 1e-3
    if dominant in ("forward", "backward_update"):
        roof = roofline_for(wl, args.precision, dominant, avg_s)
    else:
        nbytes = 4.0 * wl.I1 * wl.I2 * wl.Z * 4
        achieved = nbytes / avg_s
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None, "kernel": dominant,
                "avg_launch_ms": avg_s * 1e3}
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_cfg2_101x101x51.json")
    if args.workload == "cfg2" and os.path.exists(tfile):
        # measured bytes per launch of the kernel instance this run used
        tr = json.load(open(tfile))["bytes_per_launch"].get(args.precision, {}).get(dominant)
        fmt, bnt, _, _ = work.tile_shape()
        if tr and "read" not in tr:
            tr = tr.get(f"fwd_columns_per_wave={fmt}" if dominant == "forward" else f"bwd_phase_tiles={bnt}")
        if tr:
            roof["traffic"] = tr["read"] + tr["write"]
            roof["traffic_source"] = ("rocprofv3 L2-to-memory request counters (" + tr.get("method", "FETCH_SIZE+WRITE_SIZE")
                                      + "), profiles/hbm_traffic_cfg2_101x101x51.json")
    per_kernel = {c: {"total_ms": round(ms, 3), "launches": n,
                      "avg_ms": round(ms / n, 4) if n else None} for c, (ms, n) in ktimes.items()}

    # Post-reconstruction step of the reconstruction module (top-layer balance,
    # max normalisations, uint16 image, central-plane crop) on the reconstructed
    # volume: outside the volumes/s figure; device time from its event pair (the
    # launches only, not the copies to the host), after one untimed call.
    vr = wl.Z if wl.Z % 2 else wl.Z - 1
    work.postprocess(vr)
    work.set_timing(True)
    work.kernel_times()
    work.postprocess(vr)
    pp_ms = work.kernel_times()["postprocess"][0]
    work.set_timing(False)
    nvox = float(wl.I1) * wl.I2 * wl.Z
    pp_bytes = 4.0 * nvox + 8.0 * wl.I1 * wl.I2 * vr + 2.0 * nvox  # read X; write stack + image
    post = {"ms": round(pp_ms, 4), "valid_range": vr, "algorithmic_bytes": pp_bytes,
            "achieved_GB/s": round(pp_bytes / (pp_ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(pp_bytes / (pp_ms * 1e-3) / HBM_PEAK, 3)}
    log(f"postprocess {pp_ms:.3f} ms")

    compare = None
    if not args.no_compare and args.precision != "fp32" and not partial:
        Xb = work.get_volume()
        patch.set_image(LF)
        patch.init_volume()
        ctx.synchronize()
        t1 = time.perf_counter()
        patch.run(iters, filter=True)
        ctx.synchronize()
        t_fp32 = time.perf_counter() - t1
        Xf = patch.get_volume()
        err = float(np.linalg.norm((Xb.astype(np.float64) - Xf).ravel()) /
                    np.linalg.norm(Xf.astype(np.float64).ravel()))
        compare = {"fp32_path_volumes_per_s": 1.0 / t_fp32,
                   f"{args.precision}_over_fp32": vol_per_s * t_fp32,
                   "rel_l2_vs_fp32_path": err,
                   "note": ("error of this path against the plain FP32 GPU path on the same "
                            "inputs; the FP32 path is within 1e-6 of the FP64 oracle at the "
                            "33x33-lenslet, 11-plane shape (tests)")}
        log(f"fp32 path {1.0 / t_fp32:.4f} vol/s, rel L2 {err:.3e}")
        del Xb, Xf

    fp32_kernels = None
    if not args.no_compare and args.precision != "fp32" and not partial:
        # the plain FP32 path's per-kernel times over one more volume
        patch.set_image(LF)
        patch.init_volume()
        patch.set_timing(True, stride=8)
        patch.kernel_times()
        patch.run(iters, filter=True)
        ctx.synchronize()
        kt32 = patch.kernel_times()
        patch.set_timing(False)
        fp32_kernels = {}
        for c in ("forward", "backward_update"):
            ms, n = kt32[c]
            if n:
                a = ms / n * 1e-3
                fp32_kernels[c] = {"avg_ms": round(a * 1e3, 4),
                                   "algorithmic_GBs": round(algorithmic_bytes_per_view(wl, "fp32", c) / a / 1e9, 1),
                                   "design_GBs": round(design_bytes_per_view(wl, "fp32", c) / a / 1e9, 1),
                                   "TFLOPs": round(algorithmic_flops_per_view(wl) / a / 1e12, 1)}
        if compare is not None:
            compare["fp32_kernels"] = fp32_kernels

    check = None
    if args.oracle_views > 0 and not partial:
        log(f"FP64 oracle check: {args.oracle_views} view step(s) ...")
        check = oracle_check(wl, tabs, LF, psf, sorted({"fp32", args.precision}), args.oracle_views)
        log(f"oracle check {check}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(wl, tabs, LF, args.cpu_views)
        log(f"cpu baseline {cpu['value']:.3e} vol/s ({cpu['sample']})")

    out = {
        "metric": "deconvolved volumes/s, 101x101-lenslet x 51-plane patch, 8 RL iterations",
        "value": vol_per_s,
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16 (f32 accumulate, f32 update)",
                  "fp16": "f16 (f32 accumulate, f32 update)"}[args.precision],
        "data": "synthetic (seeded PSF family, sphere truth, Poisson measurement)",
        "config": {"workload": wl.name, "lenslets": [wl.n1, wl.n2], "planes": wl.Z,
                   "views": nviews, "rl_iterations": iters, "precision": args.precision,
                   "tile_shape": list(work.tile_shape()), "autotuned": tuned,
                   "partial_profile_run": partial},
        "roofline": roof,
        "kernels": per_kernel,
        "postprocess": post,
        "fp32_comparison": compare,
        "oracle_check": check,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    patch.close()
    if work is not patch:
        work.close()
    psf.close()
    ctx.close()


if __name__ == "__main__":
    main()

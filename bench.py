#!/usr/bin/env python3
"""Throughput of the VBM3D two-step filter on one MI355X.

One step = one full run_vbm3d (colour transform, step 1 hard-threshold
basic estimate, step 2 Wiener final estimate, inverse colour) over the
whole synthetic clip, inputs already resident in HBM.  Default workload:
60 RGB frames 1920x1080, sigma 20, synthetic generator (velocities up to
4 px/frame, seed 2).  The same clip at sigma 40 is also timed and
reported under `config.sigma40`.

Prints one JSON line.  `roofline` is the dominant kernel, the block
matching of the step that spends the most time in it: algorithmic FP32
operations per launch (3 per pixel comparison -- sub, mul, add -- plus 2
per candidate: divide, bias) over its HIP-event-timed average launch
duration, against MI355X's FP32 peak.  Its `traffic` (HBM bytes per
launch) comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
profiles/traffic.json when one matches the workload.  `hbm` is the whole
job's algorithmic HBM rate (20*C bytes per output pixel) against the HBM
peak.  `cpu_baseline` is the CPU oracle (an op-order-faithful C++/OpenMP
restatement of the reference path) timed on this host on a sample of the
same workload; `accuracy` runs the GPU path on that same sample with the
same reference grid and reports max |GPU - oracle| and the PSNR
difference (the metric's "within 0.01 dB of reference"); `fp64_error` is
the GPU path's per-step error against the oracle's FP64 replay on a small
clip.  `config.batched` is the batched form's speed on the 64-clip mixed
batch over the plain per-frame and per-clip GPU paths.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# thread placement of the CPU baseline's OpenMP run, recorded with it
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def pow2_floor(n: int) -> int:
    r = 1
    while r * 2 <= n:
        r *= 2
    return r


def cpu_baseline(noisy, sigma, frames, color=True):
    """Time the oracle on the first `frames` frames (this host's cores).
    Returns the JSON record and the oracle's outputs and counters, which
    the accuracy leg compares the GPU path against."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nb = pow2_floor(min(32, max(1, threads)))
    sample = np.ascontiguousarray(noisy[:frames])
    F, C, H, W = sample.shape
    p1, p2 = O.params_1(sigma), O.params_2(sigma)
    t0 = time.perf_counter()
    basic, final, _ = O.run_vbm3d(sigma, sample, p1, p2, color, nb, nb)
    dt = time.perf_counter() - t0
    stats = {f"step{st}": O.last_stats(st) for st in (1, 2)}
    rec = {"value": round(W * H * F / dt / 1e6, 4), "unit": "Mpixel/s", "cores": nb,
           "kind": "port", "nproc": os.cpu_count(), "threads": nb,
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
           "sample": f"first {F} of the frames, {W}x{H}x{C}, sigma {sigma:g}, both steps, "
                     f"OpenMP over nb={nb} reference tiles on {nb} threads (the reference's thread "
                     f"rule: the largest power of two <= min(32, OMP_NUM_THREADS)), {dt:.1f} s"}
    return rec, {"nb": nb, "basic": basic, "final": final, "stats": stats, "sample": sample}


def accuracy(api, synth, torch, ctx, clean, ref, sigma):
    """The GPU path on the CPU baseline's sample, with the same reference
    grid (nb), against the oracle's outputs: max |GPU - oracle| of the
    basic and final estimates (0-255 scale), PSNR of both against the
    clean clip and their difference, the candidate counts of both, and the
    oracle's near-decision counters."""
    dev = torch.device("cuda:0")
    sample = ref["sample"]
    F = sample.shape[0]
    z = torch.from_numpy(sample).to(dev)
    b, f = torch.empty_like(z), torch.empty_like(z)
    p1, p2 = api.initialize_parameters_1(sigma), api.initialize_parameters_2(sigma)
    api.profile(2, ctx)
    api.run_device(sigma, z, b, f, p1, p2, True, ref["nb"], ctx=ctx)
    torch.cuda.synchronize()
    cands = [api.profile_read("block_match", ctx, st)[2] for st in (1, 2)]
    api.profile(0, ctx)
    gb, gf = b.cpu().numpy(), f.cpu().numpy()
    cl = clean[:F]
    p_gpu, p_orc = synth.psnr(cl, gf), synth.psnr(cl, ref["final"])
    st = ref["stats"]
    return {"sample": f"the CPU baseline's {F} frames, reference grid nb={ref['nb']}",
            "basic_max_abs": float(np.abs(gb - ref["basic"]).max()),
            "final_max_abs": float(np.abs(gf - ref["final"]).max()),
            "bit_identical": bool(np.array_equal(gb.view(np.uint32), ref["basic"].view(np.uint32))
                                  and np.array_equal(gf.view(np.uint32), ref["final"].view(np.uint32))),
            "psnr_gpu_db": round(p_gpu, 4), "psnr_oracle_db": round(p_orc, 4),
            "dpsnr_db": p_gpu - p_orc,
            "candidates_gpu": [int(c) for c in cands],
            "candidates_oracle": [st["step1"]["candidates"], st["step2"]["candidates"]],
            "near_decisions": {k: {n: v[n] for n in ("refs", "near_tie_refs", "near_tau_refs",
                                                     "near_threshold_coefs", "threshold_coefs")}
                               for k, v in st.items()},
            "near_rule": "a ranking gap, |distance - tau| or ||coefficient| - threshold| within 1e-5 "
                         "(relative) in the oracle's FP32 run"}


def fp64_error(api, synth):
    """Per-step error of the GPU path against the oracle's FP64 replay of
    the same step (same groups and threshold decisions, every operation
    in double) on the 176x144x10 grey sigma-20 clip (seed 0): max |GPU -
    FP64| of the basic estimate (step 1) and of the final estimate (step
    2, both fed the GPU's basic estimate), on the 0-255 scale."""
    from oracle import oracle as O
    _, noisy = synth.noisy_clip(176, 144, 10, 1, 20.0, 0)
    p1, p2 = api.initialize_parameters_1(20.0), api.initialize_parameters_2(20.0)

    def ratio(n, d, fb):
        return np.where(d == 0, fb, n / np.where(d == 0, 1, d))

    n1, d1 = api.step(1, 20.0, noisy, None, p1)
    basic = ratio(n1, d1, noisy).astype(np.float32)
    m1, e1 = O.step_fp64(1, 20.0, noisy, None, p1)
    n2, d2 = api.step(2, 20.0, noisy, basic, p2)
    final = ratio(n2, d2, basic).astype(np.float32)
    m2, e2 = O.step_fp64(2, 20.0, noisy, basic, p2)
    return {"clip": "176x144x10 grey, sigma 20, seed 0",
            "step1_max_abs": float(np.abs(basic - ratio(m1, e1, noisy)).max()),
            "step2_max_abs": float(np.abs(final - ratio(m2, e2, basic)).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--channels", type=int, default=3)
    ap.add_argument("--sigma", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--vmax", type=int, default=4)
    ap.add_argument("--nb", type=int, default=1, help="reference grid (1 = single-thread grid)")
    ap.add_argument("--cpu-frames", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sigma40", action="store_true", help="skip the sigma-40 line")
    ap.add_argument("--batched", action="store_true",
                    help="instead: the 64-clip mixed batch, batched form against the per-clip path")
    ap.add_argument("--clips", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="batched: default launch configuration")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-form ratio in the default line (config.batched)")
    ap.add_argument("--no-flow", action="store_true", help="skip the TV-L1 flow line (config.flow)")
    ap.add_argument("--no-transforms", action="store_true",
                    help="skip the VALU-against-MFMA transform benchmark (config.transforms)")
    ap.add_argument("--flow-only", action="store_true",
                    help="instead: only the TV-L1 flow measurement (the config.flow object)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: one GPU per process; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard by independent clips or temporal slabs)", file=sys.stderr)
        sys.exit(2)

    import torch
    from vbm3d_amd import api, synth

    if args.batched:
        print(json.dumps(measure_batched(api, synth, torch, args)), flush=True)
        return
    if args.flow_only:
        print(json.dumps(measure_flow(api, synth, torch, args, cpu=not args.no_cpu)), flush=True)
        return

    W, H, F, C, sigma = args.width, args.height, args.frames, args.channels, args.sigma
    dev = torch.device("cuda:0")
    ctx = api.Context(0)
    stream = torch.cuda.current_stream()
    res = measure(api, synth, torch, ctx, stream, dev, args, sigma)
    if not args.no_sigma40 and sigma != 40.0:
        s40 = measure(api, synth, torch, ctx, stream, dev, args, 40.0)
        res["config"]["sigma40"] = {k: s40[k] for k in ("value", "unit", "ms_per_step")}
        res["config"]["sigma40"].update({k: s40["config"][k] for k in ("psnr_noisy_db", "psnr_final_db")})
        res["config"]["sigma40"]["bm_roofline_frac"] = s40["roofline"]["frac"]
    if not args.no_cpu:
        print(f"[bench] cpu baseline on {args.cpu_frames} frames", file=sys.stderr, flush=True)
        clean, noisy = synth.noisy_clip(W, H, min(args.cpu_frames, F), C, sigma, args.seed, vmax=args.vmax)
        res["cpu_baseline"], ref = cpu_baseline(noisy, sigma, min(args.cpu_frames, F))
        res["accuracy"] = accuracy(api, synth, torch, ctx, clean, ref, sigma)
        res["fp64_error"] = fp64_error(api, synth)
        if ref["nb"] != args.nb:
            # the headline clip again with the CPU run's reference grid
            print(f"[bench] GPU with the CPU's grid nb={ref['nb']}", file=sys.stderr, flush=True)
            same = measure(api, synth, torch, ctx, stream, dev, args, sigma, nb=ref["nb"], light=True)
            res["config"]["gpu_at_cpu_nb"] = {"nb": ref["nb"], "value": same["value"], "unit": "Mpixel/s",
                                              "ms_per_step": same["ms_per_step"]}
    if not args.no_batched:
        print("[bench] batched form on the 64-clip mixed batch", file=sys.stderr, flush=True)
        bargs = argparse.Namespace(**vars(args))
        bargs.steps = 2
        b = measure_batched(api, synth, torch, bargs, tune=False)
        res["config"]["batched"] = {k: b["config"][k] for k in ("workload", "per_frame_ms", "per_clip_ms",
                                                                "batched_ms", "over_per_clip", "bit_identical")}
        res["config"]["batched"]["over_per_frame"] = b["value"]
    if not args.no_transforms:
        print("[bench] DCT transforms, VALU against MFMA", file=sys.stderr, flush=True)
        res["config"]["transforms"] = measure_transforms(api)
    if not args.no_flow:
        print("[bench] TV-L1 flow", file=sys.stderr, flush=True)
        res["config"]["flow"] = measure_flow(api, synth, torch, args, cpu=not args.no_cpu)
    print(json.dumps(res), flush=True)


def measure_transforms(api, npatch=1 << 20, reps=64):
    """The group filter's k x k DCT (k = 7 for step 2, 8 for step 1) on
    the vector ALUs, as the denoiser computes it, against FP32 MFMA
    (v_mfma_f32_4x4x1f32, k padded to 8): 2D transforms per second with
    the patches on chip, and each form's largest round-trip error."""
    out = {}
    for K in (7, 8):
        row = {}
        for mode, name in ((0, "valu"), (1, "mfma")):
            ms, rate, err = api.bench_transforms(K, mode, npatch, reps)
            row[name] = {"ms": round(ms, 3), "transforms_per_s": float(f"{rate:.4g}"),
                         "round_trip_max_err": float(f"{err:.3g}")}
        row["mfma_over_valu"] = round(row["mfma"]["transforms_per_s"] / row["valu"]["transforms_per_s"], 3)
        out[f"k{K}"] = row
    out["workload"] = f"{npatch} random patches, {reps} forward + inverse 2D DCT round trips per launch"
    return out


def measure_flow(api, synth, torch, args, cpu=True):
    """The TV-L1 flow (the reference's tvl1flow with its command-line
    defaults) from frame 0 to frame 1 of the headline clip, as grey
    images (.299 R + .587 G + .114 B), inputs resident in HBM: seconds per
    flow (median of 3 after a warm-up), the iterations of every level and
    warp, and the algorithmic HBM rate of the iterations (16 floats read
    and 6 written per pixel and iteration).  cpu: the oracle (sequential,
    bit-identical to the reference's library) on a 640x360 crop of the
    same pair."""
    W, H = args.width, args.height
    _, noisy = synth.noisy_clip(W, H, 2, args.channels, args.sigma, args.seed, vmax=args.vmax)
    if noisy.shape[1] == 3:
        grey = (0.299 * noisy[:, 0].astype(np.float64) + 0.587 * noisy[:, 1] + 0.114 * noisy[:, 2]).astype(np.float32)
    else:
        grey = noisy[:, 0].copy()
    p = api.tvl1_params()
    p.nscales = api.tvl1_nscales(W, H, p.zfactor, p.nscales)
    d0 = torch.from_numpy(grey[0]).to("cuda:0")
    d1 = torch.from_numpy(grey[1]).to("cuda:0")
    du, dv = torch.empty_like(d0), torch.empty_like(d0)
    torch.cuda.synchronize()
    ctx = api.Context(0)
    iters = api.tvl1_flow_device(d0, d1, du, dv, p, ctx=ctx)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        iters = api.tvl1_flow_device(d0, d1, du, dv, p, ctx=ctx)  # returns after the flow is done
        times.append(time.perf_counter() - t0)
    ctx.close()
    dt = sorted(times)[1]
    # level sizes as the pyramid makes them (tvl1flow/zoom.c:24-36)
    sizes = [(W, H)]
    for _ in range(1, p.nscales):
        w, h = sizes[-1]
        sizes.append((int(float(np.float32(w) * np.float32(p.zfactor)) + 0.5),
                      int(float(np.float32(h) * np.float32(p.zfactor)) + 0.5)))
    solved = list(range(p.nscales - 1, p.fscale - 1, -1))
    it_bytes = sum(88.0 * sizes[lv][0] * sizes[lv][1] * int(iters[k].sum()) for k, lv in enumerate(solved))
    out = {"workload": f"TV-L1 flow, {W}x{H} grey pair (frames 0 and 1 of the headline clip), "
                       f"tau 0.25 lambda 0.15 theta 0.3, {p.nscales} levels, zfactor 0.5, 5 warps, epsilon 0.01",
           "ms": round(dt * 1e3, 3), "mpix_s": round(W * H / dt / 1e6, 3),
           "iterations_per_level": [int(x) for x in iters.sum(axis=1)],
           "iteration_bytes": it_bytes, "iteration_gbps": round(it_bytes / dt / 1e9, 1)}
    if cpu:
        from oracle import tvl1 as T
        cw, ch = min(640, W), min(360, H)
        x0, y0 = (W - cw) // 2, (H - ch) // 2
        a, b = grey[0, y0:y0 + ch, x0:x0 + cw], grey[1, y0:y0 + ch, x0:x0 + cw]
        t0 = time.perf_counter()
        T.flow(a, b)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cw * ch / tc / 1e6, 4), "unit": "Mpixel/s", "cores": 1,
                               "kind": "port", "ms": round(tc * 1e3, 1),
                               "sample": f"{cw}x{ch} centre crop of the same pair, the oracle (sequential C++, "
                                         "bit-identical to the reference's tvl1flow_lib.c)"}
    return out


def measure_batched(api, synth, torch, args, tune=None):
    """64 clips (seeds 100..163, synth.batch_spec): the plain per-frame GPU
    path (vbm3d_run_device per clip, one launch per kernel and frame) and
    the per-clip path (one launch per kernel and step for the whole clip)
    against the batched form (autotuned, graph-captured).  Every path
    restores each clip's noisy input before a run; the outputs must be
    bit-identical."""
    dev = torch.device("cuda:0")
    ctx = api.Context(0)
    ctx_frame = api.Context(0)
    ctx_frame.set_chunk_frames(1)
    specs, pristine = [], []
    for seed in range(100, 100 + args.clips):
        W, H, F, C, sigma = synth.batch_spec(seed)
        _, noisy = synth.noisy_clip(W, H, F, C, sigma, seed)
        specs.append((W, H, F, C, sigma))
        pristine.append(torch.from_numpy(noisy).to(dev))
    print(f"[bench] {len(specs)} clips generated", file=sys.stderr, flush=True)
    pix = sum(W * H * F for (W, H, F, C, s) in specs)
    items = [api.BatchItem(s, p.clone()) for (W, H, F, C, s), p in zip(specs, pristine)]
    ref = [(torch.empty_like(p), torch.empty_like(p)) for p in pristine]
    work = [p.clone() for p in pristine]

    def per_clip(c=ctx, out=ref):
        for i, ((W, H, F, C, s), p) in enumerate(zip(specs, pristine)):
            work[i].copy_(p)
            api.run_device(s, work[i], out[i][0], out[i][1], api.initialize_parameters_1(s),
                           api.initialize_parameters_2(s), True, 1, ctx=c)

    frame_out = [(torch.empty_like(p), torch.empty_like(p)) for p in pristine]
    per_clip(ctx_frame, frame_out)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per_clip(ctx_frame, frame_out)
    torch.cuda.synchronize()
    t_frame = (time.perf_counter() - t0) / args.steps
    print(f"[bench] per-frame path {t_frame * 1e3:.1f} ms", file=sys.stderr, flush=True)

    def batched():
        for it, p in zip(items, pristine):
            it.noisy.copy_(p)
        api.run_batch(items, 1, None, ctx=ctx)

    per_clip()  # warm-up (allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per_clip()
    torch.cuda.synchronize()
    t_clip = (time.perf_counter() - t0) / args.steps
    print(f"[bench] per-clip path {t_clip * 1e3:.1f} ms", file=sys.stderr, flush=True)
    tune = (not args.no_tune) if tune is None else tune
    tuned = api.autotune_batch(items, ctx=ctx) if tune else None
    print(f"[bench] autotuned {tuned}", file=sys.stderr, flush=True)
    batched()  # plans and captures the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    same = all(torch.equal(it.basic, r[0]) and torch.equal(it.denoised, r[1]) and torch.equal(it.noisy, w)
               and torch.equal(fr[0], r[0]) and torch.equal(fr[1], r[1])
               for it, r, w, fr in zip(items, ref, work, frame_out))
    return {"metric": "mixed-clip batch: batched form over the plain per-frame GPU path",
            "value": round(t_frame / t_batch, 3), "unit": "x", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": round(t_batch * 1e3, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"{len(specs)} clips, seeds 100..{99 + len(specs)}: 8-24 frames, grey or RGB, "
                                   "320x240..1280x720, sigma 10/20/30/50, both steps",
                       "megapixels": round(pix / 1e6, 3), "per_frame_ms": round(t_frame * 1e3, 2),
                       "per_clip_ms": round(t_clip * 1e3, 2), "batched_ms": round(t_batch * 1e3, 2),
                       "over_per_clip": round(t_clip / t_batch, 3),
                       "per_frame_mpix_s": round(pix / t_frame / 1e6, 2),
                       "per_clip_mpix_s": round(pix / t_clip / 1e6, 2),
                       "batched_mpix_s": round(pix / t_batch / 1e6, 2),
                       "autotuned": tuned, "bit_identical": bool(same)}}


def kernel_name(prm, C):
    # the third template argument is 1 for the search shape compiled with
    # fixed sizes (Ns 7, Npr 5, Nb 2), as the library's launcher picks it
    shape = 1 if (prm.Ns, prm.Npr, prm.Nb) == (7, 5, 2) else 0
    return f"k_bm<{prm.k}, {C}, {shape}>"


def load_traffic(workload: str, kernel: str):
    """Per-launch HBM bytes of `kernel` for `workload` from the committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE runs, calibrated)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    ent = tab.get(workload, {}).get(kernel)
    if not ent:
        return None
    return ent["fetch_bytes_per_launch"] + ent["write_bytes_per_launch"]


def measure(api, synth, torch, ctx, stream, dev, args, sigma, nb=None, light=False):
    """Times `args.steps` full runs of the clip at `sigma` on reference grid
    `nb` (default args.nb).  light: the timing alone (no host-buffer run)."""
    W, H, F, C = args.width, args.height, args.frames, args.channels
    nb = args.nb if nb is None else nb
    print(f"[bench] generating {F}x{C}x{H}x{W} sigma={sigma:g}", file=sys.stderr, flush=True)
    clean, noisy = synth.noisy_clip(W, H, F, C, sigma, args.seed, vmax=args.vmax)
    pristine = torch.from_numpy(noisy).to(dev)
    work = torch.empty_like(pristine)
    basic = torch.empty_like(pristine)
    final = torch.empty_like(pristine)
    p1 = api.initialize_parameters_1(sigma)
    p2 = api.initialize_parameters_2(sigma)

    def one_step():
        work.copy_(pristine)
        api.run_device(sigma, work, basic, final, p1, p2, True, nb, stream=stream, ctx=ctx)

    # the warmup (or an extra untimed step) counts the work per kernel; the
    # timed steps only bracket launches with events
    n_warm = max(1, args.warmup)
    api.profile(2, ctx)
    for i in range(n_warm):
        one_step()
        print(f"[bench] warmup {i + 1}/{n_warm} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    cands = [api.profile_read("block_match", ctx, st)[2] / n_warm for st in (1, 2)]
    pix = [api.profile_read("bm_pixels", ctx, st)[2] / n_warm for st in (1, 2)]
    api.profile(1, ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        one_step()
        print(f"[bench] step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    names = ("block_match", "group_filter", "aggregate", "colour")
    kern = {}
    for st in (1, 2):
        for k in names[:3]:
            ms, n, _ = api.profile_read(k, ctx, st)
            kern[f"step{st}_{k}"] = (ms, n)
    kern["colour"] = api.profile_read("colour", ctx)[:2]
    api.profile(0, ctx)
    # dominant kernel: the block matching of the step that spends longest in it
    dom = max((1, 2), key=lambda st: kern[f"step{st}_block_match"][0])
    bm_ms, bm_n = kern[f"step{dom}_block_match"]
    prm = p1 if dom == 1 else p2
    flops_launch = (3.0 * pix[dom - 1] + 2.0 * cands[dom - 1]) * args.steps / max(bm_n, 1)
    avg_ms = bm_ms / max(bm_n, 1)
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12

    out = final.cpu().numpy()
    # the host-buffer entry (vbm3d_run): the same run with the video, the
    # basic estimate and the result crossing PCIe; reported beside the
    # headline, never as it (the second of two calls, so buffers exist)
    host_ms = None
    if sigma == args.sigma and not light:
        h_basic = np.empty_like(noisy)
        h_final = np.empty_like(noisy)
        for i in range(2):
            h_noisy = noisy.copy()
            t1 = time.perf_counter()
            api.run_vbm3d(sigma, h_noisy, None, None, h_basic, h_final, p1, p2, True, nb, ctx=ctx)
            host_ms = (time.perf_counter() - t1) * 1e3
        print(f"[bench] host-buffer run {host_ms:.1f} ms", file=sys.stderr, flush=True)
        if not np.array_equal(h_final, out):
            raise RuntimeError("host-buffer run differs from the device-buffer run")
    psnr_final = synth.psnr(clean, out)
    psnr_noisy = synth.psnr(clean, noisy)
    mpix = W * H * F / dt / 1e6
    workload = (f"{F} {'RGB' if C == 3 else 'gray'} frames {W}x{H}, sigma={sigma:g}, both steps, "
                f"default parameters, reference grid nb={nb}")
    kname = kernel_name(prm, C)
    hbm_gbs = 20.0 * C * W * H * F / dt / 1e9
    return {
        "metric": "denoised Mpixel/s, 1080p RGB sigma=20, PSNR within 0.01 dB of reference",
        "value": round(mpix, 3), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": workload,
                   "frames": F, "width": W, "height": H, "channels": C, "sigma": sigma,
                   "seed": args.seed, "vmax": args.vmax,
                   "psnr_noisy_db": round(psnr_noisy, 3), "psnr_final_db": round(psnr_final, 3),
                   "kernel_ms_per_step": {k: round(v[0] / args.steps, 2) for k, v in kern.items()},
                   "kernel_launches_per_step": {k: v[1] // args.steps for k, v in kern.items()},
                   "host_buffers": None if host_ms is None else {
                       "entry": "vbm3d_run (pageable host buffers, PCIe included)",
                       "ms": round(host_ms, 2), "mpix_s": round(W * H * F / host_ms / 1e3, 3)}},
        "roofline": {"bound": "fp32-valu", "kernel": f"{kname} (block matching, step {dom})",
                     "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": load_traffic(workload, kname),
                     "peak_note": "MI355X FP32 vector peak (counts FMA as 2 and packed FP32); the kernel "
                                  "issues no MFMA and may use neither FMA nor packed FP32 without changing "
                                  "the reference's rounding, so 39.3 T non-packed operations/s is its "
                                  "practical ceiling",
                     "avg_launch_ms": round(avg_ms, 3),
                     "flops_per_launch": flops_launch,
                     "candidates_per_step": cands[dom - 1]},
        "hbm": {"achieved": round(hbm_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_gbs / HBM_PEAK_GBS, 6),
                "bytes_per_pixel": 20 * C},
    }


if __name__ == "__main__":
    main()

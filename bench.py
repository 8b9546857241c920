"""Benchmark of the NCS hot path on one MI355X.

Workload (the plain-NCS batched configuration): 10 synthetic frames of
2048 x 2048 float32 photo-electrons (WLC filament scenes, sCMOS camera maps,
ncs_amd.sim), alpha = 0.1, the 1 - OTF mask of NA 1.4 / 700 nm / 100 nm
pixels, L-BFGS history m = 5, at most 200 iterations, reference tiling
(16 x 16 windows, stride 14: 21,609 tiles per frame, 216,090 per step).

A step is one pass of the hot path over the batch: one persistent launch
that gathers every window from the frames in HBM, solves it and stitches the
interiors back.  Inputs are resident in HBM before timing starts; the timed
region is exactly --steps launches bracketed by torch.cuda.synchronize().

Prints ONE JSON line (rank 0) with the throughput, the kernel's roofline
figures (HIP events on the launch stream), the FP64 CPU oracle baseline on a
bounded sample, and cost parity of a sample of tiles against the oracle.

Two more workloads print their own line when asked for:
  --workload uneven  one launch over 8 x 512^2, 4 x 1024x2048, 2 x 2304^2 and
                     16 ROIs of 64-400 px per side, zero-padded 16 x 16 blocks,
                     heavy-tailed variance, alpha per frame in [0.05, 0.3],
                     autotuned schedule; device-resident and through the
                     captured pipeline (host memory in and out);
  --workload stream  200 frames of 2304^2 raw uint16 ADU (blinking filaments,
                     background drifting 10 -> 40 photons/pixel, fixed camera
                     maps) streamed frame by frame through two captured
                     pipelines (ADU conversion, learned start, solve, stitch)
                     with the host one frame ahead; sustained Mpix/s and
                     per-frame latency (launch to consumption) p50 / p99.
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "corrected Mpix/s at matched NCS cost: learned-warm vs plain L-BFGS, 1 GPU"
FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix peak
HBM_PEAK_GBS = 8000.0
N_FRAMES, H, W, ALPHA, HISTORY, MAX_IT = 10, 2048, 2048, 0.1, 5, 200


def _frame(seed):
    from ncs_amd import sim
    fr = sim.make_frame(seed, (H, W))
    return fr["data"], fr["gamma"]


def make_frames(n, pool_size, cache=None):
    """The batch's frames; with `cache` (an .npz path) they are read from it
    when present and written to it otherwise (profiler runs read the cache and
    start no worker processes)."""
    if cache and os.path.exists(cache):
        with numpy.load(cache) as z:
            return [(z["d%d" % i], z["g%d" % i]) for i in range(n)]
    frames = _make_frames(n, pool_size)
    if cache:
        numpy.savez(cache, **{k: v for i, (d, g) in enumerate(frames)
                              for k, v in (("d%d" % i, d), ("g%d" % i, g))})
    return frames


def _make_frames(n, pool_size):
    if pool_size > 1:
        from multiprocessing import get_context
        with get_context("fork").Pool(pool_size) as pool:
            return pool.map(_frame, range(1, n + 1))
    return [_frame(s) for s in range(1, n + 1)]


def profiled(key, kernel, path):
    """A per-launch figure of `kernel` run on `path` (plain, tuned, linear)
    for this workload from the committed rocprofv3 summaries, or None:
    "traffic_bytes_per_launch" (FETCH_SIZE and WRITE_SIZE passes calibrated
    on a device copy) or "mfma_instructions_per_launch" (SQ_INSTS_MFMA)."""
    import re
    d = os.path.join(ROOT, "profiles")
    pat = re.compile(r"^%s (\S+) (\S+): (\d+)$" % re.escape(key))
    for fn in (sorted(os.listdir(d)) if os.path.isdir(d) else []):
        with open(os.path.join(d, fn)) as fh:
            for line in fh:
                m = pat.match(line.strip())
                if m and m.group(1) == kernel and m.group(2) == path:
                    return int(m.group(3))
    return None


def flops_per_tile(iters, trials, m, warm=False):
    """Algorithmic floating-point work per tile, as defined for this path by
    the DFT-matrix formulation of the cost (per evaluation: the forward
    U = F X F^T of a real tile, 6 real 16^3 GEMMs = 49,152 flop, its adjoint
    49,152 flop, and 17 flop per pixel for the noise sum, spectrum, log,
    likelihood gradient and combination: F_eval = 102,656), plus 2 flop per
    pixel per line-search trial and (13 + 8m) flop per pixel per L-BFGS
    iteration (norms, dots, two-loop recursion):

        F_tile = F_eval (1 + T) + 512 T + (13 + 8m) 256 I

    with I iterations and T trials.  The solver executes less than this (the
    NC Hessian is diagonal in its real-DFT basis: one basis change of 4 MFMA
    16x16x32 f16 per evaluation, two per iteration); the executed MFMA
    count is reported separately from the profiler."""
    f_eval = 102656.0
    # a warm start scores both starts: one more evaluation
    return (f_eval * (1.0 + trials + (1.0 if warm else 0.0)) + 512.0 * trials
            + (13.0 + 8.0 * m) * 256.0 * iters)


def run_gpu(frames_np, args, weights, lin_coef=None):
    import torch
    from ncs_amd import _lib, masks

    dev = torch.device("cuda:0")
    ctx = _lib.context(0)
    mask = masks.otf_mask()
    ctx.set_otf_mask(masks.to_fft_order(mask).astype(numpy.float32))
    d = [torch.from_numpy(f[0]).to(dev) for f in frames_np]
    g = [torch.from_numpy(f[1]).to(dev) for f in frames_np]
    out = [torch.zeros_like(x) for x in d]
    ntile_frame = _lib.tile_count(H, W)
    total = ntile_frame * len(d)
    iters = torch.zeros(total, dtype=torch.int32, device=dev)
    status = torch.zeros_like(iters)
    cost = torch.zeros(total, dtype=torch.float32, device=dev)
    desc = [dict(data=d[i].data_ptr(), gamma=g[i].data_ptr(), out=out[i].data_ptr(),
                 height=H, width=W, alpha=ALPHA, geometry=_lib.GEOM_REFERENCE)
            for i in range(len(d))]
    stream = torch.cuda.current_stream()
    results = {}

    def timed(params, label, steps, warmup):
        # The timed launches return the image, iterations and status, as
        # the reference's reduceNoise returns the image; the per-tile costs
        # the parity checks use come from one more launch after the timing
        # (a tile's result does not depend on the launch).
        for _ in range(warmup):
            ctx.reduce_frames_device(desc, params, None, iters.data_ptr(), status.data_ptr(),
                                     None, stream.cuda_stream)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record(stream)
            ctx.reduce_frames_device(desc, params, None, iters.data_ptr(), status.data_ptr(),
                                     None, stream.cuda_stream)
            e1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        kms = [e0.elapsed_time(e1) for e0, e1 in ev]
        ctx.reduce_frames_device(desc, params, None, iters.data_ptr(), status.data_ptr(),
                                 cost.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        it = iters.cpu().numpy()
        st = status.cpu().numpy()
        results[label] = dict(wall_s=wall, kernel_ms=float(numpy.mean(kms)),
                              iters=it, status=st, cost=cost.cpu().numpy(),
                              out=[o.cpu().numpy() for o in out] if label == "plain" else None,
                              steps=steps)
        print("[bench] %s: %.2f ms/step (kernel %.2f ms), iterations mean %.2f max %d, "
              "status %s" % (label, wall * 1e3, numpy.mean(kms), it.mean(), it.max(),
                             dict(zip(*[x.tolist() for x in numpy.unique(st, return_counts=True)]))),
              file=sys.stderr, flush=True)

    # schedule: grid and workgroup size tuned on the whole batch (history
    # stays m = 5, the workload's)
    sched = None
    if args.schedule_cache and os.path.exists(args.schedule_cache):
        with open(args.schedule_cache) as fh:
            sched = json.load(fh)
    if sched is None:
        tune = ctx.autotune(desc, _lib.default_params(history=HISTORY, max_iterations=MAX_IT),
                            tune_history=False)
    else:
        tune = sched["m5"]
    print("[bench] autotune: %s" % tune, file=sys.stderr, flush=True)
    plain = _lib.default_params(history=HISTORY, max_iterations=MAX_IT,
                                waves_per_cu=tune["waves_per_cu"],
                                waves_per_group=tune["waves_per_group"])
    timed(plain, "plain", args.steps, args.warmup)
    results["tune"] = tune
    # liblbfgs's scalar initial inverse Hessian s.y / y.y (ncs.cl:1123)
    scalar = _lib.default_params(history=HISTORY, max_iterations=MAX_IT, precond=0)
    timed(scalar, "plain_scalar_h0", max(1, min(args.steps, 3)), 1)
    # the autotuner also picks the history depth (history=0 below)
    if sched is None:
        tune_h = ctx.autotune(desc, _lib.default_params(max_iterations=MAX_IT),
                              tune_history=True)
    else:
        tune_h = sched["tuned"]
    print("[bench] autotune with depth: %s" % tune_h, file=sys.stderr, flush=True)
    results["tune_history"] = tune_h
    timed(_lib.default_params(history=tune_h["history"], max_iterations=MAX_IT,
                              waves_per_cu=tune_h["waves_per_cu"],
                              waves_per_group=tune_h["waves_per_group"]),
          "tuned", args.steps, args.warmup)
    tune_l = None
    if lin_coef is not None:
        # learned linear start, fused into the solve kernel, with a schedule
        # tuned for it (its tiles take fewer iterations)
        ctx.set_linear_start(lin_coef)
        if sched is None or "linear" not in sched:
            tune_l = ctx.autotune(desc, _lib.default_params(max_iterations=MAX_IT,
                                                            start=_lib.START_LINEAR),
                                  tune_history=True)
        else:
            tune_l = sched["linear"]
        print("[bench] autotune with the linear start: %s" % tune_l, file=sys.stderr, flush=True)
        results["tune_linear"] = tune_l
        timed(_lib.default_params(history=tune_l["history"], max_iterations=MAX_IT,
                                  waves_per_cu=tune_l["waves_per_cu"],
                                  waves_per_group=tune_l["waves_per_group"],
                                  start=_lib.START_LINEAR),
              "linear", args.steps, args.warmup)
    if sched is None and args.schedule_cache:
        with open(args.schedule_cache, "w") as fh:
            json.dump({"m5": tune, "tuned": tune_h, "linear": tune_l}, fh)
    if weights is not None:
        from ncs_amd import warmstart
        warmstart.set_context_weights(weights)
        learned = _lib.default_params(history=tune_h["history"], max_iterations=MAX_IT,
                                      waves_per_cu=tune_h["waves_per_cu"],
                                      waves_per_group=tune_h["waves_per_group"],
                                      start=_lib.START_LEARNED)
        timed(learned, "learned", args.steps, args.warmup)
        # the net alone, and the fallback decision on the first frame
        u0 = torch.empty((ntile_frame * len(d), 256), dtype=torch.float32, device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ctx.predict_frames_device(desc, u0.data_ptr(), stream.cuda_stream)
        e0.record(stream)
        ctx.predict_frames_device(desc, u0.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        results["learned"]["net_ms"] = e0.elapsed_time(e1)
        from ncs_amd import warmstart as ws
        dwin = ws.reference_windows(frames_np[0][0])
        gwin = ws.reference_windows(frames_np[0][1])
        u0f = u0[:ntile_frame].view(-1, 16, 16).cpu().numpy()
        c_net, _ = ctx.eval_tiles(dwin, gwin, u0f, ALPHA)
        c_dat, _ = ctx.eval_tiles(dwin, gwin, dwin, ALPHA)
        results["learned"]["fallback_rate"] = float(numpy.mean(c_dat <= c_net))
        results["learned"]["start_gain_rel"] = float(numpy.median((c_dat - c_net) / numpy.abs(c_dat)))
    if not args.quick:
        fixed = _lib.default_params(history=HISTORY, max_iterations=MAX_IT, epsilon=0.0)
        timed(fixed, "fixed200", max(1, min(args.steps, 2)), 1)
    return results, ntile_frame, total, ctx.info()


def parity_sample(frames_np, res, ntile_frame, n_sample_frames=1):
    """Oracle vs GPU on the first frame(s): per-tile final cost and pixels."""
    import oracle
    from ncs_amd import masks
    mf = masks.to_fft_order(masks.otf_mask())
    p = oracle.default_param(m=HISTORY, max_iterations=MAX_IT)
    rel, dmax = [], 0.0
    t0 = time.perf_counter()
    for i in range(n_sample_frames):
        ref, info = oracle.reduce_noise(frames_np[i][0].astype(numpy.float64),
                                        frames_np[i][1].astype(numpy.float64), mf, ALPHA,
                                        param=p)
        gc = res["cost"][i * ntile_frame:(i + 1) * ntile_frame].astype(numpy.float64)
        rel.append(numpy.abs(gc - info["cost"]) / numpy.abs(info["cost"]))
        dmax = max(dmax, float(numpy.max(numpy.abs(res["out"][i] - ref))))
    el = time.perf_counter() - t0
    rel = numpy.concatenate(rel)
    return dict(frames=n_sample_frames, tiles=int(rel.size), max_rel_cost=float(rel.max()),
                p99_rel_cost=float(numpy.percentile(rel, 99)),
                max_abs_pixel=dmax, within_1e5=bool(rel.max() <= 1e-5)), el


def cpu_baseline(frames_np, budget_s):
    """FP64 oracle (C restatement of clib/pyCNCS/ncs.c) on all host cores,
    same parameters, whole-image path; a bounded number of frames."""
    import oracle
    from ncs_amd import masks
    mf = masks.to_fft_order(masks.otf_mask())
    p = oracle.default_param(m=HISTORY, max_iterations=MAX_IT)
    cores = min(os.cpu_count() or 1, 16)
    n, t = 0, 0.0
    while n < len(frames_np) and t < budget_s:
        t0 = time.perf_counter()
        oracle.reduce_noise(frames_np[n][0].astype(numpy.float64),
                            frames_np[n][1].astype(numpy.float64), mf, ALPHA, param=p,
                            nthreads=cores)
        t += time.perf_counter() - t0
        n += 1
    return dict(value=n * H * W / t / 1e6, unit="Mpix/s", cores=cores, kind="port",
                sample="%d frame(s) of %dx%d, reference tiling, m=%d, eps=1e-5, "
                       "More-Thuente, FP64, OpenMP over tiles" % (n, H, W, HISTORY))


# ---------------------------------------------------------------------------
# uneven batch: mixed sizes and ROIs, zero-padded blocks, per-frame alpha
# ---------------------------------------------------------------------------

def run_uneven(args):
    import torch
    from ncs_amd import _lib, masks
    from ncs_amd.workloads import block_tiles, uneven_frame, uneven_specs
    shapes, alphas, seeds = uneven_specs()
    t0 = time.perf_counter()
    from multiprocessing import get_context
    with get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        frames = pool.map(uneven_frame, list(zip(seeds, shapes)))
    print("[bench] uneven: generated %d frames in %.1f s" % (len(frames), time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    ctx = _lib.context(0)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    ctx.set_otf_mask(mf)
    G = _lib.GEOM_BLOCKS
    d = [torch.from_numpy(f[0]).to(dev) for f in frames]
    g = [torch.from_numpy(f[1]).to(dev) for f in frames]
    out = [torch.zeros_like(x) for x in d]
    desc = [dict(data=d[i].data_ptr(), gamma=g[i].data_ptr(), out=out[i].data_ptr(),
                 height=shapes[i][0], width=shapes[i][1], alpha=alphas[i], geometry=G)
            for i in range(len(d))]
    ntiles = [_lib.tile_count(h, w, G) for h, w in shapes]
    total = int(sum(ntiles))
    pix = int(sum(h * w for h, w in shapes))
    iters = torch.zeros(total, dtype=torch.int32, device=dev)
    status = torch.zeros_like(iters)
    cost = torch.zeros(total, dtype=torch.float32, device=dev)
    base = _lib.default_params(max_iterations=MAX_IT)
    tune = ctx.autotune(desc, base, tune_history=True)
    print("[bench] uneven autotune: %s" % tune, file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()

    def timed(p):
        """Timed launches return image, iterations and status; one more
        launch after the timing writes the per-tile costs."""
        def launch(cost_ptr=None):
            ctx.reduce_frames_device(desc, p, None, iters.data_ptr(), status.data_ptr(),
                                     cost_ptr, stream.cuda_stream)
        for _ in range(args.warmup):
            launch()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record(stream)
            launch()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        kms = float(numpy.mean([a.elapsed_time(b) for a, b in ev]))
        launch(cost.data_ptr())
        torch.cuda.synchronize()
        return (wall, kms, iters.cpu().numpy(), status.cpu().numpy(),
                cost.cpu().numpy().astype(numpy.float64), [o.cpu().numpy() for o in out])

    # the tuned schedule, explicitly: later autotuning replaces the
    # context's cached one
    p = _lib.default_params(max_iterations=MAX_IT, history=tune["history"],
                            waves_per_cu=tune["waves_per_cu"],
                            waves_per_group=tune["waves_per_group"])
    wall, kms, it, st, gc, outs = timed(p)
    # the learned linear start (fitted at alpha 0.1; the per-tile fallback
    # keeps the plain start wherever it scores lower), own schedule
    lin, lwall = None, None
    if not args.no_linear:
        from ncs_amd import linstart
        if os.path.exists(linstart.DEFAULT_COEF):
            ctx.set_linear_start(linstart.load())
            pl_base = _lib.default_params(max_iterations=MAX_IT, start=_lib.START_LINEAR)
            tune_l = ctx.autotune(desc, pl_base, tune_history=True)
            pl_p = _lib.default_params(max_iterations=MAX_IT, start=_lib.START_LINEAR,
                                       history=tune_l["history"],
                                       waves_per_cu=tune_l["waves_per_cu"],
                                       waves_per_group=tune_l["waves_per_group"])
            lwall, lkms, lit, lst, lgc, _ = timed(pl_p)
            lrel = numpy.abs(lgc - gc) / numpy.abs(gc)
            lin = {"value": round(pix / lwall / 1e6, 2), "unit": "Mpix/s",
                   "ms_per_step": round(lwall * 1e3, 3), "kernel_ms": round(lkms, 3),
                   "autotune": tune_l, "iterations_mean": round(float(lit.mean()), 3),
                   "cost_vs_plain_max_rel": float(lrel.max()),
                   "matched_1e5": bool(lrel.max() <= 1e-5),
                   "status": {str(k): int(v) for k, v in zip(*numpy.unique(lst, return_counts=True))}}
            print("[bench] uneven linear: %s" % lin, file=sys.stderr, flush=True)

    # the same batch through the captured pipeline, host memory in and out
    pl = _lib.Pipeline(ctx, shapes, alphas, params=p, geometry=G)
    hin, hout = [], []
    for i, (h, w) in enumerate(shapes):
        a = torch.empty((h, w), dtype=torch.float32, pin_memory=True)
        a.numpy()[...] = frames[i][0]
        b = torch.empty((h, w), dtype=torch.float32, pin_memory=True)
        hin.append(a)
        hout.append(b)
        pl.set_gamma(i, frames[i][1])
    hin_np = [a.numpy() for a in hin]
    hout_np = [b.numpy() for b in hout]
    pl.launch(hin_np, hout_np)
    pl.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pl.launch(hin_np, hout_np)
        pl.wait()
    pwall = (time.perf_counter() - t0) / args.steps
    same = all(numpy.array_equal(hout_np[i], outs[i]) for i in range(len(shapes)))
    pl.close()

    # parity: every tile of the 16 ROIs and of the first 512^2 frame against
    # the masked-likelihood oracle (same history, FP64)
    import oracle
    firsts = numpy.concatenate([[0], numpy.cumsum(ntiles)])
    rel = []
    t0 = time.perf_counter()
    for i in [0] + list(range(14, len(shapes))):
        dt, m = block_tiles(frames[i][0])
        gt, _ = block_tiles(frames[i][1])
        ref = oracle.solve_tiles_masked(dt, gt, m, mf, alphas[i],
                                        param=oracle.default_param(m=tune["history"],
                                                                   max_iterations=MAX_IT),
                                        nthreads=min(16, os.cpu_count() or 1))
        c = gc[firsts[i]:firsts[i + 1]]
        rel.append(numpy.abs(c - ref["cost"]) / numpy.abs(ref["cost"]))
    rel = numpy.concatenate(rel)
    print("[bench] uneven parity: %d tiles in %.1f s, max rel %.2e"
          % (rel.size, time.perf_counter() - t0, rel.max()), file=sys.stderr, flush=True)
    # headline: the faster path whose per-tile costs match the plain run's
    best_wall, best_path = wall, "plain"
    if lin is not None and lin["matched_1e5"] and lwall < wall:
        best_wall, best_path = lwall, "linear"
    line = {
        "metric": "corrected Mpix/s, uneven batch in one launch, 1 GPU",
        "value": round(pix / best_wall / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(best_wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (WLC filaments, sCMOS maps with 1%% lognormal(7, 0.5) variance; "
                "seeds %d-%d)" % (seeds[0], seeds[-1]),
        "config": {"workload": "uneven batch: 8x512^2, 4x1024x2048, 2x2304^2, 16 ROIs "
                               "64-400 px, zero-padded 16x16 blocks, alpha U(0.05, 0.3) "
                               "per frame, eps 1e-5, autotuned schedule",
                   "frames": len(shapes), "pixels": pix, "tiles": total, "path": best_path,
                   "rois": [list(s) for s in shapes[14:]]},
        "plain": {"value": round(pix / wall / 1e6, 2), "ms_per_step": round(wall * 1e3, 3),
                  "kernel_ms": round(kms, 3)},
        "tiles_per_s": round(total / best_wall, 1),
        "pipeline_host_to_host": {"value": round(pix / pwall / 1e6, 2), "unit": "Mpix/s",
                                  "ms_per_step": round(pwall * 1e3, 3),
                                  "identical_to_device_resident": bool(same)},
        "autotune": tune,
        "linear_start": lin,
        "iterations": {"mean": round(float(it.mean()), 3), "max": int(it.max())},
        "status": {str(k): int(v) for k, v in zip(*numpy.unique(st, return_counts=True))},
        "parity": {"tiles": int(rel.size), "max_rel_cost": float(rel.max()),
                   "p99_rel_cost": float(numpy.percentile(rel, 99)),
                   "within_1e5": bool(rel.max() <= 1e-5),
                   "oracle": "FP64 masked-likelihood oracle, same history"},
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# time series streamed through the captured pipelines
# ---------------------------------------------------------------------------

def run_stream(args, weights, lin_coef=None):
    import torch
    from ncs_amd import _lib, masks
    n = args.stream_frames
    shape = (2304, 2304)
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    from ncs_amd.workloads import stream_series
    adu, owner, offset, gain, var = stream_series(n, shape, 5000, dev)
    print("[bench] stream: rendered %d frames of %s in %.1f s" % (n, shape, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    ctx = _lib.context(0)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    ctx.set_otf_mask(mf)
    off32, gain32, var32 = (x.astype(numpy.float32) for x in (offset, gain, var))
    results = {}
    modes = [("plain", _lib.START_PLAIN)]
    if lin_coef is not None:
        ctx.set_linear_start(lin_coef)
        modes.append(("linear", _lib.START_LINEAR))
    if weights is not None:
        from ncs_amd import warmstart
        warmstart.set_context_weights(weights)
        modes.append(("learned", _lib.START_LEARNED))
    check_frames = sorted({0, n // 2, n - 1})
    depth, lag = args.stream_depth, args.stream_lag
    if depth < 1 or lag < 0 or lag >= depth * _lib.PIPELINE_REPLAYS:
        raise SystemExit("--stream-depth >= 1 and 0 <= --stream-lag < depth * %d"
                         % _lib.PIPELINE_REPLAYS)
    n_out = lag + 2  # an output buffer is reused only after its frame was consumed
    # each start's schedule is autotuned on the first frame's photons, as the
    # batch workloads tune theirs
    ph0 = (adu[0].astype(numpy.float32) - off32) / gain32
    ph0[ph0 <= 0] = numpy.float32(1e-6)
    td = torch.from_numpy(ph0).to(dev)
    tg = torch.from_numpy(var32 / (gain32 * gain32)).to(dev)
    tout = torch.zeros_like(td)
    desc0 = [dict(data=td.data_ptr(), gamma=tg.data_ptr(), out=tout.data_ptr(),
                  height=shape[0], width=shape[1], alpha=ALPHA, geometry=_lib.GEOM_REFERENCE)]
    tunes = {}
    for label, start in modes:
        if start == _lib.START_LEARNED:
            p = _lib.default_params(max_iterations=MAX_IT, start=start)
        else:
            tune = ctx.autotune(desc0, _lib.default_params(max_iterations=MAX_IT, start=start),
                                tune_history=True)
            tunes[label] = tune
            print("[bench] stream %s autotune: %s" % (label, tune), file=sys.stderr, flush=True)
            p = _lib.default_params(max_iterations=MAX_IT, start=start, history=tune["history"],
                                    waves_per_cu=tune["waves_per_cu"],
                                    waves_per_group=tune["waves_per_group"])
        pls = [_lib.Pipeline(ctx, [shape], [ALPHA], params=p, input_format=_lib.INPUT_ADU_U16,
                             keep_tile_results=True) for _ in range(depth)]
        for pl in pls:
            pl.set_camera(0, off32, gain32, var32)
        outs = [torch.empty(shape, dtype=torch.float32, pin_memory=True) for _ in range(n_out)]
        outs_np = [o.numpy() for o in outs]
        kept, costs = {}, {}
        # warm-up replay on each pipeline (graph capture and instantiation)
        for pl in pls:
            pl.launch([adu[0]], [outs_np[0]])
            pl.wait()
        lat = numpy.zeros(n)
        t_launch = numpy.zeros(n)
        seq = numpy.zeros(n, dtype=numpy.int64)

        def finish(k):
            # frame k is consumed `lag` launches after its own
            pl = pls[k % depth]
            pl.wait(seq[k])
            lat[k] = time.perf_counter() - t_launch[k]
            if k in check_frames:
                kept[k] = outs_np[k % n_out].copy()
                costs[k] = pl.tile_results(seq[k])[2]

        t0 = time.perf_counter()
        for k in range(n):
            t_launch[k] = time.perf_counter()
            seq[k] = pls[k % depth].launch([adu[k]], [outs_np[k % n_out]])
            if k >= lag:
                finish(k - lag)
        for k in range(max(0, n - lag), n):
            finish(k)
        wall = time.perf_counter() - t0
        for pl in pls:
            pl.close()
        results[label] = dict(wall=wall, lat=lat, kept=kept, costs=costs, params=p.as_dict())
        print("[bench] stream %s: %.1f Mpix/s, latency p50 %.2f ms p99 %.2f ms"
              % (label, n * shape[0] * shape[1] / wall / 1e6, numpy.median(lat) * 1e3,
                 numpy.percentile(lat, 99) * 1e3), file=sys.stderr, flush=True)
    # parity: the checked frames against the eager plain batch on the same
    # photons (host conversion in float32)
    par = {}
    # the eager batch with the plain mode's schedule: bit-exact for that mode
    p0 = _lib.default_params(**results["plain"]["params"])
    for k in check_frames:
        ph = (adu[k].astype(numpy.float32) - off32) / gain32
        ph[ph <= 0] = numpy.float32(1e-6)
        o = numpy.zeros(shape, numpy.float32)
        _, _, c = ctx.reduce_frames([ph], [var32 / (gain32 * gain32)], [o], [ALPHA],
                                    _lib.GEOM_REFERENCE, p0)
        for label in results:
            r = results[label]
            rel = numpy.abs(r["costs"][k].astype(numpy.float64) - c) / numpy.abs(c)
            e = par.setdefault(label, dict(max_rel_cost=0.0, max_abs_pixel=0.0,
                                           bit_exact_plain=True))
            e["max_rel_cost"] = max(e["max_rel_cost"], float(rel.max()))
            e["max_abs_pixel"] = max(e["max_abs_pixel"],
                                     float(numpy.max(numpy.abs(r["kept"][k] - o))))
            e["bit_exact_plain"] = e["bit_exact_plain"] and bool(numpy.array_equal(r["kept"][k], o))
    best = "plain"
    for label in results:
        if par[label]["max_rel_cost"] <= 1e-5 and results[label]["wall"] < results[best]["wall"]:
            best = label
    pix = n * shape[0] * shape[1]

    def summary(label):
        r = results[label]
        return {"value": round(pix / r["wall"] / 1e6, 2), "unit": "Mpix/s",
                "latency_ms_p50": round(float(numpy.median(r["lat"])) * 1e3, 3),
                "latency_ms_p99": round(float(numpy.percentile(r["lat"], 99)) * 1e3, 3),
                "schedule": {k: r["params"][k] for k in ("history", "waves_per_cu",
                                                         "waves_per_group")},
                "parity_vs_eager_plain": par[label]}

    line = {
        "metric": "sustained corrected Mpix/s, 2304x2304 raw-ADU time series through the "
                  "captured pipeline, 1 GPU",
        "value": summary(best)["value"], "unit": "Mpix/s", "n_gpus": 1, "steps": n,
        "warmup": 2, "ms_per_step": round(results[best]["wall"] / n * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic time series (seed 5000): blinking WLC filaments p=0.3 per 1 um "
                "segment, background 10->40 photons/px, fixed sCMOS maps, uint16 ADU",
        "config": {"workload": "%d frames of 2304x2304 uint16 ADU streamed frame by frame "
                               "through %d captured pipelines (H2D, ADU->photons, %s start, "
                               "solve+stitch, D2H), host %d frames ahead, alpha %.1f, "
                               "autotuned schedule, eps 1e-5" % (n, depth, best, lag, ALPHA),
                   "frames": n, "height": shape[0], "width": shape[1], "start": best,
                   "pipelines": depth, "frames_ahead": lag},
        "latency_ms_p50": summary(best)["latency_ms_p50"],
        "latency_ms_p99": summary(best)["latency_ms_p99"],
        "modes": {label: summary(label) for label in results},
    }
    print(json.dumps(line), flush=True)


def _scalar_summary(res, plain, pix):
    """The plain start with liblbfgs's scalar initial Hessian (precond=0)."""
    s = res.get("plain_scalar_h0")
    if s is None:
        return None
    rel = numpy.abs(s["cost"].astype(numpy.float64) - plain["cost"]) / numpy.abs(plain["cost"])
    return {"value": round(pix / s["wall_s"] / 1e6, 2), "unit": "Mpix/s",
            "ms_per_step": round(s["wall_s"] * 1e3, 3), "kernel_ms": round(s["kernel_ms"], 3),
            "iterations_mean": round(float(s["iters"].mean()), 3),
            "cost_vs_spectral_max_rel": float(rel.max()),
            "status": {str(k): int(v) for k, v in zip(*numpy.unique(s["status"],
                                                                     return_counts=True))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=N_FRAMES)
    ap.add_argument("--quick", action="store_true", help="skip the fixed-iteration run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-learned", action="store_true", help="skip the conv-net warm start")
    ap.add_argument("--no-linear", action="store_true", help="skip the learned linear start")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (PMC byte-counter calibration)")
    ap.add_argument("--workload", choices=("batch", "uneven", "stream"), default="batch")
    ap.add_argument("--stream-frames", type=int, default=200)
    ap.add_argument("--stream-depth", type=int, default=2,
                    help="captured pipelines the stream workload alternates between")
    ap.add_argument("--stream-lag", type=int, default=1,
                    help="frames the host enqueues ahead of the one it consumes")
    ap.add_argument("--frames-cache", default=None,
                    help="read the batch from this .npz if present, else write it there")
    ap.add_argument("--schedule-cache", default=None,
                    help="read the autotuned schedules from this JSON if present (no "
                         "autotuning), else write them there")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the NCS path shards by frame with no exchange "
              "during the solve, and this build runs and reports a single MI355X only "
              "(--gpus %d requested)." % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven":
        run_uneven(args)
        return
    if args.workload == "stream":
        weights = None
        if not args.no_learned:
            from ncs_amd import warmstart
            if os.path.exists(warmstart.DEFAULT_WEIGHTS):
                weights = warmstart.load()
        lin_coef = None
        if not args.no_linear:
            from ncs_amd import linstart
            if os.path.exists(linstart.DEFAULT_COEF):
                lin_coef = linstart.load()
        run_stream(args, weights, lin_coef)
        return

    # learned warm start weights (trained on the GPU by ncs_amd.warmstart);
    # process pools run before the GPU is initialised
    weights, train_windows = None, None
    if not args.no_learned:
        from ncs_amd import warmstart
        if os.path.exists(warmstart.DEFAULT_WEIGHTS):
            weights = warmstart.load()
        else:
            train_windows = warmstart.gen_windows(50_000, workers=min(16, os.cpu_count() or 1),
                                                  log=lambda *a: print("[bench]", *a,
                                                                       file=sys.stderr))
    t0 = time.perf_counter()
    frames_np = make_frames(args.frames, min(args.frames, 16, os.cpu_count() or 1),
                            args.frames_cache)
    print("[bench] generated %d frames in %.1f s" % (len(frames_np), time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    if args.copy_probe:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("[bench] copy probe: 1 GiB read + 1 GiB written", file=sys.stderr, flush=True)
        del a, b
    weights_note = "ncs_amd/weights (trained on MI355X by the package's own kernels)"
    if train_windows is not None:
        from ncs_amd import warmstart
        d, g, u = warmstart.solve_pairs(*train_windows, alpha=ALPHA,
                                        log=lambda *a: print("[bench]", *a, file=sys.stderr))
        p, _ = warmstart.train(d, g, u, steps=1500,
                               log=lambda *a: print("[bench]", *a, file=sys.stderr))
        weights = p.cpu().numpy()
        weights_note = "trained in this run (50k pairs, 1500 steps): no stored weights found"
    lin_coef, lin_note = None, None
    if not args.no_linear:
        from ncs_amd import linstart
        if os.path.exists(linstart.DEFAULT_COEF):
            lin_coef = linstart.load()
            lin_note = ("ncs_amd/weights/%s (fitted in closed form on MI355X solver pairs)"
                        % os.path.basename(linstart.DEFAULT_COEF))
    res, ntile_frame, total, devinfo = run_gpu(frames_np, args, weights, lin_coef)
    plain = res["plain"]
    pix = len(frames_np) * H * W
    mpix = pix / plain["wall_s"] / 1e6
    alg_bytes = pix * 12.0  # data + gamma in, image out (float32)
    it = plain["iters"].astype(numpy.int64)

    def roofline_of(run, m, path, warm=False):
        kms = run["kernel_ms"]
        it = run["iters"].astype(numpy.float64)
        # trials T = I: one trial per iteration, a lower bound (the first
        # trial is accepted on almost every iteration)
        flops = float(numpy.sum(flops_per_tile(it, it, m, warm)))
        achieved = flops / (kms * 1e-3) / 1e12
        kernel = "solve_kernel<%d,false,false>" % m
        return {"bound": "mfma", "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": profiled("traffic_bytes_per_launch", kernel, path),
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                                "profiles/)",
                "flop_model": "algorithmic: F_eval (1 + T) + 512 T + (13 + 8m) 256 I per tile, "
                              "F_eval = 102,656 (DFT-matrix evaluation), T = I",
                "kernel": kernel, "kernel_ms": round(kms, 3),
                # the basis changes run MFMA 16x16x32 f16 (16,384 flops each)
                "executed_mfma_tflops_f16": (
                    round(profiled("mfma_instructions_per_launch", kernel, path) * 16384.0
                          / (kms * 1e-3) / 1e12, 2)
                    if profiled("mfma_instructions_per_launch", kernel, path) else None),
                "flop_per_launch": flops,
                "hbm_gbs_algorithmic": round(alg_bytes / (kms * 1e-3) / 1e9, 2)}

    parity, _ = parity_sample(frames_np, plain, ntile_frame, 1)
    value, start_used = mpix, "plain"
    candidates = {"plain": (mpix, True, plain["wall_s"])}
    tuned_info = None
    if "tuned" in res:
        T = res["tuned"]
        tm = pix / T["wall_s"] / 1e6
        rel = numpy.abs(T["cost"].astype(numpy.float64) - plain["cost"]) / numpy.abs(plain["cost"])
        tuned_info = {"value": round(tm, 2), "unit": "Mpix/s",
                      "ms_per_step": round(T["wall_s"] * 1e3, 3),
                      "kernel_ms": round(T["kernel_ms"], 3),
                      "history": res["tune_history"]["history"],
                      "waves_per_cu": res["tune_history"]["waves_per_cu"],
                      "waves_per_group": res["tune_history"]["waves_per_group"],
                      "iterations_mean": round(float(T["iters"].mean()), 3),
                      "cost_vs_plain_max_rel": float(rel.max()),
                      "matched_1e5": bool(rel.max() <= 1e-5)}
        candidates["tuned"] = (tm, tuned_info["matched_1e5"], T["wall_s"])
    linear_info = None
    if "linear" in res:
        L = res["linear"]
        lm = pix / L["wall_s"] / 1e6
        rel = numpy.abs(L["cost"].astype(numpy.float64) - plain["cost"]) / numpy.abs(plain["cost"])
        TL = res["tune_linear"]
        linear_info = {
            "value": round(lm, 2), "unit": "Mpix/s", "ms_per_step": round(L["wall_s"] * 1e3, 3),
            "kernel_ms": round(L["kernel_ms"], 3),
            "history": TL["history"], "waves_per_cu": TL["waves_per_cu"],
            "waves_per_group": TL["waves_per_group"],
            "iterations_mean": round(float(L["iters"].mean()), 3),
            "iterations_saved_vs_plain_m5": round(float(plain["iters"].mean() - L["iters"].mean()), 3),
            "cost_vs_plain_max_rel": float(rel.max()),
            "cost_vs_plain_p99_rel": float(numpy.percentile(rel, 99)),
            "matched_1e5": bool(rel.max() <= 1e-5),
            "coefficients": lin_note,
            "status": {str(k): int(v) for k, v in zip(*numpy.unique(L["status"], return_counts=True))}}
        candidates["linear"] = (lm, linear_info["matched_1e5"], L["wall_s"])
    learned_info = None
    if "learned" in res:
        L = res["learned"]
        lm = pix / L["wall_s"] / 1e6
        rel = numpy.abs(L["cost"].astype(numpy.float64) - plain["cost"]) / numpy.abs(plain["cost"])
        matched = bool(rel.max() <= 1e-5)
        learned_info = {
            "value": round(lm, 2), "unit": "Mpix/s", "ms_per_step": round(L["wall_s"] * 1e3, 3),
            "kernel_ms_net_plus_solve": round(L["kernel_ms"], 3),
            "net_ms": round(L.get("net_ms", float("nan")), 3),
            "iterations_mean": round(float(L["iters"].mean()), 3),
            "iterations_saved": round(float(plain["iters"].mean() - L["iters"].mean()), 3),
            "fallback_rate_frame1": round(L.get("fallback_rate", float("nan")), 4),
            "median_start_cost_gain_rel": L.get("start_gain_rel"),
            "cost_vs_plain_max_rel": float(rel.max()),
            "cost_vs_plain_p99_rel": float(numpy.percentile(rel, 99)),
            "matched_1e5": matched,
            "speedup_vs_plain": round(lm / mpix, 3),
            "speedup_vs_plain_scalar_h0": (round(lm * res["plain_scalar_h0"]["wall_s"] / pix * 1e6, 3)
                                           if "plain_scalar_h0" in res else None),
            "weights": weights_note,
            "status": {str(k): int(v) for k, v in zip(*numpy.unique(L["status"], return_counts=True))}}
        candidates["learned"] = (lm, matched, L["wall_s"])
    # headline: the fastest path whose per-tile costs match the plain run's
    # (m = 5, spectral H0) within 1e-5
    for name, (v, ok, _) in candidates.items():
        if ok and v > value:
            value, start_used = v, name

    def headline_roofline():
        if start_used == "tuned":
            return roofline_of(res["tuned"], res["tune_history"]["history"], "tuned")
        if start_used == "linear":
            return roofline_of(res["linear"], res["tune_linear"]["history"], "linear", warm=True)
        # the learned conv-net path runs two kernels; its solve is reported by the plain one
        return roofline_of(plain, HISTORY, "plain")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(candidates[start_used][2] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (WLC filament scenes, sCMOS camera maps; seeds 1-%d)" % len(frames_np),
        "config": {"workload": "plain NCS batched: %d frames of %dx%d, alpha %.1f, m=%d "
                               "reference run, <=%d iterations, eps 1e-5, reference tiling; "
                               "headline path: %s"
                               % (len(frames_np), H, W, ALPHA, HISTORY, MAX_IT, start_used),
                   "frames": len(frames_np), "height": H, "width": W, "alpha": ALPHA,
                   "history": HISTORY, "max_iterations": MAX_IT, "epsilon": 1e-5,
                   "tiles": int(total), "path": start_used},
        "plain": {"value": round(mpix, 2), "unit": "Mpix/s",
                  "ms_per_step": round(plain["wall_s"] * 1e3, 3),
                  "kernel_ms": round(plain["kernel_ms"], 3),
                  "initial_hessian": "spectral (precond=1)"},
        "plain_scalar_h0": _scalar_summary(res, plain, pix),
        "tuned": tuned_info,
        "autotune": res.get("tune"),
        "autotune_history": res.get("tune_history"),
        "autotune_linear": res.get("tune_linear"),
        "linear_start": linear_info,
        "learned_warm": learned_info,
        "roofline": headline_roofline(),
        "tiles_per_s": round(total / candidates[start_used][2], 1),
        "tiles_per_s_nominal_16x16": round(len(frames_np) * (H // 16) * (W // 16)
                                           / candidates[start_used][2], 1),
        "iterations": {"mean": round(float(it.mean()), 3), "p50": float(numpy.median(it)),
                       "max": int(it.max())},
        "status": {str(k): int(v) for k, v in zip(*numpy.unique(plain["status"], return_counts=True))},
        "parity": parity,
        "device": devinfo,
    }
    if "fixed200" in res:
        f = res["fixed200"]
        line["fixed_iterations_mode"] = {
            "value": round(pix / f["wall_s"] / 1e6, 2), "unit": "Mpix/s",
            "ms_per_step": round(f["wall_s"] * 1e3, 3), "kernel_ms": round(f["kernel_ms"], 3),
            "iterations_mean": round(float(f["iters"].mean()), 2),
            "status": {str(k): int(v) for k, v in zip(*numpy.unique(f["status"], return_counts=True))}}
    if not args.no_cpu:
        cb = cpu_baseline(frames_np, args.cpu_budget)
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

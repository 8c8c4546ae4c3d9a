#!/usr/bin/env python3
"""Throughput of the phase-linking path on one MI355X.

Workload (one step = one pass of the path over the whole frame, inputs
resident in HBM): N=30 dates, 2048 x 2048 pixels, Voronoi decorrelation
scene (16 cells, gamma0 ~ U(0.05, 0.95), tau ~ U(24, 120) d, seed 0),
21x7 window, KS test (alpha 0.05), EVD, min_shp 30, bias correction,
consecutive pair maps. The timed path is tcoh_run_device (prep, SHP,
covariance, eigen) in the mixed precision form; the plain FP32 form runs
the same workload for the ratio.

Reported beside the GPU figure:
  roofline      the dominant kernel's algorithmic FLOP rate (from HIP
                events on the handle's stream over the timed steps)
                against the MI355X FP32 peak
  cpu_baseline  the CPU oracle (plain C restatement of the reference
                kernel, OpenMP over rows) on a band of the same frame
  accuracy      max/RMS |d gamma| and |d pair| of both forms against the
                FP64 oracle on a crop, and SHP-count mismatches there
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))  # the synthetic scene generators

import synth  # noqa: E402

METRIC = "Mpixel/s at N=30, 21x7 KS+EVD: mixed-prec vs plain FP32 GPU, |dγ|<=1e-4"
METRIC_EMI = "Mpixel/s at N=30, 21x7 AD+EMI, all pairs: mixed-prec vs plain FP32 GPU"
METRIC_BATCH = ("aggregate Mpixel/s of six stacks N in {8,12,20,30,45,60}, 21x7 KS/tlog + "
                "EVD, batched mixed-prec (one graph) vs per-stack sequential plain FP32 GPU")
# The six stacks of the batched workload: depth, frame, SHP test.
C4_STACKS = [(8, 2048, 3072, "ks"), (12, 2048, 2048, "tlog"), (20, 1536, 2048, "ks"),
             (30, 1024, 1024, "tlog"), (45, 768, 768, "ks"), (60, 512, 512, "tlog")]
METRIC_GEO = ("Mpixel/s at N=60, 21x7 KS+EVD, geometric phase removed: mixed-prec vs "
              "plain FP32 GPU, one GPU's 1024 x 25000 share")
# HBM bytes per k_link launch (512 rows x 2048 px, N=30) of the split form
# TRAFFIC_FORM, from the two counter passes of
# profiles/k_link_n30_2048x2048_kernel_stats.md (FETCH_SIZE KiB x 2.0 +
# WRITE_SIZE KiB x 1.0) x 1024 B.
TRAFFIC_FORM = "bf16"
TRAFFIC_N30 = (257472 * 2.0 + 600416 * 1.0) * 1024
PEAK_FP64 = 78.6e12    # MI355X FP64 vector (TFLOP/s)
PEAK_FP32 = 157.3e12   # MI355X FP32 (TFLOP/s)
PEAK_F16_MFMA = 2.5e15  # MI355X dense FP16/BF16 matrix


def mfma_per_chunk_of(n, form):
    """v_mfma_f32_16x16x32 instructions the mixed covariance executes per
    32-candidate chunk (16*16*32*2 flop each), each product as hi.hi +
    hi.lo + lo.hi: Re Re^T + Im Im^T on the upper 16 x 16 date block pairs;
    C_im as X - X^T from X = Im Re^T on all block pairs up to 32 date
    slots (BF16; FP16 where the pixel's T area holds the transposition
    tiles, N >= 29), else Im Re^T - Re Im^T on the upper pairs."""
    nbk = (16 if n <= 16 else (32 if n <= 32 else 64)) // 16
    nup = nbk * (nbk + 1) // 2
    xt = nbk <= 2 and (form == "bf16" or (nbk == 2 and n >= 29))
    return 6 * nup + (3 * nbk * nbk if xt else 6 * nup)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prepare_on_device(tcoh, torch, s, i, q, ha, hr, geometry=None, heights=None):
    """read_strip (main.c:52-127) on the GPU for one whole-frame strip, by
    the library's strip kernel (tcoh_prepare_strip_device) from the date
    planes i, q (n, rows, cols) float32 on the device, with the geometric
    phase of `geometry` at `heights` (rows, cols) float64 removed when
    given; returns the strip buffers and the kernel's device time (ms, HIP
    events on the stream it runs on, second of two calls)."""
    n, rows, cols = i.shape
    pr, pc = rows + 2 * ha, cols + 2 * hr
    slc = torch.empty((pr, pc, n, 2), dtype=torch.float32, device=i.device)
    amp = torch.empty((pr, pc, n), dtype=torch.float32, device=i.device)
    valid = torch.empty((pr, pc), dtype=torch.uint8, device=i.device)
    # a stream of its own: a null handle would mean the library handle's
    # stream, which events recorded by torch would not bracket
    st = torch.cuda.Stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        for rep in range(2):
            e0.record(st)
            be.prepare_strip_device(i.data_ptr(), q.data_ptr(), rows, cols, -ha, pr, hr,
                                    slc.data_ptr(), amp.data_ptr(), valid.data_ptr(),
                                    geometry=geometry,
                                    heights_ptr=None if heights is None else
                                    heights.data_ptr(),
                                    stream=st.cuda_stream)
            e1.record(st)
        torch.cuda.synchronize()
    return slc, amp, valid, e0.elapsed_time(e1)


def config5_inputs(tcoh, torch, s, n, rows, cols, hr):
    """The stack of the geometry workload: the Voronoi decorrelation scene
    times exp(-i phi_geo), phi_geo the flat-earth and topographic phase of
    60 dates (circular polar orbits 700 km up, baselines ~ N(0, 100 m),
    reference date 0, C band) at a smooth random DEM of 0-500 m (about 2 km
    correlation at 20 m pixels). The phase grid comes from the library's
    FP64 geometry, and the library's own strip kernel applies the forward
    rotation (the grid negated). Returns I/Q planes on the GPU, the heights
    on the GPU and the geometry struct the timed preparation removes."""
    import torch.nn.functional as F
    reg, params = synth.config_voronoi(n=n, rows=rows, cols=cols, seed=0,
                                            gamma0=(0.05, 0.95))
    stack = synth.scene_stack_torch(n, rows, cols, reg, params, seed=0)
    i = stack.real.contiguous()
    q = stack.imag.contiguous()
    del stack
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    coarse = torch.rand((1, 1, rows // 100 + 3, cols // 100 + 3), generator=gen,
                        device="cuda", dtype=torch.float64)
    dem = F.interpolate(coarse, size=(rows, cols), mode="bicubic", align_corners=True)[0, 0]
    dem = (dem - dem.min()) / (dem.max() - dem.min()) * 500.0
    heights = dem.contiguous()
    orbits, scene, ref = synth.geometry_case(n=n, rows=rows, cols=cols, ref=0, seed=5)
    grid, hs = tcoh.api.geometry_grid(orbits, scene, ref, 0.0, 500.0)
    g = tcoh.api.make_geometry(grid, hs)
    fwd = tcoh.api.make_geometry(-grid, hs)
    pc = cols + 2 * hr
    slc = torch.empty((rows, pc, n, 2), dtype=torch.float32, device="cuda")
    amp = torch.empty((rows, pc, n), dtype=torch.float32, device="cuda")
    valid = torch.empty((rows, pc), dtype=torch.uint8, device="cuda")
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        be.prepare_strip_device(i.data_ptr(), q.data_ptr(), rows, cols, 0, rows, hr,
                                slc.data_ptr(), amp.data_ptr(), valid.data_ptr(),
                                geometry=fwd, heights_ptr=heights.data_ptr())
        torch.cuda.synchronize()
    del amp, valid
    obs = slc[:, hr:hr + cols].permute(2, 3, 0, 1)
    i = obs[:, 0].contiguous()
    q = obs[:, 1].contiguous()
    del slc, obs
    torch.cuda.synchronize()
    return i, q, heights, g


def cpu_threads():
    """The CPU baseline's threads: every core this process may use, at most
    OMP_NUM_THREADS when the environment sets it (the GPU box's CPU share)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def flops_eigen(n):
    """Algorithmic FLOPs of one estimated pixel's eigen stage: Hermitian
    tridiagonal reduction (16/3 N^3) plus normalisation and fit 8N(N-1)."""
    return 16.0 / 3.0 * n ** 3 + 8.0 * n * (n - 1)


def run_form(tcoh, s, pcols, bufs, rows, cols, precision, steps, warmup, torch):
    slc, amp, valid, coh, count, pairs = bufs
    be = tcoh.Backend(s, pcols, platform=0, device=0, precision=precision)
    pr = slc.shape[0]
    args = (slc.data_ptr(), amp.data_ptr(), valid.data_ptr(), pr, rows, cols,
            coh.data_ptr(), count.data_ptr(), pairs.data_ptr())
    for _ in range(warmup):
        be.run_device(*args)
    torch.cuda.synchronize()
    be.set_timing(True)
    stage_ms = np.zeros(4)
    launches = np.zeros(4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        be.run_device(*args)
        ms, nl = be.stage_times()
        stage_ms += ms
        launches += nl
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    be.close()
    ms_step = (t1 - t0) * 1e3 / steps
    return ms_step, stage_ms / steps, launches / steps


def run_graph(tcoh, s, pcols, bufs, rows, cols, precision, steps, torch, want):
    """The same step in graph mode (one captured HIP graph replayed per
    step, tcoh_set_graph) on a stream of torch's: ms per step, and whether
    the replayed maps equal the eager ones bit for bit."""
    slc, amp, valid, coh, count, pairs = bufs
    st = torch.cuda.Stream()
    be = tcoh.Backend(s, pcols, platform=0, device=0, precision=precision)
    be.set_graph(True)
    args = (slc.data_ptr(), amp.data_ptr(), valid.data_ptr(), slc.shape[0], rows, cols,
            coh.data_ptr(), count.data_ptr(), pairs.data_ptr())
    torch.cuda.synchronize()
    for _ in range(2):  # eager + capture, then one replay
        be.run_device(*args, stream=st.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        be.run_device(*args, stream=st.cuda_stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    be.close()
    same = bool(np.array_equal(coh.cpu().numpy().view(np.uint32), want[0].view(np.uint32))
                and np.array_equal(count.cpu().numpy(), want[1]))
    return (t1 - t0) * 1e3 / steps, same


def link_bound(n, looks, est):
    """k_link's algorithmic FLOPs of a strip's estimated pixels (looks:
    their SHP counts) and its time at the bound of the split forms: the
    three split products of the covariance on the matrix cores plus the
    linking on the FP32 vector ALUs."""
    flop_cov = float(np.sum(4.0 * n * (n + 1) * looks))
    flop_link = est * flops_eigen(n)
    return flop_cov + flop_link, 3.0 * flop_cov / PEAK_F16_MFMA + flop_link / PEAK_FP32


def mfma_executed(n, looks):
    """Matrix-core FLOPs the split forms execute for these pixels: per
    pixel ceil(L/32) chunks of 32 candidates, 12 v_mfma_f32_16x16x32 per
    upper 16 x 16 date block pair."""
    nbk = (16 if n <= 16 else (32 if n <= 32 else 64)) // 16
    return float(np.sum(np.ceil(looks / 32.0))) * 12 * nbk * (nbk + 1) // 2 * 16 * 16 * 32 * 2


def timed_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def config4(args, tcoh, torch, cref, fp64):
    """The batched workload: six independent stacks (C4_STACKS; the
    quadrant scene of gamma0 {0.9, 0.6, 0.3, 0.1}, tau 48 d, seed 100 + k
    for stack k), 21x7 windows, alpha 0.05, min_shp 20, EVD, no pair maps.
    One step = every stack once. Timed: the batch (tcoh_batch_run) in the
    split FP16 form replayed as one captured graph; beside it the same
    batch eagerly, the stacks one after the other in the same form, and the
    per-stack sequential plain FP32 path the metric compares with."""
    wa, wr = 21, 7
    ha, hr = wa // 2, wr // 2
    st = torch.cuda.Stream()
    stacks = []
    t0 = time.perf_counter()
    for k, (n, rows, cols, test) in enumerate(C4_STACKS):
        s_opts = dict(window=(wa, wr), shp_test=test, alpha=0.05, min_shp=20,
                      estimator="evd", bias=False, pairs=None)
        s = tcoh.make_settings(n, **s_opts)
        reg, params = synth.quadrant_regions(rows, cols)
        stack = synth.scene_stack_torch(n, rows, cols, reg, params, seed=100 + k)
        di, dq = stack.real.contiguous(), stack.imag.contiguous()
        del stack
        slc, amp, valid, _ = prepare_on_device(tcoh, torch, s, di, dq, ha, hr)
        del di, dq
        stacks.append(dict(
            n=n, rows=rows, cols=cols, test=test, s=s, s_opts=s_opts, pcols=cols + 2 * hr,
            slc=slc, amp=amp, valid=valid,
            coh=torch.empty((rows, cols), dtype=torch.float32, device="cuda"),
            count=torch.empty((rows, cols), dtype=torch.int32, device="cuda"),
            pairs=torch.empty((1, rows, cols), dtype=torch.float32, device="cuda")))
    torch.cuda.synchronize()
    total_px = sum(c["rows"] * c["cols"] for c in stacks)
    log("six stacks (%d px) ready in %.1f s" % (total_px, time.perf_counter() - t0))

    def args_of(c):
        return (c["slc"].data_ptr(), c["amp"].data_ptr(), c["valid"].data_ptr(),
                c["slc"].shape[0], c["rows"], c["cols"], c["coh"].data_ptr(),
                c["count"].data_ptr(), c["pairs"].data_ptr())

    def snapshot():
        return [(c["coh"].cpu().numpy().view(np.uint32).copy(), c["count"].cpu().numpy())
                for c in stacks]

    def same_maps(a, b):
        return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
                   for x, y in zip(a, b))

    out_forms = {}
    per_stack = []
    maps = {}
    for form in (args.precision, "plain"):
        bes = [tcoh.Backend(c["s"], c["pcols"], platform=0, device=0, precision=form)
               for c in stacks]

        def sequential():
            for be, c in zip(bes, stacks):
                be.run_device(*args_of(c), stream=st.cuda_stream)
        steps = args.steps if form == args.precision else max(2, args.steps // 2)
        seq_ms = timed_steps(sequential, steps, args.warmup, torch)
        maps[form] = snapshot()
        res = {"sequential_ms": seq_ms, "sequential_value": total_px / (seq_ms * 1e-3) / 1e6,
               "steps": steps}
        if form == args.precision:
            # per-stack stage times (HIP events on the run's stream), alone
            for be, c in zip(bes, stacks):
                be.set_timing(True)
                be.run_device(*args_of(c), stream=st.cuda_stream)
                ms, nl = be.stage_times()
                be.set_timing(False)
                cnt = maps[form][len(per_stack)][1]
                looks = cnt[cnt >= c["s"].min_looks].astype(np.float64)
                flop, tb = link_bound(c["n"], looks, looks.size)
                per_stack.append(dict(
                    n=c["n"], frame=[c["rows"], c["cols"]], shp_test=c["test"],
                    tiles=list(be.get_tiles()), estimated_pixels=int(looks.size),
                    mean_shp=float(looks.mean()), shp_range=[int(cnt[cnt >= 0].min()),
                                                            int(cnt.max())],
                    stage_ms=dict(zip(["prep", "shp", "link"], np.round(ms[:3], 3).tolist())),
                    link_launches=int(nl[2]), link_flop=flop, link_bound_ms=tb * 1e3,
                    mfma_flop=mfma_executed(c["n"], looks)))
            with tcoh.Batch() as batch:
                items = [(be,) + args_of(c) for be, c in zip(bes, stacks)]
                res["batch_eager_ms"] = timed_steps(
                    lambda: batch.run(items, stream=st.cuda_stream), steps, args.warmup, torch)
                order, lane = batch.plan()
                res["schedule"] = {"order_n": [stacks[i]["n"] for i in order],
                                   "lane_of_n": {str(stacks[i]["n"]): lane[i]
                                                 for i in range(len(stacks))},
                                   "streams": 4}
            same_eager = same_maps(snapshot(), maps[form])
            with tcoh.Batch(graph=True) as batch:
                items = [(be,) + args_of(c) for be, c in zip(bes, stacks)]
                res["batch_graph_ms"] = timed_steps(
                    lambda: batch.run(items, stream=st.cuda_stream), steps,
                    max(2, args.warmup), torch)
            got = snapshot()
            res["batch_maps_equal_sequential"] = bool(same_eager and same_maps(got, maps[form]))
            log("%s: sequential %.2f ms, batch eager %.2f ms, batch graph %.2f ms, maps "
                "equal: %s" % (form, seq_ms, res["batch_eager_ms"], res["batch_graph_ms"],
                               res["batch_maps_equal_sequential"]))
        else:
            log("%s: sequential %.2f ms" % (form, seq_ms))
        for be in bes:
            be.close()
        out_forms[form] = res

    mix = out_forms[args.precision]
    ms = mix["batch_graph_ms"]
    plain_ms = out_forms["plain"]["sequential_ms"]
    flop = sum(p["link_flop"] for p in per_stack)
    link_s = sum(p["stage_ms"]["link"] for p in per_stack) * 1e-3
    launches = sum(p["link_launches"] for p in per_stack)
    bound_s = sum(p["link_bound_ms"] for p in per_stack) * 1e-3
    mfma = sum(p["mfma_flop"] for p in per_stack)
    roofline = {"bound": "mfma", "achieved": flop / link_s / 1e12, "peak": flop / bound_s / 1e12,
                "unit": "TFLOP/s", "frac": bound_s / link_s, "traffic": None,
                "kernel": "k_link %s over the six stacks (split-16-bit MFMA covariance + FP32 "
                          "linking), each stack's launches timed alone" % args.precision,
                "avg_launch_ms": link_s * 1e3 / max(1, launches),
                "launches_per_step": launches,
                "mfma": {"executed_flop_per_step": mfma,
                         "util_of_2.5PF_over_k_link": mfma / link_s / PEAK_F16_MFMA}}
    out = {
        "metric": METRIC_BATCH, "value": total_px / (ms * 1e-3) / 1e6, "unit": "Mpixel/s",
        "n_gpus": 1, "steps": mix["steps"], "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 hi/lo split, f32 accumulate" if args.precision == "f16"
                 else "bf16 hi/lo split, f32 accumulate",
        "data": "synthetic",
        "config": {"workload": "six stacks N=8 2048x3072 ks, N=12 2048x2048 tlog, N=20 "
                               "1536x2048 ks, N=30 1024x1024 tlog, N=45 768x768 ks, N=60 "
                               "512x512 tlog; quadrant scene gamma0 {0.9,0.6,0.3,0.1}, tau 48 d, "
                               "seeds 100..105; 21x7 window, alpha 0.05, min_shp 20, EVD",
                   "pixels_per_step": total_px, "stacks": per_stack},
        "batch": dict(mix, path="tcoh_batch_run in graph mode: the six strips forked onto 4 "
                                "streams in a longest-first schedule, captured as one graph"),
        "plain_sequential": dict(out_forms["plain"],
                                 value=total_px / (plain_ms * 1e-3) / 1e6),
        "speedup_vs_plain": plain_ms / ms,
        "roofline": roofline,
    }
    if args.crop > 0:
        crop = min(args.crop, 64)
        acc = {}
        t0 = time.perf_counter()
        for c, (coh_m, cnt_m) in zip(stacks, maps[args.precision]):
            rows, cols, n = c["rows"], c["cols"], c["n"]
            y0, x0 = rows // 2 - crop // 2, cols // 2 - crop // 2
            bc, bn, bp, bph = band_maps(tcoh, torch, c["s_opts"], n, c["slc"], c["amp"],
                                        c["valid"], c["pcols"], y0, crop, cols,
                                        args.precision)
            gm = {args.precision: (coh_m.view(np.float32).reshape(rows, cols),
                                   cnt_m.reshape(rows, cols), np.zeros((0, rows, cols)))}
            a = crop_accuracy(tcoh, cref, fp64, c["slc"].cpu().numpy(), c["amp"].cpu().numpy(),
                              c["valid"].cpu().numpy(), c["pcols"], n, c["s_opts"], y0, x0,
                              crop, gm, {args.precision: (M2C_BAND[args.precision], bp, bph)})
            f = a[args.precision]
            acc["N=%d" % n] = {k: f[k] for k in (
                "max_abs_dgamma", "rms_dgamma", "pixels_over_1e-4", "gate_1e-4",
                "max_abs_dphase_rad", "rms_dphase_rad", "shp_count_mismatches",
                "nan_mask_equal")}
            acc["N=%d" % n]["band_rerun_equals_timed_run"] = bool(
                np.array_equal(bc.view(np.uint32), coh_m.reshape(rows, cols)[y0:y0 + crop]) and
                np.array_equal(bn, cnt_m.reshape(rows, cols)[y0:y0 + crop]))
        acc["crop"] = "%d x %d at each frame's centre" % (crop, crop)
        acc["gate_1e-4"] = all(v["gate_1e-4"] for k, v in acc.items() if k.startswith("N="))
        out["accuracy"] = acc
        log("accuracy crops in %.1f s" % (time.perf_counter() - t0))
    if not args.no_cpu:
        threads = cpu_threads()
        band = args.cpu_rows or 32
        est_s, sample = 0.0, []
        for c in stacks:
            so = cref.settings(c["n"], **c["s_opts"])
            pr_b = band + wa - 1
            t0 = time.perf_counter()
            cref.run(so, c["pcols"], c["slc"][:pr_b].cpu().numpy(), c["amp"][:pr_b].cpu().numpy(),
                     c["valid"][:pr_b].cpu().numpy(), pr_b, band, c["cols"], with_mask=False,
                     nthreads=threads)
            dt = time.perf_counter() - t0
            est_s += dt * c["rows"] / band
            sample.append("N=%d %d x %d px in %.2f s" % (c["n"], band, c["cols"], dt))
        out["cpu_baseline"] = {
            "value": total_px / est_s / 1e6, "unit": "Mpixel/s", "cores": threads,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": "the first %d output rows of each stack (%s), FP32 C restatement of "
                      "the reference kernel, OpenMP over rows; the batch's CPU time is each "
                      "stack's band time scaled to its rows" % (band, "; ".join(sample))}
    print(json.dumps(out))


# Bias-threshold band of each form's m2c rounding (as tests/test_gpu_parity.py
# derives it): the oracle takes the form's own decision inside it.
M2C_BAND = {"plain": 1e-6, "f16": 1e-6, "bf16": 1e-5, "emi": 1e-10}


def pixel_conditioning(fp64, slc_c, count, mask, sd, pcols, size, k):
    """Why crop pixel k may differ: in double, its smallest |m2c| (bias
    threshold), the relative gap of T's two largest eigenvalues, and its
    smallest phasor magnitude in units of the 1e-3 floor."""
    n, wa, wr = sd["ndates"], sd["win_az"], sd["win_rg"]
    y, x = divmod(int(k), size)
    acc = fp64.mask_bits(mask[[k]], wa * wr)
    c = fp64.covariance(slc_c, acc, np.array([y * pcols + x]),
                        fp64.window_offsets(wa, wr, pcols))
    t = fp64.coherence(c, count[[k]], sd["bias"])
    w, v = np.linalg.eigh(t)
    mag = np.abs(v[0, :, -1])
    up = np.triu(np.ones((n, n), bool), 1)
    out = {"pixel": [y, x], "count": int(count[k]),
           "top_eig_gap": float((w[0, -1] - w[0, -2]) / w[0, -1]),
           "min_phasor_over_floor": float(np.min(mag) / (1e-3 * mag.max()))}
    if sd["bias"]:
        out["min_abs_m2c"] = float(np.min(np.abs(fp64.bias_m2c(c, count[[k]])[0][up])))
    return out


def band_maps(tcoh, torch, s_opts, n, slc, amp, valid, pcols, y0, size, cols, precision):
    """The crop's rows of the frame again, through tcoh_run_device_phases on
    the frame's own device buffers (a full-width band of output rows y0 ..
    y0 + size - 1) with all pairs: |T_ij| of every pair i < j and the
    linked phases of that band. Each output pixel depends on its window
    only (also in the split forms), so the band's gamma equals the frame
    run's bit for bit, which the caller checks."""
    wa = s_opts["window"][0]
    s_all = tcoh.make_settings(n, **dict(s_opts, pairs="all"))
    nt = n * (n - 1) // 2
    coh = torch.empty((size, cols), dtype=torch.float32, device="cuda")
    count = torch.empty((size, cols), dtype=torch.int32, device="cuda")
    pairs = torch.empty((nt, size, cols), dtype=torch.float32, device="cuda")
    ph = torch.empty((n, size, cols), dtype=torch.float32, device="cuda")
    off = y0 * pcols
    st = torch.cuda.current_stream()
    with tcoh.Backend(s_all, pcols, platform=0, device=0, precision=precision) as be:
        be.run_device(slc.data_ptr() + off * n * 8, amp.data_ptr() + off * n * 4,
                      valid.data_ptr() + off, size + wa - 1, size, cols, coh.data_ptr(),
                      count.data_ptr(), pairs.data_ptr(), stream=st.cuda_stream,
                      phases_ptr=ph.data_ptr())
        torch.cuda.synchronize()
    return (coh.cpu().numpy(), count.cpu().numpy(), pairs.cpu().numpy(), ph.cpu().numpy())


def crop_accuracy(tcoh, cref, fp64, slc, amp, valid, pcols, n, s_opts, y0, x0, size,
                  gpu_maps, band_of):
    """Oracle SHP counts and FP64 gamma, pairs and linked phases on a size x
    size crop, against each GPU form: the FP64 oracle takes the form's own
    zero / non-zero bias decision where |m2c| is within the form's rounding
    of the threshold (band_of[form][0]: band, all-pairs magnitudes and
    phases of the crop's rows); everything else is FP64. The FP32 C
    oracle's own gamma error is reported beside the forms'."""
    wa, wr = s_opts["window"]
    pr_c, pc_c = size + wa - 1, size + wr - 1
    slc3 = slc.reshape(-1, pcols, n * 2)[y0:y0 + pr_c, x0:x0 + pc_c]
    amp3 = amp.reshape(-1, pcols, n)[y0:y0 + pr_c, x0:x0 + pc_c]
    val3 = valid.reshape(-1, pcols)[y0:y0 + pr_c, x0:x0 + pc_c]
    so = cref.settings(n, **s_opts)
    sd = cref.settings_dict(so)
    ocoh, ocount, _, mask, _ = cref.run(so, pc_c, slc3, amp3, val3, pr_c, size, size,
                                        nthreads=min(16, os.cpu_count() or 1))
    log("accuracy: oracle SHP counts and FP32 gamma of the %d^2 crop" % size)
    s32 = np.ascontiguousarray(slc3, np.float32).reshape(-1, n, 2)
    slc_c = s32[..., 0].astype(np.float64) + 1j * s32[..., 1].astype(np.float64)
    c64_plain = fp64.run(slc_c, ocount, mask, sd, pc_c, size, size)[0]
    log("accuracy: FP64 oracle of the crop")
    d32 = np.abs(ocoh - c64_plain)
    out = {"oracle_fp32": {"max_abs_dgamma": float(np.nanmax(d32)),
                           "pixels_over_1e-4": int(np.sum(d32 > 1e-4))}}
    for form, (coh, count, pairs) in gpu_maps.items():
        band, allp, ph = band_of[form]
        cc = coh[y0:y0 + size, x0:x0 + size].reshape(-1)
        kc = count[y0:y0 + size, x0:x0 + size].reshape(-1)
        pc = pairs[:, y0:y0 + size, x0:x0 + size].reshape(pairs.shape[0], size * size).T
        adopt = allp[:, :, x0:x0 + size].reshape(allp.shape[0], -1).T.astype(np.float64)
        phc = ph[:, :, x0:x0 + size].reshape(n, -1)
        z64 = np.empty((size * size, n), complex)
        zmag = np.empty((size * size, n))
        c64, p64, near, _ = fp64.run(slc_c, ocount, mask, sd, pc_c, size, size, band=band,
                                     adopt=adopt if sd["bias"] else None, z_out=z64,
                                     zmag_out=zmag)
        log("accuracy: FP64 oracle with the %s form's bias decisions" % form)
        d = np.abs(cc - c64)
        dp = np.abs(pc - p64)
        both = np.isfinite(phc.T) & np.isfinite(z64.real)
        clear = np.isfinite(zmag) & ((zmag > 2e-3) | (zmag < 0.5e-3))
        dph = np.abs(np.angle(np.exp(1j * np.where(both, phc.T, 0.0)) *
                              np.conj(np.where(both, z64, 1.0))))[both & clear]
        worst = int(np.nanargmax(d))
        over = int(np.sum(d > 1e-4))
        out[form] = {
            "max_abs_dgamma": float(np.nanmax(d)),
            "rms_dgamma": float(np.sqrt(np.nanmean(d ** 2))),
            "pixels_over_1e-4": over,
            "gate_1e-4": bool(over == 0 and np.array_equal(np.isnan(cc), np.isnan(c64))),
            "max_abs_dpair": float(np.nanmax(dp)) if dp.size else None,
            "rms_dpair": float(np.sqrt(np.nanmean(dp ** 2))) if dp.size else None,
            "max_abs_dphase_rad": float(dph.max()) if dph.size else None,
            "rms_dphase_rad": float(np.sqrt(np.mean(dph ** 2))) if dph.size else None,
            "phases_compared": int(dph.size),
            "bias_band_m2c": band,
            "entries_in_band": int(near.sum()),
            "nan_mask_equal": bool(np.array_equal(np.isnan(cc), np.isnan(c64))),
            "shp_count_mismatches": int(np.sum(kc != ocount)),
            "worst": dict(pixel_conditioning(fp64, slc_c, ocount, mask, sd, pc_c,
                                             size, worst),
                          gpu=float(cc[worst]), fp64=float(c64[worst]),
                          oracle_fp32=float(ocoh[worst])),
        }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="bf16", choices=["f16", "bf16"],
                    help="the timed mixed form (config 2 names split BF16)")
    ap.add_argument("--no-alt", action="store_true",
                    help="skip the other split form (reported beside the headline)")
    ap.add_argument("--rows", type=int, default=None,
                    help="default 2048 (configs 2, 3), 1024 (config 5)")
    ap.add_argument("--cols", type=int, default=None,
                    help="default 2048 (configs 2, 3), 25000 (config 5)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=None,
                    help="rows of the CPU baseline band (default 384, 64, 4 for "
                         "configs 2, 3, 5)")
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph-replay rate (tcoh_set_graph)")
    ap.add_argument("--no-region", action="store_true",
                    help="skip the host-to-host rate through tcoh_run_region")
    ap.add_argument("--config", type=int, default=2, choices=[2, 3, 4, 5],
                    help="2: N=30 KS+EVD, bias, consecutive pairs, gamma0 ~ U(0.05, "
                         "0.95) (the headline); 3: N=30 AD+EMI, bias, all 435 pairs, "
                         "gamma0 ~ U(0.6, 0.99); 4: six stacks N in {8, 12, 20, 30, 45, "
                         "60} batched (KS/tlog + EVD); 5: N=60 KS+EVD, bias, flat-earth and "
                         "topographic phase removed at DEM heights, one GPU's share "
                         "(1024 x 25000) of 8192 x 25000")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device three times: a kernel of "
                         "known traffic to calibrate HBM counters against")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every output pixel is independent; rows would shard with a "
              "20-row halo). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge
    from oracle import cref, fp64

    tcoh = ge.load_package()
    if args.config == 4:
        config4(args, tcoh, torch, cref, fp64)
        return
    c5 = args.config == 5
    n = 60 if c5 else 30
    rows = args.rows or (1024 if c5 else 2048)
    cols = args.cols or (25000 if c5 else 2048)
    cpu_rows = args.cpu_rows or {2: 384, 3: 64, 5: 4}[args.config]
    wa, wr = 21, 7
    ha, hr = wa // 2, wr // 2
    if c5:
        s_opts = dict(window=(wa, wr), shp_test="ks", alpha=0.05, min_shp=20,
                      estimator="evd", bias=True, pairs=None)
        gamma0, label = (0.05, 0.95), "KS alpha 0.05, EVD, min_shp 20 (min_looks 60), " \
                                      "bias correction, no pair maps; flat-earth and " \
                                      "topographic phase of a 0-500 m DEM removed"
    elif args.config == 2:
        s_opts = dict(window=(wa, wr), shp_test="ks", alpha=0.05, min_shp=30,
                      estimator="evd", bias=True, pairs="consecutive")
        gamma0, label = (0.05, 0.95), "KS alpha 0.05, EVD, min_shp 30, bias correction, " \
                                      "consecutive pairs"
    else:
        s_opts = dict(window=(wa, wr), shp_test="ad", alpha=0.05, min_shp=20,
                      estimator="emi", bias=True, pairs="all")
        gamma0, label = (0.6, 0.99), "AD alpha 0.05, EMI (FP64 Cholesky of |T|), " \
                                     "min_shp 20, bias correction, all 435 pairs"
    s = tcoh.make_settings(n, **s_opts)

    if args.copy_probe:
        nb = args.copy_probe << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x %d bytes read + written" % nb)
        del src, dst

    t0 = time.perf_counter()
    geo, dheights = None, None
    if c5:
        di, dq, dheights, geo = config5_inputs(tcoh, torch, s, n, rows, cols, hr)
    else:
        reg, params = synth.config_voronoi(n=n, rows=rows, cols=cols, seed=0,
                                                gamma0=gamma0)
        stack = synth.scene_stack_torch(n, rows, cols, reg, params, seed=0)
        di, dq = stack.real.contiguous(), stack.imag.contiguous()
        del stack
    slc, amp, valid, prep_ms = prepare_on_device(tcoh, torch, s, di, dq, ha, hr,
                                                 geometry=geo, heights=dheights)
    host_planes = None
    if not args.no_region:
        host_planes = (di.cpu().numpy(), dq.cpu().numpy(),
                       None if dheights is None else dheights.cpu().numpy())
    del di, dq
    pcols = cols + 2 * hr
    coh = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
    count = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    pairs = torch.empty((s.npairs, rows, cols), dtype=torch.float32, device="cuda")
    bufs = (slc, amp, valid, coh, count, pairs)
    torch.cuda.synchronize()
    log("scene %dx%d N=%d ready in %.1f s" % (rows, cols, n, time.perf_counter() - t0))

    emi = s_opts["estimator"] == "emi"
    forms = {}
    maps = {}
    band_of = {}
    band_same = {}
    graph = {}
    alt = "f16" if args.precision == "bf16" else "bf16"
    order = [args.precision] + ([] if args.no_alt or emi else [alt]) + \
            ([] if args.no_plain else ["plain"])
    crop = min(args.crop, rows, cols)
    cy0 = max(0, rows // 2 - crop // 2)
    cx0 = max(0, cols // 2 - crop // 2)
    for form in order:
        steps = args.steps if form == args.precision else max(2, args.steps // 2)
        ms, st_ms, nl = run_form(tcoh, s, pcols, bufs, rows, cols, form, steps,
                                 args.warmup, torch)
        forms[form] = (ms, st_ms, nl, steps)
        maps[form] = (coh.cpu().numpy(), count.cpu().numpy(), pairs.cpu().numpy())
        log("%s: %.2f ms/step, stages prep/shp/link ms %s" %
            (form, ms, np.round(st_ms, 3).tolist()))
        if form == args.precision and not args.no_graph:
            gms, same = run_graph(tcoh, s, pcols, bufs, rows, cols, form, steps, torch,
                                  maps[form])
            graph = {"ms_per_step": gms, "value": rows * cols / (gms * 1e-3) / 1e6,
                     "maps_equal_eager": same, "steps": steps}
            log("%s graph replay: %.2f ms/step, maps equal: %s" % (form, gms, same))
        if crop > 0:
            bc, bn, bp, bph = band_maps(tcoh, torch, s_opts, n, slc, amp, valid, pcols,
                                        cy0, crop, cols, form)
            band_same[form] = bool(
                np.array_equal(bc.view(np.uint32),
                               maps[form][0][cy0:cy0 + crop].view(np.uint32)) and
                np.array_equal(bn, maps[form][1][cy0:cy0 + crop]))
            band_of[form] = (M2C_BAND["emi" if emi else form], bp, bph)

    ms, st_ms, nl, steps = forms[args.precision]
    npix = rows * cols
    value = npix / (ms * 1e-3) / 1e6
    count_h = maps[args.precision][1]
    est = int(np.sum(count_h >= s.min_looks))
    looks = count_h[count_h >= s.min_looks].astype(np.float64)

    # The dominant kernel is the fused covariance + linking kernel k_link.
    # Algorithmic FLOPs per estimated pixel (L = its SHP count):
    # covariance 4N(N+1)L, linking (16/3)N^3 + 8N(N-1), and with EMI N^3
    # FP64 for the Cholesky inverse of |T|. For the mixed form the bound is
    # the matrix-core time of the three split products of the covariance
    # plus the FP32 (and FP64) vector time of the linking; "peak" is the
    # FLOP rate at that bound, so frac = bound time / measured time.
    stage_names = ["prep", "shp", "link", "unused"]
    dom = 2
    flop_cov = float(np.sum(4.0 * n * (n + 1) * looks))
    flop_link = est * flops_eigen(n)
    flop_f64 = est * float(n) ** 3 if s_opts["estimator"] == "emi" else 0.0
    if emi:
        # EMI: covariance sums in FP64 on the vector ALUs (every form), the
        # linking in FP32, the Cholesky of |T| in FP64
        t_bound = flop_cov / PEAK_FP64 + flop_link / PEAK_FP32 + flop_f64 / PEAK_FP64
        kname = "k_link EMI (FP64 covariance sums + FP64 Cholesky + FP32 linking) " \
                "and its two queue kernels"
        bound = "valu"
    else:
        t_bound = 3.0 * flop_cov / PEAK_F16_MFMA + flop_link / PEAK_FP32
        kname = "k_link %s (split-16-bit MFMA covariance + FP32 linking)" % args.precision
        bound = "mfma"
    flop = flop_cov + flop_link + flop_f64
    avg_launch_s = st_ms[dom] * 1e-3 / max(1.0, nl[dom])
    achieved = flop / max(1.0, nl[dom]) / avg_launch_s
    peak = flop / t_bound
    # the HBM traffic was measured (rocprofv3 counters) on this workload only
    traffic_measured = (args.config == 2 and args.precision == TRAFFIC_FORM
                        and rows == 2048 and cols == 2048)
    mfma_per_chunk = mfma_per_chunk_of(n, args.precision)
    mfma = None
    if not emi:
        executed = float(np.sum(np.ceil(looks / 32.0))) * mfma_per_chunk * 16 * 16 * 32 * 2
        mfma = {"executed_flop_per_launch": executed / max(1.0, nl[dom]),
                "instructions_per_launch": executed / (16 * 16 * 32 * 2) / max(1.0, nl[dom]),
                "util_of_2.5PF_over_k_link": executed / max(1.0, nl[dom]) / avg_launch_s /
                PEAK_F16_MFMA,
                "algorithmic_covariance_share": 3.0 * flop_cov / executed,
                "note": "executed MFMA flop rate over the whole fused kernel's time "
                        "(covariance and linking); the rocprofv3 SQ_INSTS_MFMA count of "
                        "the same kernel is in profiles/"}
    roofline = {"bound": bound, "achieved": achieved / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": TRAFFIC_N30 if traffic_measured else None,
                "traffic_source": ("profiles/k_link_n30_2048x2048_kernel_stats.md "
                                   "(rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, "
                                   "calibrated on a 1 GiB copy)") if traffic_measured else None,
                "kernel": kname, "avg_launch_ms": avg_launch_s * 1e3,
                "launches_per_step": float(nl[dom])}
    if mfma and args.precision != "plain":
        roofline["mfma"] = mfma

    out = {
        "metric": {2: METRIC, 3: METRIC_EMI, 5: METRIC_GEO}[args.config],
        "value": value,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 hi/lo split, f32 accumulate" if args.precision == "f16"
                 else "bf16 hi/lo split, f32 accumulate",
        "data": "synthetic",
        "config": {"workload": "N=%d dates, %dx%d px, 21x7 window, %s; Voronoi "
                               "decorrelation scene (16 cells, gamma0 ~ U(%g, %g)) seed 0"
                               % (n, rows, cols, label, gamma0[0], gamma0[1]),
                   "estimated_pixels": est, "mean_shp": float(looks.mean()),
                   "strip_prep_ms": round(prep_ms, 3),
                   "stage_ms": dict(zip(stage_names, np.round(st_ms, 3).tolist()))},
        "roofline": roofline,
    }
    if graph:
        out["graph"] = dict(graph, path="tcoh_set_graph: the strip replayed as one "
                                        "captured HIP graph per step")
    if alt in forms:
        ams, ast, _, asteps = forms[alt]
        out[alt] = {"value": npix / (ams * 1e-3) / 1e6, "ms_per_step": ams,
                    "steps": asteps, "speedup_vs_plain": None,
                    "stage_ms": dict(zip(stage_names, np.round(ast, 3).tolist()))}
    if "plain" in forms:
        pms, pst, _, psteps = forms["plain"]
        out["plain"] = {"value": npix / (pms * 1e-3) / 1e6, "ms_per_step": pms,
                        "steps": psteps,
                        "stage_ms": dict(zip(stage_names, np.round(pst, 3).tolist()))}
        out["speedup_vs_plain"] = pms / ms
        if alt in out:
            out[alt]["speedup_vs_plain"] = pms / out[alt]["ms_per_step"]

    if host_planes is not None:
        # The whole frame from host date planes through the streaming entry
        # point (pinned staging, copies overlapping computation, strips of
        # about a million pixels): the rate a caller holding the stack in
        # host memory sees, PCIe transfers included. Not the headline value.
        hi, hq, hh = host_planes
        with tcoh.Backend(s, pcols, platform=0, device=0,
                          precision=args.precision) as be:
            reps = 1 if c5 else 2
            if not c5:
                be.run_region(hi, hq, geometry=geo, heights=hh)
            t0 = time.perf_counter()
            for _ in range(reps):
                rc, rn, rp = be.run_region(hi, hq, geometry=geo, heights=hh)
            dt = (time.perf_counter() - t0) / reps
        same = bool(np.array_equal(rn.reshape(rows, cols), maps[args.precision][1]) and
                    np.array_equal(rc.reshape(rows, cols).view(np.uint32),
                                   maps[args.precision][0].view(np.uint32)))
        out["host_to_host"] = {
            "value": npix / dt / 1e6, "unit": "Mpixel/s", "ms_per_frame": dt * 1e3,
            "path": "tcoh_run_region: I/Q planes in host memory -> maps in host memory",
            "bytes_in": int(2 * hi.nbytes + (0 if hh is None else hh.nbytes)),
            "maps_equal_device_run": same}
        log("host to host: %.1f ms per frame (%.1f Mpx/s), maps equal: %s" %
            (dt * 1e3, npix / dt / 1e6, same))
        del host_planes, hi, hq, hh

    if crop > 0:
        t0 = time.perf_counter()
        out["accuracy"] = crop_accuracy(tcoh, cref, fp64, slc.cpu().numpy(),
                                        amp.cpu().numpy(), valid.cpu().numpy(), pcols,
                                        n, s_opts, cy0, cx0, crop, maps, band_of)
        out["accuracy"]["crop"] = [cy0, cx0, crop]
        for form, same in band_same.items():
            out["accuracy"][form]["band_rerun_equals_timed_run"] = same
        log("accuracy crop in %.1f s" % (time.perf_counter() - t0))

    if not args.no_cpu:
        band = min(cpu_rows, rows)
        threads = cpu_threads()
        so = cref.settings(n, **s_opts)
        pr_b = band + wa - 1
        sl = slc[:pr_b].cpu().numpy()
        am = amp[:pr_b].cpu().numpy()
        va = valid[:pr_b].cpu().numpy()
        t0 = time.perf_counter()
        cref.run(so, pcols, sl, am, va, pr_b, band, cols, with_mask=False,
                 nthreads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": band * cols / dt / 1e6, "unit": "Mpixel/s", "cores": threads,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "threads_rule": "every core this process may run on (sched_getaffinity), "
                            "at most OMP_NUM_THREADS when that is set",
            "sample": "%d x %d px band of the same frame (%.1f s), FP32 C restatement "
                      "of the reference kernel, OpenMP over rows" % (band, cols, dt)}
        log("cpu baseline: %.4f Mpx/s on %d threads (%.1f s)"
            % (out["cpu_baseline"]["value"], threads, dt))
    print(json.dumps(out))


if __name__ == "__main__":
    main()

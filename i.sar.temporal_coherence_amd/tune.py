#!/usr/bin/env python3
"""Launch-shape tuner of libtcoh_hip.so, run on an MI355X.

For each (estimator, covariance form, stack depth) it times one whole
strip of a synthetic distributed-scatterer stack through tcoh_run_device
under candidate launch shapes (tcoh_set_tiles) and keeps the fastest:

  link_stripes the number of XCD stripes a row is cut into by the fused
               covariance + linking launch: 1 (the whole row, every XCD a
               band of rows) .. 64; the built-in 8 is one stripe per XCD
  shp_stripes  the same for the SHP launch
  band_px      output pixels of one linking launch (256 Ki .. 2 Mi)

searched one coordinate at a time (link stripe, then SHP stripe, then
band), each step from the best so far, median of `--reps` timed runs after
a warm-up; a candidate replaces the best only when faster by MARGIN. Every candidate's maps must equal the built-in shape's bit for
bit (the shapes change only the order in which tiles run). --write puts the
chosen entries in csrc/tiles.h (or --json keeps the results, and
--from-json writes their table later); rebuild the library afterwards
(only backend.o depends on the table).

  python i.sar.temporal_coherence_amd/tune.py --write
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TILES_H = os.path.join(HERE, "csrc", "tiles.h")

# (estimator, form, N, rows, cols, SHP test): the depths of the batched
# six-stack workload at its frame sizes, N=30 at the headline frame, and
# EMI at N=30 on a smaller frame
WORKLOADS = [
    ("evd", "f16", 8, 2048, 3072, "ks"),
    ("evd", "f16", 12, 2048, 2048, "tlog"),
    ("evd", "f16", 20, 1536, 2048, "ks"),
    ("evd", "f16", 30, 2048, 2048, "ks"),
    ("evd", "f16", 45, 768, 768, "ks"),
    ("evd", "f16", 60, 512, 512, "tlog"),
    ("evd", "plain", 8, 2048, 3072, "ks"),
    ("evd", "plain", 12, 2048, 2048, "tlog"),
    ("evd", "plain", 20, 1536, 2048, "ks"),
    ("evd", "plain", 30, 2048, 2048, "ks"),
    ("evd", "plain", 45, 768, 768, "ks"),
    ("evd", "plain", 60, 512, 512, "tlog"),
    ("emi", "plain", 30, 1024, 1024, "ad"),
]
BANDS = [1 << 18, 1 << 19, 1 << 20, 1 << 21]
STRIPES = [1, 2, 4, 8, 16, 32, 64]
# a candidate must beat the best so far by this fraction (run-to-run noise)
MARGIN = 0.015


def load_package():
    import importlib.util
    if "tcoh_amd" in sys.modules:
        return sys.modules["tcoh_amd"]
    spec = importlib.util.spec_from_file_location(
        "tcoh_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tcoh_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def stack_planes(torch, n, rows, cols, seed):
    """I/Q planes (n, rows, cols) on the GPU: s_k = sqrt(g) z0 e^{j k phi}
    + sqrt(1 - g) z_k with the coherence g rising from 0.1 to 0.9 across
    the columns and phi = 0.3 rad a date."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.linspace(0.1, 0.9, cols, device="cuda").expand(rows, cols)

    def cn(*shape):
        return torch.complex(torch.randn(*shape, device="cuda", generator=gen),
                             torch.randn(*shape, device="cuda", generator=gen)) / 2 ** 0.5
    z0 = cn(rows, cols)
    k = torch.arange(n, device="cuda", dtype=torch.float32).view(n, 1, 1)
    rot = torch.polar(torch.ones_like(k), 0.3 * k)
    s = torch.sqrt(g) * z0 * rot + torch.sqrt(1 - g) * cn(n, rows, cols)
    return s.real.contiguous(), s.imag.contiguous()


def tile_widths(n, cols):
    """Tiles per row of the SHP launch (64 pixels) and of the linking
    launch (a group's pixels)."""
    ppw = 64 // (8 if n <= 16 else (16 if n <= 32 else 64))
    return (cols + 63) // 64, (cols + ppw - 1) // ppw


def candidates(nbx):
    """Stripe counts that give distinct widths on a row of nbx tiles."""
    out, widths = [], set()
    for k in STRIPES:
        w = (nbx + k - 1) // k
        if k <= nbx and w not in widths:
            widths.add(w)
            out.append(k)
    return out


def tune_one(tcoh, torch, est, form, n, rows, cols, test, reps):
    s = tcoh.make_settings(n, window=(21, 7), shp_test=test, min_shp=20, estimator=est,
                           bias=True, pairs="consecutive")
    wa, wr = 21, 7
    ha, hr = wa // 2, wr // 2
    pr, pc = rows + 2 * ha, cols + 2 * hr
    i, q = stack_planes(torch, n, rows, cols, seed=n)
    slc = torch.empty((pr, pc, n, 2), dtype=torch.float32, device="cuda")
    amp = torch.empty((pr, pc, n), dtype=torch.float32, device="cuda")
    valid = torch.empty((pr, pc), dtype=torch.uint8, device="cuda")
    coh = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
    count = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    pairs = torch.empty((s.npairs, rows, cols), dtype=torch.float32, device="cuda")
    st = torch.cuda.Stream()
    with tcoh.Backend(s, pc, platform=0, device=0, precision=form) as be:
        be.prepare_strip_device(i.data_ptr(), q.data_ptr(), rows, cols, -ha, pr, hr,
                                slc.data_ptr(), amp.data_ptr(), valid.data_ptr(),
                                stream=st.cuda_stream)
        del i, q
        args = (slc.data_ptr(), amp.data_ptr(), valid.data_ptr(), pr, rows, cols,
                coh.data_ptr(), count.data_ptr(), pairs.data_ptr())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def timed(tiles):
            be.set_tiles(tiles)
            be.run_device(*args, stream=st.cuda_stream)
            ms = []
            for _ in range(reps):
                e0.record(st)
                be.run_device(*args, stream=st.cuda_stream)
                e1.record(st)
                e1.synchronize()
                ms.append(e0.elapsed_time(e1))
            maps = (coh.cpu().numpy().view(np.uint32).copy(), count.cpu().numpy().copy())
            return float(np.median(ms)), maps

        nbs, nbl = tile_widths(n, cols)
        built_in = [8, 8, 1 << 20]
        t0, want = timed(built_in)
        best, best_ms = list(built_in), t0
        tried = {tuple(built_in): t0}

        def step(k, values):
            nonlocal best, best_ms
            for v in values:
                cand = list(best)
                cand[k] = v
                if tuple(cand) in tried:
                    continue
                ms, maps = timed(cand)
                tried[tuple(cand)] = ms
                if not (np.array_equal(maps[0], want[0]) and np.array_equal(maps[1], want[1])):
                    raise SystemExit("launch shape %s changed the maps (N=%d %s %s)"
                                     % (cand, n, est, form))
                if ms < best_ms * (1.0 - MARGIN):
                    best, best_ms = cand, ms

        step(1, candidates(nbl))
        step(0, candidates(nbs))
        step(2, [b for b in BANDS if b <= 2 * rows * cols])
        be.set_tiles(None)
    # the built-in count is stored as 0
    for k in (0, 1):
        if best[k] == 8:
            best[k] = 0
    return dict(estimator=est, form=form, n=n, rows=rows, cols=cols, test=test,
                built_in_ms=t0, best_ms=best_ms, tiles=best, tried=len(tried))


def write_table(results):
    text = open(TILES_H).read()
    lines = ["    /* emi mixed  n  shp  link  band */"]
    for r in sorted(results, key=lambda r: (r["estimator"] == "emi", r["form"], r["n"])):
        emi = 1 if r["estimator"] == "emi" else 0
        mixed = 0 if (emi or r["form"] == "plain") else 1
        shp, link, band = r["tiles"]
        lines.append("    {%d, %d, %d, %d, %d, %d}, /* %s %s %dx%d: %.2f -> %.2f ms */"
                     % (emi, mixed, r["n"], shp, link, band, r["estimator"], r["form"],
                        r["rows"], r["cols"], r["built_in_ms"], r["best_ms"]))
    lines.append("    {0, 0, 0, 0, 0, 0},")
    new = re.sub(r"(TILE_TABLE\[\] = \{\n)(.*?)(\n\};)",
                 lambda m: m.group(1) + "\n".join(lines) + m.group(3), text, flags=re.S)
    open(TILES_H, "w").write(new)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--only", default="", help="comma-separated depths to tune")
    ap.add_argument("--write", action="store_true", help="write the table into tiles.h")
    ap.add_argument("--json", default="", help="also write the results here")
    ap.add_argument("--from-json", default="",
                    help="write the table of an earlier run's --json output (no GPU)")
    args = ap.parse_args()
    if args.from_json:
        with open(args.from_json) as f:
            write_table(json.load(f))
        return
    import torch
    torch.zeros(1, device="cuda")  # torch's HIP runtime first (the library shares it)
    tcoh = load_package()
    only = {int(v) for v in args.only.split(",") if v}
    results = []
    for est, form, n, rows, cols, test in WORKLOADS:
        if only and n not in only:
            continue
        t = time.perf_counter()
        r = tune_one(tcoh, torch, est, form, n, rows, cols, test, args.reps)
        results.append(r)
        print("%s %s N=%d %dx%d: built-in %.2f ms, best %.2f ms with %s (%d shapes, %.0f s)"
              % (est, form, n, rows, cols, r["built_in_ms"], r["best_ms"], r["tiles"],
                 r["tried"], time.perf_counter() - t), flush=True)
        torch.cuda.empty_cache()
    if args.json:
        with open(args.json, "w") as f:
            json.dump(results, f, indent=1)
    if args.write:
        write_table(results)


if __name__ == "__main__":
    main()

"""Benchmark of the points / line aggregation path on one MI355X.

Default workload (the headline): 100M float32 points, 90% in one 0.01-sigma Gaussian and
10% uniform, value lognormal(0, 2), Canvas 1920x1080 with the data extent as ranges,
``summary(count=count(), mean=mean('v'))``.  A step is one ``Canvas.points`` call on the
HBM-resident columns (plane creation, the aggregation kernel, finalize, result wrap).

Prints ONE JSON line: throughput in Gpoints/s, the aggregation kernel's achieved HBM
bandwidth (algorithmic bytes per launch / average launch time from HIP events on the
kernel's stream) against the 8 TB/s peak, and the CPU oracle timed on the host's cores.

Other workloads: --workload points800 | points_skewed | bycat | lines | serving | shade |
stream | rowindex.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class KernelTimer:
    """HIP events recorded on the current stream around each aggregation launch."""

    def __init__(self, torch):
        self.torch = torch
        self.pairs = []

    def start(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.torch.cuda.current_stream())
        self.pairs.append([e, None])

    def stop(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.torch.cuda.current_stream())
        self.pairs[-1][1] = e

    def mean_ms(self):
        self.torch.cuda.synchronize()
        d = [a.elapsed_time(b) for a, b in self.pairs]
        return sum(d) / len(d) if d else float("nan")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the multi-threaded CPU baseline: the process's CPU share
    (OMP_NUM_THREADS when the host sets it), else the CPUs it may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_forms(run, rows_1t, rows_pt, what, reps=5):
    """Time the CPU baseline (oracle/cpu_baseline.c) in its two forms: one thread (the
    reference's pandas path) on the first rows_1t rows, and P threads (its dask path,
    npartitions = P, private canvases, binary combine tree) on the first rows_pt rows.
    run(rows, threads) performs one whole call.  One untimed warm-up call each, then the
    median of `reps`.  Returns (rows/s at P threads, P, rows/s at 1 thread, sample text,
    host info)."""
    from oracle import baseline

    def med(rows, threads):
        run(rows, threads)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run(rows, threads)
            times.append(time.perf_counter() - t0)
        return statistics.median(times)

    P = cpu_threads()
    t1 = med(rows_1t, 1)
    tp = med(rows_pt, P)
    sample = (f"{what}: the CPU baseline (specialised C, -O3) on the first {rows_pt} rows "
              f"with {P} threads (dask form: {P} partitions, private canvases, binary "
              f"combine tree) and on the first {rows_1t} rows with 1 thread (pandas form); "
              f"median of {reps} after one warm-up")
    return rows_pt / tp, P, rows_1t / t1, sample, baseline.host_info()


def build_workload(name, torch, scale=1.0):
    """Returns (callable step, rows per step, algorithmic bytes per aggregation launch,
    config dict, cpu_baseline callable or None)."""
    import numpy as np

    import datashader_amd as dsa
    from datashader_amd import synthetic
    from datashader_amd import _device as dev

    if name == "points_skewed":
        n = int(100_000_000 * scale)
        W, H = 1920, 1080
        x, y, v = synthetic.points_skewed_dev(n, seed=2)
        frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
        xr = dev.maybe_expand_bounds(dev.bounds(x))
        yr = dev.maybe_expand_bounds(dev.bounds(y))
        canvas = dsa.Canvas(W, H, x_range=xr, y_range=yr)
        agg = dsa.summary(count=dsa.count(), mean=dsa.mean("v"))

        def step():
            return canvas.points(frame, "x", "y", agg)

        algo = 12 * n + (4 + 8) * W * H
        cfg = {"workload": "points 1920x1080 count+mean, 100M float32 rows (90% in a "
                           "0.01-sigma Gaussian, 10% uniform), value lognormal(0,2)",
               "rows": n, "canvas": [W, H], "aggs": ["count()", "mean('v')"]}

        def cpu(sample_rows, res=None):
            import oracle
            from oracle import baseline
            m1, mp = min(sample_rows, n), n
            xs, ys, vs = (t_[:mp].cpu().numpy() for t_ in (x, y, v))
            sx, tx = canvas.x_axis.compute_scale_and_translate(xr, W)
            sy, ty = canvas.y_axis.compute_scale_and_translate(yr, H)
            view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1])
            last = {}

            def run(m, th):
                last[m] = baseline.points_count_mean_f32(xs[:m], ys[:m], vs[:m], view, W, H, th)

            out = cpu_forms(run, m1, mp, "count + mean")
            if res is not None:  # the GPU call's planes against the baseline's, all rows
                ref = last[mp]
                cnt = res["count"].values
                gm, rm = res["mean"].values, ref["mean"]
                ok = ~np.isnan(rm)
                rel = float(np.max(np.abs(gm[ok] - rm[ok]) / np.abs(rm[ok]))) if ok.any() else 0.0
                out[4]["agreement"] = {
                    "count_mismatched_cells": int(np.count_nonzero(cnt != ref["count"])),
                    "mean_nan_pattern_equal": bool(np.array_equal(np.isnan(gm), np.isnan(rm))),
                    "mean_max_rel_err": rel}
            return out

        def check(res):
            c = res["count"].data
            return int(c.sum().item())

        return step, n, algo, cfg, cpu, check

    if name == "stream":
        import pandas as pd
        n = int(400_000_000 * scale)
        W, H = 1920, 1080
        chunk = 1 << 25
        x, y, v = synthetic.points_skewed_dev(n, seed=2)
        xr = dev.maybe_expand_bounds(dev.bounds(x))
        yr = dev.maybe_expand_bounds(dev.bounds(y))
        host = {"x": x.cpu().numpy(), "y": y.cpu().numpy(), "v": v.cpu().numpy()}
        del x, y, v
        torch.cuda.empty_cache()
        source = dsa.ChunkedFrame(pd.DataFrame(host), chunk_rows=chunk)
        canvas = dsa.Canvas(W, H, x_range=xr, y_range=yr)
        agg = dsa.summary(count=dsa.count(), mean=dsa.mean("v"))

        def step():
            return canvas.points(source, "x", "y", agg)

        nchunks = -(-n // chunk)
        algo = 12 * n // nchunks + (4 + 8) * W * H  # per aggregation launch (one chunk)
        cfg = {"workload": "points 1920x1080 count+mean streamed from host memory "
                           "(ChunkedFrame, 2^25-row chunks, PCIe included): 400M float32 "
                           "rows (90% in a 0.01-sigma Gaussian, 10% uniform), value "
                           "lognormal(0,2)", "rows": n, "chunk_rows": chunk,
               "canvas": [W, H], "aggs": ["count()", "mean('v')"]}

        def cpu(sample_rows, res=None):
            import oracle
            from oracle import baseline
            m1, mp = min(sample_rows, n), min(n, 100_000_000)
            xs, ys, vs = (host[k] for k in ("x", "y", "v"))
            sx, tx = canvas.x_axis.compute_scale_and_translate(xr, W)
            sy, ty = canvas.y_axis.compute_scale_and_translate(yr, H)
            view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1])
            return cpu_forms(lambda m, th: baseline.points_count_mean_f32(
                xs[:m], ys[:m], vs[:m], view, W, H, th), m1, mp, "count + mean")

        def check(res):
            return int(res["count"].values.sum(dtype=np.int64))

        return step, n, algo, cfg, cpu, check

    if name == "points800":
        n = int(10_000_000 * scale)
        W, H = 800, 600
        x, y, _ = synthetic.points_mixture_dev(n, seed=1, dtype=torch.float64)
        frame = dsa.DeviceFrame({"x": x, "y": y})
        canvas = dsa.Canvas(W, H, x_range=(-10, 10), y_range=(-10, 10))
        agg = dsa.count()

        def step():
            return canvas.points(frame, "x", "y", agg)

        algo = 16 * n + 4 * W * H
        cfg = {"workload": "points 800x600 count, 10M float64 rows (5-Gaussian mixture, "
                           "0.1% NaN rows)", "rows": n, "canvas": [W, H], "aggs": ["count()"]}

        def cpu(sample_rows, res=None):
            import oracle
            from oracle import baseline
            m1, mp = min(sample_rows, n), n
            xs, ys = x.cpu().numpy(), y.cpu().numpy()
            view = oracle.View(40.0, 400.0, 30.0, 300.0, -10, 10, -10, 10)
            return cpu_forms(lambda m, th: baseline.points_count_f64(
                xs[:m], ys[:m], view, W, H, th), m1, mp, "count")

        def check(res):
            return int(res.data.sum().item())

        return step, n, algo, cfg, cpu, check

    if name == "bycat":
        n = int(500_000_000 * scale)
        W = H = 1024
        x, y, cat = synthetic.points_ridges_dev(n, seed=3)
        cats = [f"c{k}" for k in range(16)]
        frame = dsa.DeviceFrame({"x": x, "y": y, "cat": cat}, categories={"cat": cats})
        xr = dev.maybe_expand_bounds(dev.bounds(x))
        yr = dev.maybe_expand_bounds(dev.bounds(y))
        canvas = dsa.Canvas(W, H, x_range=xr, y_range=yr)
        agg = dsa.by("cat", dsa.count())

        def step():
            return canvas.points(frame, "x", "y", agg)

        algo = 9 * n + 64 * W * H
        cfg = {"workload": "points 1024x1024 by('cat', count()), 500M float32 rows "
                           "(3 sinusoidal ridges), 16 Zipf(1.2) categories",
               "rows": n, "canvas": [W, H], "aggs": ["by('cat', count())"]}

        def cpu(sample_rows, res=None):
            import oracle
            from oracle import baseline
            m1, mp = min(sample_rows, n), min(n, 100_000_000)
            xs, ys = x[:mp].cpu().numpy(), y[:mp].cpu().numpy()
            cs = cat[:mp].cpu().numpy().astype(np.uint8)
            sx, tx = canvas.x_axis.compute_scale_and_translate(xr, W)
            sy, ty = canvas.y_axis.compute_scale_and_translate(yr, H)
            view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1])
            return cpu_forms(lambda m, th: baseline.points_count_cat_f32(
                xs[:m], ys[:m], cs[:m], view, W, H, 16, th), m1, mp, "by(cat, count())")

        def check(res):
            return int(res.data.sum().item())

        return step, n, algo, cfg, cpu, check

    if name == "lines":
        nlines = int(1_000_000 * scale)
        W, H = 2048, 1024
        xs, block = synthetic.random_walks_dev(nlines, 256, seed=4)
        names = [f"y{j}" for j in range(256)]
        frame = dsa.DeviceFrame.from_block(names, block)
        # autorange, as LinesAxis1XConstant computes it (line.py:328-331, 249-255)
        xr = dev.maybe_expand_bounds((np.nanmin(xs), np.nanmax(xs)))
        ylo, yhi = dev.bounds(block)
        yr = dev.maybe_expand_bounds((ylo, yhi))
        canvas = dsa.Canvas(W, H, x_range=xr, y_range=yr)
        agg = dsa.count()

        def step():
            return canvas.line(frame, x=xs, y=names, agg=agg, axis=1)

        algo = 4 * 256 * nlines + 8 * 256 + 4 * W * H
        cfg = {"workload": "lines 2048x1024 count, 1M random walks x 256 float32 samples, "
                           "shared x = linspace(0,1,256)", "lines": nlines, "canvas": [W, H],
               "aggs": ["count()"]}

        def cpu(sample_rows, res=None):
            import oracle
            from oracle import baseline
            # series, row-major float32 as the reference's to_numpy() gives them
            m1 = max(1, min(sample_rows // 256, nlines))
            mp = max(1, min(nlines, 8 * m1))
            ys = block[:, :mp].t().contiguous().cpu().numpy()
            sx, tx = canvas.x_axis.compute_scale_and_translate(xr, W)
            sy, ty = canvas.y_axis.compute_scale_and_translate(yr, H)
            view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1], yf32=True)
            xs64 = np.asarray(xs, np.float64)
            rate_p, P, rate_1, sample, info = cpu_forms(
                lambda m, th: baseline.lines_xconst_f32(xs64, ys[:m], view, W, H, th),
                m1, mp, "lines count (series of 256 samples)", reps=3)
            return rate_p * 256, P, rate_1 * 256, sample.replace("rows", "series"), info

        def check(res):
            return int(res.data.sum().item())

        return step, nlines * 256, algo, cfg, cpu, check

    raise SystemExit(f"unknown workload {name!r}")


def _traffic(workload):
    """HBM bytes per aggregation call from the committed counter summary (two rocprofv3
    --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return None if entry is None else entry.get("bytes_per_call")


def run_serving(args, torch):
    """Viewport serving (count + mean) over a 1B-row resident set, three forms on the same
    request sample: plain (one Canvas.points call over the whole set per request, one
    at a time), indexed (the same call over the lattice-sorted set, per request, on four
    streams) and batched (datashader_amd.batching: continuous batching, each batch one
    graph replay of the work-queue kernels, knobs from the autotuner).  Each form runs
    saturated (all requests queued at the start) and under the Poisson arrivals; prints
    one JSON line."""
    import numpy as np

    import datashader_amd as dsa
    from datashader_amd import batching, serving, synthetic

    n = int(1_000_000_000 * args.scale)
    _log(f"[bench] serving: generating {n} resident rows")
    x, y, v = synthetic.points_mixture_dev(n, seed=5, nan_frac=0.0, dtype=torch.float32,
                                           with_value=True)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    index = serving.PointIndex(x, y, v, grid=args.grid)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    _log(f"[bench] index (grid {args.grid}) built in {build_s:.2f} s")
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    pick = torch.randint(0, n, (100_000,), generator=g, device="cuda")
    centres = (x[pick].cpu().numpy(), y[pick].cpu().numpy())
    extent = ((index.x0, index.x1), (index.y0, index.y1))
    reqs = serving.make_requests(args.requests, centres, extent, seed=5, rate=200.0)
    est = batching.RowEstimator(index)
    scanned = sum(12 * est.rows(r.x_range, r.y_range) + 12 * r.width * r.height for r in reqs)

    def pct(a):
        return {"p50_ms": float(np.percentile(a, 50) * 1e3),
                "p99_ms": float(np.percentile(a, 99) * 1e3)}

    # plain: one call per request over the unsorted set, FIFO on one stream
    for r in reqs[:3]:
        serving.run_plain(frame, r)
    torch.cuda.synchronize()

    def plain(realtime):
        lat = np.zeros(len(reqs))
        s0 = time.perf_counter()
        for i, r in enumerate(reqs):
            if realtime:
                wait = r.arrival - (time.perf_counter() - s0)
                if wait > 0:
                    time.sleep(wait)
            serving.run_plain(frame, r)
            torch.cuda.synchronize()
            lat[i] = time.perf_counter() - s0 - (r.arrival if realtime else 0.0)
        return lat, time.perf_counter() - s0

    plat_sat, pwall = plain(False)
    plain_rps = len(reqs) / pwall
    _log(f"[bench] plain saturated: {plain_rps:.1f} req/s")
    plat_rt, _ = plain(True)
    _log(f"[bench] plain Poisson 200/s: p99 {np.percentile(plat_rt, 99) * 1e3:.1f} ms")

    # indexed, per request (no batching) on four streams: dsx_points_agg_indexed reads
    # lattices of up to 1024 cells a side
    index1k = index if index.grid <= 1024 else serving.PointIndex(x, y, v, grid=1024)
    vs = serving.ViewportServer(index1k, slots=4)
    for r in reqs[:3]:
        serving.run_indexed(index1k, r)
    torch.cuda.synchronize()
    isat = vs.serve(reqs, realtime=False)
    irt = vs.serve(reqs, realtime=True)
    idx_rps = len(reqs) / isat["wall"]
    _log(f"[bench] indexed per request: {idx_rps:.1f} req/s saturated")
    if index1k is not index:
        del vs, index1k
        torch.cuda.empty_cache()

    # batched: tune on a separate sample, then serve
    tsample = serving.make_requests(args.tune_requests, centres, extent, seed=55, rate=200.0)
    t1 = time.perf_counter()
    tuning = batching.BatchTuner(index).tune(tsample, max_batch=args.max_batch)
    tune_s = time.perf_counter() - t1
    _log(f"[bench] tuned in {tune_s:.1f} s: threads {tuning.threads}, workgroups "
         f"{tuning.workgroups}, {len(tuning.per_class)} classes")
    server = batching.BatchServer(index, engines=2, max_batch=args.max_batch, tuning=tuning)
    server.serve(reqs[:64], realtime=False)  # warm (allocator pools, graph upload)
    bsat = server.serve(reqs, realtime=False)
    brt = server.serve(reqs, realtime=True)
    b_rps = len(reqs) / bsat["wall"]
    _log(f"[bench] batched: {b_rps:.1f} req/s saturated ({bsat['batches']} batches, mean "
         f"{bsat['mean_batch']:.1f}); Poisson p99 {np.percentile(brt['latency'], 99) * 1e3:.1f} ms")
    server.close()

    out = {
        "metric": "viewport req/s (count + mean), batched served form vs plain per-request form",
        "value": b_rps, "unit": "req/s", "n_gpus": 1, "steps": len(reqs), "warmup": 3,
        "ms_per_step": 1e3 / b_rps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated on the device)",
        "config": {"workload": "serving: 1B resident float32 rows (5-Gaussian mixture), value "
                               "U(0,1); Poisson 200 req/s, zoom 2^U(0,12), centres at random "
                               "rows, canvases 256^2/512^2/1024x768/1920x1080",
                   "rows": n, "requests": len(reqs), "grid": index.grid,
                   "index_build_s": build_s, "max_batch": args.max_batch, "engines": 2,
                   "tuning": tuning.to_dict(), "tune_s": tune_s,
                   "tune_requests": len(tsample)},
        "plain": {"saturated_req_per_s": plain_rps, **pct(plat_sat),
                  "poisson_200": pct(plat_rt)},
        "indexed_per_request": {"saturated_req_per_s": idx_rps, **pct(isat["latency"]),
                                "poisson_200": pct(irt["latency"])},
        "batched": {"saturated_req_per_s": b_rps, **pct(bsat["latency"]),
                    "batches": bsat["batches"], "mean_batch": bsat["mean_batch"],
                    "poisson_200": {**pct(brt["latency"]), "batches": brt["batches"],
                                    "mean_batch": brt["mean_batch"]},
                    "algorithmic_bytes": scanned,
                    "achieved_gbs": scanned / bsat["wall"] / 1e9},
        "speedup_vs_plain": b_rps / plain_rps,
        "p99_poisson_batched_vs_plain": float(np.percentile(brt["latency"], 99) /
                                              np.percentile(plat_rt, 99)),
    }
    print(json.dumps(out), flush=True)


def run_shade(args, torch):
    """tf.shade of the headline's aggregates on the device (eq_hist of the 1920x1080 count
    plane, linear / log of the mean plane) and of a 1024x1024x16 categorical count plane;
    the numpy oracle on the host copies is the CPU baseline.  Prints one JSON line."""
    import numpy as np

    import datashader_amd as dsa
    import datashader_amd.transfer_functions as tf
    from datashader_amd import _device as dev
    from datashader_amd import synthetic

    n = int(100_000_000 * args.scale)
    x, y, v = synthetic.points_skewed_dev(n, seed=2)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
    xr = dev.maybe_expand_bounds(dev.bounds(x))
    yr = dev.maybe_expand_bounds(dev.bounds(y))
    cvs = dsa.Canvas(1920, 1080, x_range=xr, y_range=yr)
    aggs = cvs.points(frame, "x", "y", dsa.summary(count=dsa.count(), mean=dsa.mean("v")))
    count, mean = aggs["count"], aggs["mean"]
    nc = int(50_000_000 * args.scale)
    cx, cy, cc = synthetic.points_ridges_dev(nc, seed=3)
    cats = [f"c{k}" for k in range(16)]
    cframe = dsa.DeviceFrame({"x": cx, "y": cy, "cat": cc}, categories={"cat": cats})
    ccvs = dsa.Canvas(1024, 1024, x_range=(0, 1), y_range=(0, 1))
    byc = ccvs.points(cframe, "x", "y", dsa.by("cat", dsa.count()))
    del x, y, v, frame, cx, cy, cc, cframe
    torch.cuda.synchronize()
    cases = [("count eq_hist", count, {}, 4), ("mean linear", mean, {"how": "linear"}, 8),
             ("mean log", mean, {"how": "log"}, 8), ("by-cat eq_hist", byc, {}, 64)]
    results = {}
    for label, agg, kw, in_bytes in cases:
        for _ in range(args.warmup):
            tf.shade(agg, **kw)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(torch.cuda.current_stream())
        for _ in range(args.steps):
            img = tf.shade(agg, **kw)
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        px = agg.shape[0] * agg.shape[1]
        ms = (t1 - t0) * 1e3 / args.steps
        dev_ms = e0.elapsed_time(e1) / args.steps
        entry = {"pixels": px, "ms_per_call": ms, "device_ms_per_call": dev_ms,
                 "mpix_per_s": px / (ms * 1e-3) / 1e6,
                 "algorithmic_bytes": px * (in_bytes + 4),
                 "achieved_gbs": px * (in_bytes + 4) / (dev_ms * 1e-3) / 1e9,
                 "checksum": int(img.values.astype(np.uint64).sum())}
        if not args.no_cpu:
            from datashader_amd.colors import Sets1to3, rgb
            from oracle import shade as osh
            host = agg.values
            how = kw.get("how", "eq_hist")
            times = []
            for _ in range(3):
                c0 = time.perf_counter()
                if host.ndim == 3:
                    osh.colorize(host, [rgb(c) for c in Sets1to3[:host.shape[2]]], how=how)
                else:
                    osh.interpolate(host, [rgb("lightblue"), rgb("darkblue")], how=how)
                times.append(time.perf_counter() - c0)
            entry["cpu_ms_per_call"] = statistics.median(times) * 1e3
        results[label] = entry
        _log(f"[bench] shade {label}: {ms:.3f} ms/call (device {dev_ms:.3f} ms)")
    head = results["count eq_hist"]
    out = {
        "metric": "Mpixels/s shaded (tf.shade on the device)", "value": head["mpix_per_s"],
        "unit": "Mpixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_call"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated on the device)",
        "config": {"workload": "shade: eq_hist of the 1920x1080 count plane of 100M skewed "
                               "points (lightblue-darkblue); also the mean plane (linear, log) "
                               "and a 1024x1024x16 by-category count plane (Sets1to3)"},
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": None, "kernel_ms": head["device_ms_per_call"],
                     "algorithmic_bytes": head["algorithmic_bytes"]},
        "cpu_baseline": None if args.no_cpu else {
            "value": head["pixels"] / (head["cpu_ms_per_call"] * 1e-3) / 1e6,
            "unit": "Mpixels/s", "cores": 1, "kind": "port",
            "sample": "the same planes shaded by the numpy oracle on the host, median of 3"},
        "cases": results,
    }
    print(json.dumps(out), flush=True)


def run_rowindex(args, torch):
    """Row-index reductions over the headline's 100M rows (1920x1080): first, last,
    where(max), where(min) (one atomic pass each after the max / min pass) and max_n(3)
    (a device sort of the kept rows); the sequential oracle on a sample is the CPU
    baseline.  Prints one JSON line."""
    import numpy as np

    import datashader_amd as dsa
    from datashader_amd import _device as dev
    from datashader_amd import synthetic

    n = int(100_000_000 * args.scale)
    W, H = 1920, 1080
    x, y, v = synthetic.points_skewed_dev(n, seed=2)
    w = torch.arange(n, device="cuda", dtype=torch.float32)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v, "w": w})
    xr = dev.maybe_expand_bounds(dev.bounds(x))
    yr = dev.maybe_expand_bounds(dev.bounds(y))
    cvs = dsa.Canvas(W, H, x_range=xr, y_range=yr)
    cases = [("first", dsa.first("v")), ("last", dsa.last("v")),
             ("where_max", dsa.where(dsa.max("v"), "w")),
             ("where_min", dsa.where(dsa.min("v"))), ("max_n3", dsa.max_n("v", n=3))]
    results = {}
    for label, red in cases:
        for _ in range(args.warmup):
            cvs.points(frame, "x", "y", red)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = cvs.points(frame, "x", "y", red)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        results[label] = {"ms_per_call": ms, "gpoints_per_s": n / (ms * 1e-3) / 1e9,
                          "checksum": float(np.nansum(res.values.astype(np.float64)))}
        _log(f"[bench] rowindex {label}: {ms:.3f} ms/call")
    cpu = None
    if not args.no_cpu:
        from oracle import closed_forms as cf
        from oracle import rowindex as ri
        m = min(args.cpu_sample, 5_000_000)
        xs, ys, vs = (t_[:m].double().cpu().numpy() for t_ in (x, y, v))
        sx, tx = cvs.x_axis.compute_scale_and_translate(xr, W)
        sy, ty = cvs.y_axis.compute_scale_and_translate(yr, H)
        import oracle
        view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1])
        t0 = time.perf_counter()
        cell = cf.point_bins(xs, ys, view, W, H)
        ri.select_rows(cell, W * H, "first", vs)
        sec = time.perf_counter() - t0
        cpu = {"value": m / sec / 1e9, "unit": "Gpoints/s", "cores": 1, "kind": "port",
               "sample": f"first {m} rows, first('v') by the sequential Python oracle"}
    head = results["where_max"]
    out = {
        "metric": "Gpoints/s aggregated (row-index reductions)", "value": head["gpoints_per_s"],
        "unit": "Gpoints/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_call"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated on the device)",
        "config": {"workload": "row-index reductions over 100M float32 skewed points into "
                               "1920x1080: where(max('v'), 'w') (value), and first, last, "
                               "where(min('v')), max_n('v', 3)", "rows": n},
        "cpu_baseline": cpu, "cases": results,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="points_skewed",
                    choices=["points_skewed", "points800", "bycat", "lines", "serving",
                             "shade", "stream", "rowindex"])
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload size")
    ap.add_argument("--cpu-sample", type=int, default=20_000_000,
                    help="rows (or line samples) of the CPU oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one plain 1 GiB device copy (the reference dispatch that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--requests", type=int, default=2000, help="serving: viewport requests")
    ap.add_argument("--grid", type=int, default=4096,
                    help="serving: cells per side of the index's lattice")
    ap.add_argument("--tune-requests", type=int, default=96,
                    help="serving: sample requests the autotuner times")
    ap.add_argument("--max-batch", type=int, default=32,
                    help="serving: requests per batch")
    ap.add_argument("--strategy", default="auto", choices=["auto", "window", "tiles"],
                    help="points: force how a call splits its rows (auto: picked from a "
                         "sample by the first call)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "form of this path is a design in DESIGN.md and is not built", file=sys.stderr)
        sys.exit(2)

    import torch

    from datashader_amd import backend

    torch.cuda.set_device(0)
    if args.workload == "serving":
        run_serving(args, torch)
        return
    if args.workload == "shade":
        run_shade(args, torch)
        return
    if args.workload == "rowindex":
        run_rowindex(args, torch)
        return
    from datashader_amd import _abi
    backend.FORCE_STRATEGY = {"auto": None, "window": _abi.STRATEGY_WINDOW,
                              "tiles": _abi.STRATEGY_TILES}[args.strategy]
    _log(f"[bench] building workload {args.workload} (scale {args.scale})")
    step, rows, algo, cfg, cpu, check = build_workload(args.workload, torch, args.scale)
    if args.strategy != "auto":
        cfg["strategy"] = args.strategy
    torch.cuda.synchronize()
    for i in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    _log(f"[bench] warmup done ({args.warmup} steps)")
    if args.copy_calib:
        src = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        _log("[bench] 1 GiB copy done")

    timer = KernelTimer(torch)
    backend.set_kernel_timer(timer)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    backend.set_kernel_timer(None)
    kern_ms = timer.mean_ms()
    ms = (t1 - t0) * 1e3 / args.steps
    checksum = check(res)
    _log(f"[bench] {args.steps} steps: {ms:.3f} ms/step, kernel {kern_ms:.3f} ms, "
         f"checksum {checksum}")

    achieved = algo / (kern_ms * 1e-3) / 1e9
    unit_rows = "Gpoints/s" if args.workload != "lines" else "Gsamples/s"
    value = rows * args.steps / (t1 - t0) / 1e9

    cpu_baseline = None
    if cpu is not None and not args.no_cpu:
        _log("[bench] CPU baseline (1 thread and the host's CPU share)")
        rate_p, P, rate_1, sample, info = cpu(args.cpu_sample, res)
        cpu_baseline = {"value": rate_p / 1e9, "unit": unit_rows, "cores": P, "kind": "port",
                        "sample": sample, "threads": P, "single_thread": rate_1 / 1e9,
                        **info}
        _log(f"[bench] CPU baseline: {rate_1 / 1e9:.3f} {unit_rows} on 1 thread, "
             f"{rate_p / 1e9:.3f} on {P} threads ({info['cpu_model']}, nproc {info['nproc']})")

    out = {
        "metric": "Gpoints/s aggregated + % HBM roofline" if args.workload != "lines"
        else "Gsamples/s aggregated (line vertices) + % HBM roofline",
        "value": value, "unit": unit_rows, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded, generated on the device)", "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": _traffic(args.workload),
                     "kernel_ms": kern_ms, "algorithmic_bytes": algo},
        "cpu_baseline": cpu_baseline,
        "checksum": checksum,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

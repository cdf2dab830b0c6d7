"""Batched viewport serving: many ``Canvas.points`` requests per GPU launch.

The reference answers every ``Canvas.points`` call by scanning its whole source
(``core.py:1334-1359`` -> ``data_libraries/pandas.py:23-66``) and has no serving layer.
Here a resident point set sorted by a lattice (:class:`serving.PointIndex`) answers a
stream of viewport requests -- each its own ranges and canvas size -- in batches:

* :class:`BatchEngine` owns one ``dsx_batch`` engine (``include/dsx.h``): a device
  workspace, a pinned request table and a HIP graph captured once, in which persistent
  workgroups drain a work queue of (request, canvas tile, row part) items.  One
  :meth:`BatchEngine.submit` is one graph replay for up to ``max_requests`` requests.
* :class:`BatchTuner` picks the engine's workgroup size and persistent grid and, per
  class of request (canvas size x candidate rows per pixel), the LDS tile shape and the
  rows per work item, by timing candidates on sample requests.  Results do not depend on
  any of these choices.
* :class:`BatchServer` serves continuously: arrived requests wait in a queue; whenever
  an engine is idle it takes every waiting request (up to its table, the tile budget and
  the output memory budget) as one batch on its own stream, so batches overlap on the
  device while new requests keep arriving.

Each request's result is what ``Canvas.points`` returns for the whole set (counts
bit-exact; float64 sums in another order, within 1e-12 relative for the float32 values
of the serving workload).  Supported reductions: ``count()``, ``count(v)``, ``sum(v)``,
``mean(v)`` and a ``summary`` of them over the index's value column.
"""
from __future__ import annotations

import ctypes
import time
from collections import deque

import numpy as np

from . import _abi
from . import _device as dev
from . import reductions as rd

_PLANE_BITS = {"count": _abi.BATCH_COUNT, "countv": _abi.BATCH_COUNTV,
               "sum": _abi.BATCH_SUM, "mean": _abi.BATCH_MEAN}
_PLANE_BYTES = {"count": 4, "countv": 4, "sum": 8, "mean": 8}


def batch_planes(agg, value="v"):
    """[(name, reduction, plane)] for an aggregation the engine computes, else None."""
    leaves = rd.leaves(agg)
    out = []
    for name, red in leaves:
        if type(red) is rd.count and red.column is None:
            out.append((name, red, "count"))
        elif type(red) is rd.count and red.column == value:
            out.append((name, red, "countv"))
        elif type(red) is rd.sum and red.column == value:
            out.append((name, red, "sum"))
        elif type(red) is rd.mean and red.column == value:
            out.append((name, red, "mean"))
        else:
            return None
    return out


def _device_count():
    t = dev.torch()
    return t.cuda.get_device_properties(t.cuda.current_device()).multi_processor_count


class BatchEngine:
    """One batch engine over a :class:`serving.PointIndex` (float32 columns)."""

    def __init__(self, index, planes=("count", "mean"), max_requests=32, max_tiles=65536,
                 max_regions=4096, threads=1024, workgroups=None, use_graph=True):
        t = dev.torch()
        dev.require_gpu()
        frame = index.frame
        x, y = frame["x"], frame["y"]
        if x.dtype != t.float32:
            raise ValueError("the batch engine serves float32 indexes")
        bits = 0
        for p in planes:
            bits |= _PLANE_BITS[p]
        self.planes = tuple(planes)
        self.index = index
        v = frame["v"] if "v" in frame.columns else None
        if v is None and any(p != "count" for p in planes):
            raise ValueError("count(v) / sum(v) / mean(v) need the index's value column")
        lib = self.lib = _abi.load()
        c = self.cfg = _abi.BatchConfig()
        c.n = index.n
        c.x, c.y, c.val = dev.ptr(x), dev.ptr(y), (dev.ptr(v) if v is not None else None)
        c.grid, c.planes = index.grid, bits
        c.x0, c.x1, c.y0, c.y1 = index.x0, index.x1, index.y0, index.y1
        c.cell_start = dev.ptr(index.frame.grid_index["cell_start"])
        c.max_requests, c.max_tiles, c.max_regions = max_requests, max_tiles, max_regions
        c.threads = threads
        c.workgroups = workgroups or _device_count()
        self.max_requests, self.max_tiles = max_requests, max_tiles
        ws = int(lib.dsx_batch_workspace_size(c))
        if ws == 0:
            raise _abi.DsxError("dsx_batch_workspace_size rejected the configuration")
        self.tile_cells = int(lib.dsx_batch_tile_cells(c))
        self._ws = t.empty(ws, dtype=t.uint8, device=x.device)
        tb = int(lib.dsx_batch_table_bytes(max_requests))
        self._table = t.empty(tb, dtype=t.uint8, pin_memory=True)  # page-locked: copied per batch
        h = ctypes.c_void_p()
        _abi.check(lib.dsx_batch_create(c, dev.ptr(self._ws), ws,
                                        ctypes.c_void_p(self._table.data_ptr()),
                                        1 if use_graph else 0, ctypes.byref(h)),
                   "dsx_batch_create")
        self._h = h
        self._slots = lib.dsx_batch_table(h)
        self.stream = t.cuda.Stream()
        self.event = None
        self.batch = None

    def close(self):
        if getattr(self, "_h", None):
            self.lib.dsx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, vw, tile_h, tile_cells):
        """(tile_w, tile_h) as dsx_batch_submit narrows them: at most the LDS tile's
        cells, at most the canvas."""
        W, H = vw["W"], vw["H"]
        cells = max(1, min(int(tile_cells or self.tile_cells), self.tile_cells))
        th = max(1, min(int(tile_h) if tile_h > 0 else 64, H, cells))
        tw = max(1, min(cells // th, W))
        return tw, th

    def tiles_of(self, vw, tile_h, tile_cells=0):
        """Tiles a request takes with this tile shape."""
        tw, th = self.shape(vw, tile_h, tile_cells)
        return -(-vw["W"] // tw) * -(-vw["H"] // th)

    def submit(self, jobs):
        """Launch one batch on the engine's stream.  ``jobs``: list of
        (view dict, {plane: device tensor}, knobs dict); returns at once."""
        if len(jobs) > self.max_requests:
            raise ValueError("more requests than the engine's table holds")
        s = self._slots
        for k, (vw, outs, knobs) in enumerate(jobs):
            tw, th = self.shape(vw, knobs["tile_h"], knobs.get("tile_cells", 0))
            part_rows = knobs["part_rows"]
            q = s[k]
            q.sx, q.tx, q.sy, q.ty = vw["sx"], vw["tx"], vw["sy"], vw["ty"]
            q.xmin, q.xmax, q.ymin, q.ymax = vw["xmin"], vw["xmax"], vw["ymin"], vw["ymax"]
            q.width, q.height = vw["W"], vw["H"]
            q.tile_h, q.tile_w = int(th), int(tw)
            q.part_rows = int(part_rows)
            q.count = outs["count"].data_ptr() if "count" in outs else None
            q.countv = outs["countv"].data_ptr() if "countv" in outs else None
            q.sum = outs["sum"].data_ptr() if "sum" in outs else None
            q.mean = outs["mean"].data_ptr() if "mean" in outs else None
        _abi.check(self.lib.dsx_batch_submit(self._h, len(jobs),
                                             ctypes.c_void_p(self.stream.cuda_stream)),
                   "dsx_batch_submit")
        t = dev.torch()
        self.event = t.cuda.Event()
        self.event.record(self.stream)

    def done(self):
        return self.event is None or self.event.query()



class Tuning:
    """The tuned knobs: global (threads, workgroups) and per request class (tile_h,
    tile_cells -- the tile's pixels, at most the LDS tile --, part_rows).  Classes: canvas
    cells (log2 bucket) x candidate rows per canvas cell (log4 bucket): dense views want
    small tiles (each reads only the lattice cells under it) and long parts, sparse ones
    large tiles."""

    DEFAULT = {"tile_h": 64, "tile_cells": 0, "part_rows": 1 << 18}

    def __init__(self, threads=1024, workgroups=None, per_class=None):
        self.threads, self.workgroups = threads, workgroups
        self.per_class = dict(per_class or {})

    @staticmethod
    def key(W, H, rows):
        cells = W * H
        dens = rows / cells if cells else 0.0
        dc = int(np.clip(np.floor(np.log(max(dens, 1e-6)) / np.log(4.0)), -4, 6))
        return (int(np.log2(cells)), dc)

    def knobs(self, W, H, rows):
        return self.per_class.get(self.key(W, H, rows), self.DEFAULT)

    def to_dict(self):
        return {"threads": self.threads, "workgroups": self.workgroups,
                "per_class": {f"{k[0]},{k[1]}": v for k, v in self.per_class.items()}}


class RowEstimator:
    """Candidate rows of a view from a summed-area table of the lattice's cell counts
    (host side, O(1) per request): the rows of the cells the view's ranges map to."""

    def __init__(self, index):
        G = index.grid
        cs = np.asarray(index.cell_start, dtype=np.int64)
        counts = np.diff(cs[: G * G + 1]).reshape(G, G)
        sat = np.zeros((G + 1, G + 1), np.int64)
        sat[1:, 1:] = counts.cumsum(0).cumsum(1)
        self.sat, self.G = sat, G
        self.x0, self.sx, self.y0, self.sy = index.x0, index.sx, index.y0, index.sy

    def rows(self, x_range, y_range):
        G = self.G

        def cell(v, v0, s):
            return int(min(max((v - v0) * s, 0.0), G - 1))

        gx0, gx1 = cell(x_range[0], self.x0, self.sx), cell(x_range[1], self.x0, self.sx)
        gy0, gy1 = cell(y_range[0], self.y0, self.sy), cell(y_range[1], self.y0, self.sy)
        s = self.sat
        return int(s[gy1 + 1, gx1 + 1] - s[gy0, gx1 + 1] - s[gy1 + 1, gx0] + s[gy0, gx0])


def _view_of(req):
    from .backend import make_view
    return make_view(req.canvas(), req.x_range, req.y_range)


def _alloc_outputs(planes, W, H, device):
    t = dev.torch()
    dts = {"count": t.int32, "countv": t.int32, "sum": t.float64, "mean": t.float64}
    return {p: t.empty(H * W, dtype=dts[p], device=device) for p in planes}


class BatchTuner:
    """Autotuner of the batch engine, keyed on canvas size and point skew (candidate rows
    per canvas cell of a request).

    For every class of the sample requests it times each candidate (tile_h, tile_cells,
    part_rows) on the class's requests, one request per batch (HIP events on the engine's
    stream), and keeps the fastest; then it times the whole sample in batches under each
    candidate (threads, workgroups) and keeps the fastest.  The result is a
    :class:`Tuning`."""

    TILE_H = (16, 64)
    TILE_DIV = (1, 4, 16, 64)      # tile_cells = LDS tile cells / TILE_DIV
    PART_ROWS = (1 << 17, 1 << 19, 1 << 21)
    THREADS = (512, 1024)
    WG_PER_CU = (1, 2)

    def __init__(self, index, planes=("count", "mean")):
        self.index, self.planes = index, tuple(planes)
        self.est = RowEstimator(index)

    def _time(self, eng, jobs_list, reps=2):
        t = dev.torch()
        torch_stream = eng.stream
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(reps):
            e0.record(torch_stream)
            for jobs in jobs_list:
                eng.submit(jobs)
            e1.record(torch_stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    def tune(self, requests, max_batch=32, verbose=False):
        t = dev.torch()
        device = self.index.frame["x"].device
        views = [_view_of(r) for r in requests]
        rows = [self.est.rows(r.x_range, r.y_range) for r in requests]
        outs = [_alloc_outputs(self.planes, v["W"], v["H"], device) for v in views]
        classes = {}
        for i, (v, n) in enumerate(zip(views, rows)):
            classes.setdefault(Tuning.key(v["W"], v["H"], n), []).append(i)
        eng = BatchEngine(self.index, self.planes, max_requests=max_batch)
        per_class = {}
        for key, idx in sorted(classes.items()):
            best, best_t = None, float("inf")
            for div in self.TILE_DIV:
                for th in self.TILE_H:
                    for pr in self.PART_ROWS:
                        kn = {"tile_h": th, "tile_cells": eng.tile_cells // div,
                              "part_rows": pr}
                        ms = self._time(eng, [[(views[i], outs[i], kn)] for i in idx])
                        if ms < best_t:
                            best, best_t = kn, ms
            per_class[key] = best
            if verbose:
                print(f"[tune] class {key}: {len(idx)} requests -> {best} ({best_t:.3f} ms)",
                      flush=True)
        eng.close()
        tuning = Tuning(per_class=per_class)
        cus = _device_count()
        best_g, best_t = (1024, cus), float("inf")
        for th in self.THREADS:
            for k in self.WG_PER_CU:
                e = BatchEngine(self.index, self.planes, max_requests=max_batch, threads=th,
                                workgroups=k * cus)
                jl, cur, ntl = [], [], 0
                for i in range(len(views)):  # batches within the table and the tile budget
                    kn = tuning.knobs(views[i]["W"], views[i]["H"], rows[i])
                    nt = e.tiles_of(views[i], kn["tile_h"], kn.get("tile_cells", 0))
                    if cur and (len(cur) == max_batch or ntl + nt > e.max_tiles):
                        jl.append(cur)
                        cur, ntl = [], 0
                    cur.append((views[i], outs[i], kn))
                    ntl += nt
                if cur:
                    jl.append(cur)
                ms = self._time(e, jl)
                e.close()
                if verbose:
                    print(f"[tune] threads {th}, workgroups {k * cus}: {ms:.3f} ms", flush=True)
                if ms < best_t:
                    best_g, best_t = (th, k * cus), ms
        tuning.threads, tuning.workgroups = best_g
        t.cuda.synchronize()
        return tuning


class BatchServer:
    """Continuous batching of viewport requests over a :class:`serving.PointIndex`.

    ``engines`` batch engines, each with its own stream, pinned table and captured graph.
    A request waits from its arrival until an engine is idle; an idle engine takes every
    waiting request (up to ``max_batch``, the engine's tile budget and the output memory
    budget ``mem_budget`` over all requests in flight) as one batch.  So requests that
    arrive while other batches run are admitted as soon as an engine and memory free up,
    and up to ``engines`` batches overlap on the device."""

    def __init__(self, index, agg=None, engines=2, max_batch=32, tuning=None,
                 mem_budget=8 << 30, use_graph=True, value="v"):
        agg = agg if agg is not None else rd.summary(count=rd.count(), mean=rd.mean(value))
        leaves = batch_planes(agg, value)
        if leaves is None:
            raise ValueError("the batch engine computes count(), count(v), sum(v) and mean(v) "
                             "of the index's value column only")
        self.agg, self.leaves = agg, leaves
        self.planes = tuple(dict.fromkeys(p for _, _, p in leaves))
        self.index = index
        self.tuning = tuning or Tuning()
        self.est = RowEstimator(index)
        self.mem_budget = int(mem_budget)
        self.max_batch = int(max_batch)
        self.engines = [BatchEngine(index, self.planes, max_requests=max_batch,
                                    threads=self.tuning.threads,
                                    workgroups=self.tuning.workgroups, use_graph=use_graph)
                        for _ in range(engines)]
        self.device = index.frame["x"].device

    def close(self):
        for e in self.engines:
            e.close()

    def _job(self, req):
        vw = _view_of(req)
        W, H = vw["W"], vw["H"]
        return vw, self.tuning.knobs(W, H, self.est.rows(req.x_range, req.y_range))

    def result(self, req, vw, outs):
        """The Canvas.points result (DataArray / Dataset) of a served request."""
        from .backend import wrap_results
        from .glyphs import Point
        res = []
        for name, red, plane in self.leaves:
            tt = outs[plane]
            if plane in ("count", "countv"):
                tt = tt.view(dev.torch().uint32)
            res.append((name, red, tt.reshape(vw["H"], vw["W"])))
        return wrap_results(self.agg, res, req.canvas(), Point("x", "y"), vw, req.x_range,
                            req.y_range, None, False)

    def serve(self, requests, realtime=True, keep_results=False):
        """Serve ``requests`` (sorted by arrival, seconds from the start); returns
        {"latency", "done", "results", "wall", "batches"} as ViewportServer.serve, the
        latency of a request running from its arrival (realtime) or from the start to the
        completion of its batch on the device."""
        t = dev.torch()
        n = len(requests)
        lat = np.zeros(n)
        done_at = np.zeros(n)
        results = [None] * n if keep_results else None
        waiting = deque()
        inflight = {}          # engine index -> (request ids, outputs, views, bytes)
        held = 0
        nxt = 0
        nbatches = 0
        sizes = []
        t0 = time.perf_counter()
        while nxt < n or waiting or inflight:
            now = time.perf_counter() - t0
            while nxt < n and (not realtime or requests[nxt].arrival <= now):
                waiting.append(nxt)
                nxt += 1
            progressed = False
            for ei, eng in enumerate(self.engines):
                if ei in inflight or not waiting:
                    continue
                ids, jobs, outs_l, views, nbytes, tiles = [], [], [], [], 0, 0
                while waiting and len(ids) < self.max_batch:
                    i = waiting[0]
                    vw, kn = self._job(requests[i])
                    need = vw["W"] * vw["H"] * sum(_PLANE_BYTES[p] for p in self.planes)
                    nt = eng.tiles_of(vw, kn["tile_h"], kn.get("tile_cells", 0))
                    if ids and (held + nbytes + need > self.mem_budget or
                                tiles + nt > eng.max_tiles):
                        break
                    waiting.popleft()
                    with t.cuda.stream(eng.stream):
                        outs = _alloc_outputs(self.planes, vw["W"], vw["H"], self.device)
                    ids.append(i)
                    jobs.append((vw, outs, kn))
                    outs_l.append(outs)
                    views.append(vw)
                    nbytes += need
                    tiles += nt
                if not ids:
                    continue
                eng.submit(jobs)
                inflight[ei] = (ids, outs_l, views, nbytes)
                held += nbytes
                nbatches += 1
                sizes.append(len(ids))
                progressed = True
            for ei in list(inflight):
                eng = self.engines[ei]
                if eng.done():
                    ids, outs_l, views, nbytes = inflight.pop(ei)
                    tnow = time.perf_counter() - t0
                    for i, outs, vw in zip(ids, outs_l, views):
                        done_at[i] = tnow
                        lat[i] = tnow - (requests[i].arrival if realtime else 0.0)
                        if keep_results:
                            results[i] = self.result(requests[i], vw, outs)
                    held -= nbytes
                    progressed = True
            if not progressed:
                if inflight:
                    time.sleep(20e-6)
                elif nxt < n and realtime:
                    time.sleep(min(max(requests[nxt].arrival - (time.perf_counter() - t0), 0.0),
                                   1e-3))
        t.cuda.synchronize()
        return {"latency": lat, "done": done_at, "results": results,
                "wall": time.perf_counter() - t0, "batches": nbatches,
                "mean_batch": float(np.mean(sizes)) if sizes else 0.0}

"""Viewport serving over a resident point set.

The reference answers every ``Canvas.points`` call by scanning every row of its source.
A viewport server keeps one point set resident in HBM and answers a stream of viewport
requests (ranges, canvas size, reductions) against it:

* :class:`PointIndex` sorts the rows once by the cell of a ``grid x grid`` lattice over
  the set's extent (``dsx_grid_index_build``: a rocPRIM radix sort of packed rows), so
  that a viewport reads only the rows of the cells it overlaps;
* a request gathers those rows (``dsx_gather_rows``) and aggregates them with the same
  ``Canvas.points`` path as any call -- the viewport's own range test decides which rows
  land, so the result is the reference's for the whole set (counts bit-exact, float sums
  to the accumulation-order tolerance);
* :class:`ViewportServer` runs requests continuously: each admitted request is launched
  at once on a free slot (its own HIP stream), so requests that arrive while others are
  on the GPU overlap with them; admission is bounded by the slots and by the device
  memory the in-flight requests hold.

Nothing here has a counterpart in the reference (it has no serving layer); the
aggregation itself is ``glyphs/points.py:188-221`` as everywhere else in this package.
"""
from __future__ import annotations

import time
from collections import OrderedDict

import numpy as np

from . import _abi
from . import _device as dev
from .frame import DSX_DTYPE, DeviceFrame

_ITEM_ROWS = 1 << 16       # rows per gather work item
FULL_SCAN_FRACTION = 0.25  # above this share of the rows a request scans the whole set


class PointIndex:
    """A resident point set sorted by grid cell.

    ``x``, ``y`` and the optional value column ``val`` are 1-D device tensors of one float
    dtype.  ``frame`` is the sorted set as a :class:`DeviceFrame` with columns ``x``, ``y``
    (and ``v``); ``cell_start`` (host, int64) gives the rows of each cell."""

    def __init__(self, x, y, val=None, grid=1024, extent=None):
        t = dev.torch()
        dev.require_gpu()
        if x.dtype != y.dtype or (val is not None and val.dtype != x.dtype):
            raise ValueError("x, y and the value column must share one float dtype")
        kind = {t.float32: "f4", t.float64: "f8"}.get(x.dtype)
        if kind is None:
            raise ValueError("the index holds float32 or float64 columns")
        n = int(x.shape[0])
        if n < 1:
            raise ValueError("an index needs at least one row")
        if extent is None:
            extent = (dev.maybe_expand_bounds(dev.bounds(x)), dev.maybe_expand_bounds(dev.bounds(y)))
        (x0, x1), (y0, y1) = extent
        self.grid, self.n = int(grid), n
        self.x0, self.x1, self.y0, self.y1 = float(x0), float(x1), float(y0), float(y1)
        self.sx = self.grid / (self.x1 - self.x0)
        self.sy = self.grid / (self.y1 - self.y0)
        lib = _abi.load()
        ox, oy = t.empty_like(x), t.empty_like(y)
        ov = t.empty_like(val) if val is not None else None
        cs = t.empty(self.grid * self.grid + 1, dtype=t.int64, device=x.device)
        g = _abi.GridIndexDesc()
        g.n, g.grid, g.dtype = n, self.grid, DSX_DTYPE[kind]
        g.x, g.y, g.val = dev.ptr(x), dev.ptr(y), dev.ptr(val) if val is not None else None
        g.x0, g.x1, g.y0, g.y1 = self.x0, self.x1, self.y0, self.y1
        g.out_x, g.out_y = dev.ptr(ox), dev.ptr(oy)
        g.out_val = dev.ptr(ov) if ov is not None else None
        g.cell_start = dev.ptr(cs)
        nbytes = lib.dsx_grid_index_workspace_size(g)
        if nbytes == 0:
            raise _abi.DsxError("dsx_grid_index_workspace_size rejected the index")
        ws = t.empty(nbytes, dtype=t.uint8, device=x.device)
        _abi.check(lib.dsx_grid_index_build(g, dev.ptr(ws), nbytes, dev.stream_handle()),
                   "dsx_grid_index_build")
        self.cell_start = cs.cpu().numpy()
        del ws
        cols = {"x": ox, "y": oy}
        if ov is not None:
            cols["v"] = ov
        self.frame = DeviceFrame(cols)
        # Canvas.points over this frame reads the lattice: each canvas tile reads only
        # the rows of the cells under it (dsx_points_agg_indexed)
        self.frame.grid_index = {"x": "x", "y": "y", "grid": self.grid,
                                 "extent": (self.x0, self.x1, self.y0, self.y1),
                                 "cell_start": cs}
        self.indexed = int(self.cell_start[-1])  # rows with both coordinates finite

    def _cell(self, v, v0, s):
        f = (float(v) - v0) * s
        return int(min(max(f, 0.0), self.grid - 1))

    def ranges(self, x_range, y_range):
        """Row intervals [start, end) of the sorted set holding every row that can fall
        in the viewport (the cells its ranges map to, one more cell on each side)."""
        if x_range[1] < self.x0 or x_range[0] > self.x1 or y_range[1] < self.y0 or \
                y_range[0] > self.y1:
            return []
        G, cs = self.grid, self.cell_start
        gx0 = max(self._cell(x_range[0], self.x0, self.sx) - 1, 0)
        gx1 = min(self._cell(x_range[1], self.x0, self.sx) + 1, G - 1)
        gy0 = max(self._cell(y_range[0], self.y0, self.sy) - 1, 0)
        gy1 = min(self._cell(y_range[1], self.y0, self.sy) + 1, G - 1)
        if gx0 == 0 and gx1 == G - 1:
            s, e = int(cs[gy0 * G]), int(cs[(gy1 + 1) * G])
            return [(s, e)] if e > s else []
        rows = np.arange(gy0, gy1 + 1) * G
        starts = cs[rows + gx0]
        ends = cs[rows + gx1 + 1]
        out = []
        for s, e in zip(starts.tolist(), ends.tolist()):
            if e <= s:
                continue
            if out and out[-1][1] == s:
                out[-1] = (out[-1][0], e)
            else:
                out.append((s, e))
        return out

    def candidates(self, x_range, y_range):
        """DeviceFrame of the rows that can fall in the viewport (for drill-down: the
        rows behind a view): the whole sorted set when they are more than
        FULL_SCAN_FRACTION of it, else a gathered copy (on the current stream)."""
        t = dev.torch()
        iv = self.ranges(x_range, y_range)
        total = sum(e - s for s, e in iv)
        if total > FULL_SCAN_FRACTION * self.n:
            return self.frame, total
        names = list(self.frame.columns)
        src = [self.frame[c] for c in names]
        dst = [t.empty(total, dtype=c.dtype, device=c.device) for c in src]
        if total:
            items = []
            off = 0
            for s, e in iv:
                for a in range(s, e, _ITEM_ROWS):
                    ln = min(_ITEM_ROWS, e - a)
                    items += (a, off, ln)
                    off += ln
            host = t.tensor(items, dtype=t.int64).pin_memory()
            dev_items = host.to(src[0].device, non_blocking=True)
            d = _abi.GatherDesc()
            d.ncols = len(src)
            for k, (a, b) in enumerate(zip(src, dst)):
                d.elem_bytes[k] = a.element_size()
                d.src[k], d.dst[k] = dev.ptr(a), dev.ptr(b)
            d.items, d.nitems = dev.ptr(dev_items), len(items) // 3
            # (the pinned staging buffer and the device item list are stream-ordered
            # allocations: torch reuses them only after this stream's work on them)
            _abi.check(_abi.load().dsx_gather_rows(d, dev.stream_handle()), "dsx_gather_rows")
        f = DeviceFrame(dict(zip(names, dst)))
        f.persistent = False
        return f, total


class Request:
    """One viewport: ranges, canvas size and the reductions to compute."""

    def __init__(self, x_range, y_range, width, height, agg, arrival=0.0):
        self.x_range, self.y_range = tuple(x_range), tuple(y_range)
        self.width, self.height = int(width), int(height)
        self.agg = agg
        self.arrival = float(arrival)

    def canvas(self):
        from .core import Canvas
        return Canvas(self.width, self.height, x_range=self.x_range, y_range=self.y_range)

    def plane_bytes(self):
        return self.width * self.height * 32


def run_plain(frame, req):
    """The plain per-request form: one Canvas.points call over the whole set."""
    return req.canvas().points(frame, "x", "y", req.agg)


def run_indexed(index, req):
    """One request through the index (asynchronous on the current stream): the same
    Canvas.points call, over the sorted frame, which the backend aggregates tile by tile
    from the lattice cells under each tile."""
    return req.canvas().points(index.frame, "x", "y", req.agg)


class ViewportServer:
    """Continuous serving of viewport requests over a :class:`PointIndex`.

    Up to ``slots`` requests are on the GPU at once, each on its own stream; a request is
    admitted when it has arrived, a slot is free and the device memory held by the
    in-flight requests (gathered rows, planes and workspace, estimated) stays within
    ``mem_budget`` bytes."""

    def __init__(self, index, slots=4, mem_budget=32 << 30):
        t = dev.torch()
        self.index = index
        self.streams = [t.cuda.Stream() for _ in range(slots)]
        self.mem_budget = int(mem_budget)

    def _estimate(self, req):
        # planes and finalized outputs, plus the indexed call's workspace (the tile
        # parts' merge regions: about the canvas planes plus two regions per CU)
        return 2 * req.plane_bytes() + (128 << 20)

    def serve(self, requests, realtime=True, keep_results=False):
        """Serve ``requests`` (sorted by arrival, seconds from the start) and return a
        dict with per-request latency (completion - arrival, seconds), completion times
        and, with ``keep_results``, the results in request order."""
        t = dev.torch()
        n = len(requests)
        lat = np.zeros(n)
        done_at = np.zeros(n)
        results = [None] * n if keep_results else None
        free = list(range(len(self.streams)))
        inflight = OrderedDict()   # slot -> (request index, event, result, bytes)
        held = 0
        nxt = 0
        t0 = time.perf_counter()
        while nxt < n or inflight:
            now = time.perf_counter() - t0
            launched = False
            while nxt < n and free and (not realtime or requests[nxt].arrival <= now):
                req = requests[nxt]
                need = self._estimate(req)
                if inflight and held + need > self.mem_budget:
                    break
                slot = free.pop()
                with t.cuda.stream(self.streams[slot]):
                    res = run_indexed(self.index, req)
                    ev = t.cuda.Event()
                    ev.record(self.streams[slot])
                inflight[slot] = (nxt, ev, res, need)
                held += need
                nxt += 1
                launched = True
            for slot in list(inflight):
                i, ev, res, need = inflight[slot]
                if ev.query():
                    tnow = time.perf_counter() - t0
                    done_at[i] = tnow
                    lat[i] = tnow - (requests[i].arrival if realtime else 0.0)
                    if keep_results:
                        results[i] = res
                    del inflight[slot]
                    held -= need
                    free.append(slot)
            if not launched and inflight:
                time.sleep(20e-6)
            elif not launched and nxt < n and realtime:
                time.sleep(min(max(requests[nxt].arrival - (time.perf_counter() - t0), 0.0), 1e-3))
        return {"latency": lat, "done": done_at, "results": results,
                "wall": time.perf_counter() - t0}


def make_requests(n, index_or_frame_rows, extent, seed=5, rate=200.0, agg=None,
                  sizes=((256, 256), (512, 512), (1024, 768), (1920, 1080)), zoom_max_log2=12):
    """The viewport stream of the serving benchmark: Poisson arrivals at ``rate`` per
    second, zoom 2**U(0, zoom_max_log2), centre at a random resident row (so centres
    follow the data density), canvas size drawn from ``sizes``.  ``index_or_frame_rows``
    is (x, y) host arrays of candidate centres."""
    from . import reductions as rd
    rng = np.random.default_rng(seed)
    cx, cy = index_or_frame_rows
    (x0, x1), (y0, y1) = extent
    w, h = x1 - x0, y1 - y0
    agg = agg if agg is not None else rd.summary(count=rd.count(), mean=rd.mean("v"))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for k in range(n):
        z = 2.0 ** rng.uniform(0, zoom_max_log2)
        j = rng.integers(0, len(cx))
        W, H = sizes[rng.integers(0, len(sizes))]
        hw, hh = w / z / 2, h / z / 2
        xr = (float(cx[j] - hw), float(cx[j] + hw))
        yr = (float(cy[j] - hh), float(cy[j] + hh))
        out.append(Request(xr, yr, W, H, agg, arrival=float(arrivals[k])))
    return out

"""Out-of-core aggregation: a host-resident table streamed through the GPU in chunks.

The reference aggregates tables larger than memory with its dask backend
(``data_libraries/dask.py:86-219``): every partition is aggregated into its own canvas and
the partial canvases are merged with the reductions' ``_combine``
(``compiler.py:478-507``).  ``ChunkedFrame`` is the analogue here for tables that live in
host memory (larger than HBM, or simply not worth keeping resident): ``Canvas.points`` /
``Canvas.line`` over a ``ChunkedFrame`` upload it chunk by chunk and aggregate each
chunk with the same kernels as a resident ``DeviceFrame``, merging the chunk planes on
the device with ``dsx_combine``.

The upload is pipelined: two pinned staging slots and two device slots, the host copy
of chunk i+1 into pinned memory (parallel threads) and its DMA on a copy stream overlap
the aggregation of chunk i on the compute stream.  With ``x_range`` / ``y_range`` left
to autorange, a first streamed pass computes the bounds (as dask computes them over all
partitions first).  var / std stream the table twice: the second pass accumulates the
squared deviations from the final per-cell means (the deviations of every chunk from the
same mean add, so the merge is exact in the arithmetic sense).

Points and the axis-1 line layouts (one line per row) have independent rows.  Axis-0
lines join consecutive rows, so each chunk after the first is uploaded with the two rows
before it in front: the first is context only (``DSX_PLOT_CONTEXT``: not drawn, its NaN
test gives the next segment its segment_start), so the segment that joins the chunks is
drawn once, exactly as in the unchunked line (a chunk boundary is invisible).  Between
partitions the reference's dask stitching is followed (``data_libraries/dask.py:222-266``):
a ``LineAxis0`` partition is drawn as the previous partition's last row followed by its
own rows, with ``plot_start=False``; ``LinesAxis0Multi`` partitions are drawn one by one
(dask's default path, ``dask.py:86-219``).  Row-index reductions (first, last, where,
the n-value reductions) are not streamed (NotImplementedError).
"""
from __future__ import annotations

from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pandas as pd

from . import _abi
from . import _device as dev
from . import backend
from .core import _cols_to_keep, bypixel
from .frame import DeviceFrame, Schema
from .glyphs import (LineAxis0, LineAxis0Multi, LinesAxis1, LinesAxis1XConstant,
                     LinesAxis1YConstant, Point)


class ChunkedFrame:
    """A host-resident table (one pandas DataFrame, or a list of DataFrame partitions
    with the same columns) that ``Canvas`` aggregates ``chunk_rows`` rows at a time."""

    streams_in_chunks = True

    def __init__(self, data, chunk_rows=1 << 25):
        parts = [data] if isinstance(data, pd.DataFrame) else list(data)
        if not parts or not all(isinstance(p, pd.DataFrame) for p in parts):
            raise ValueError("ChunkedFrame takes a pandas DataFrame or a list of them")
        first = Schema.from_pandas(parts[0])
        for p in parts[1:]:
            s = Schema.from_pandas(p)
            if s.dict != first.dict:
                raise ValueError("all partitions of a ChunkedFrame must have the same columns "
                                 "and dtypes")
            for c, k in first.dict.items():
                if k == "category" and s.categories(c) != first.categories(c):
                    raise ValueError(f"partitions disagree on the categories of {c!r}")
        if int(chunk_rows) < 1:
            raise ValueError("chunk_rows must be positive")
        self.partitions = parts
        self.chunk_rows = int(chunk_rows)
        self._schema = first

    @property
    def schema(self):
        return self._schema

    @property
    def columns(self):
        return list(self.partitions[0].columns)

    def __len__(self):
        return sum(len(p) for p in self.partitions)

    def _host_columns(self, part, names):
        out = {}
        for n in names:
            s = part[n]
            if isinstance(s.dtype, pd.CategoricalDtype):
                out[n] = np.ascontiguousarray(s.cat.codes.to_numpy())
            else:
                out[n] = np.ascontiguousarray(s.to_numpy())
        return out

    def chunks(self, names, stitch=None):
        """(cols, head, plot_start) per chunk, partition by partition: ``cols`` is
        {name: numpy view} of the chunk's rows; ``head`` is None or {name: array of 1 or
        2 rows} to upload in front of them; ``plot_start`` is the line kernel's
        plot_start for the uploaded rows (axis-0 lines).

        ``stitch``: None (rows independent), "axis0" (``LineAxis0``: partitions joined as
        dask joins them) or "multi" (``LinesAxis0Multi``: partitions drawn one by one).
        Within a partition a chunk gets the two rows before it in front, the first as
        context (``DSX_PLOT_CONTEXT``), or the one row before it with the partition's own
        plot_start when that row opens the partition's line."""
        prev_last = None  # last row of the previous partition (dask's df.iloc[-1:])
        for ip, part in enumerate(self.partitions):
            cols = self._host_columns(part, names)
            n = len(part)
            # LineAxis0: partition i > 0 is drawn as concat(previous partition's last
            # row, its rows) with plot_start False (dask.py:242-248), also when the
            # previous partition is empty and hands on no row
            tail = [prev_last] if stitch == "axis0" and prev_last is not None else []
            part_start = not (stitch == "axis0" and ip > 0)
            for a in range(0, n, self.chunk_rows):
                b = min(n, a + self.chunk_rows)
                chunk = {k: v[a:b] for k, v in cols.items()}
                if stitch is None or not tail:
                    yield chunk, None, part_start
                else:
                    head = {k: np.concatenate([r[k] for r in tail]) for k in cols}
                    yield chunk, head, (_abi.PLOT_CONTEXT if len(tail) == 2 else part_start)
                if stitch is not None:
                    rows = [{k: v[j:j + 1].copy() for k, v in cols.items()}
                            for j in range(max(a, b - 2), b)]
                    tail = (tail + rows)[-2:]
            prev_last = {k: v[n - 1:n].copy() for k, v in cols.items()} if n else None


class _Uploader:
    """Two pinned staging slots + two device slots; yields one DeviceFrame per chunk
    with its upload overlapped with the previous chunk's aggregation."""

    def __init__(self, source, names, workers=8, stitch=None):
        t = dev.torch()
        self.t = t
        self.source = source
        self.names = list(names)
        self.stitch = stitch
        part0 = source.partitions[0]
        probe = source._host_columns(part0.iloc[:1], self.names)
        self.dtypes = {n: probe[n].dtype for n in self.names}
        self.cats = {n: source.schema.categories(n) for n in self.names
                     if source.schema[n] == "category"}
        longest = max(len(p) for p in source.partitions)
        # (+2: an axis-0 chunk is uploaded behind the two rows before it)
        self.cap = max(1, min(source.chunk_rows, longest)) + (2 if stitch else 0)
        groups = OrderedDict()
        for n in self.names:
            groups.setdefault(self.dtypes[n], []).append(n)
        self.groups = groups
        self.pinned, self.device = [], []
        for _ in range(2):
            hp, dv = {}, {}
            for dt, ns in groups.items():
                tdt = t.from_numpy(np.zeros(1, dtype=dt)).dtype
                hp[dt] = t.empty((len(ns), self.cap), dtype=tdt, pin_memory=True)
                # one [ncols, cap] block per dtype: columns are equally spaced rows, which
                # the line kernels read in place as one strided matrix
                dv[dt] = t.empty((len(ns), self.cap), dtype=tdt, device="cuda")
            self.pinned.append(hp)
            self.device.append(dv)
        self.copy_stream = t.cuda.Stream()
        self.h2d = [None, None]
        self.used = [None, None]
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.workers = workers

    def _stage(self, slot, chunk, head=None):
        t = self.t
        h = 0 if head is None else len(next(iter(head.values())))
        m = len(next(iter(chunk.values()))) + h
        if self.h2d[slot] is not None:
            self.h2d[slot].synchronize()  # the pinned slot's previous DMA has finished
        jobs = []
        for dt, ns in self.groups.items():
            host = self.pinned[slot][dt].numpy()
            for j, n in enumerate(ns):
                if h:
                    host[j][:h] = head[n]
                src, dst = chunk[n], host[j][h:]
                m0 = m - h
                step = max(1, -(-m0 // self.workers))
                for a in range(0, m0, step):
                    b = min(m0, a + step)
                    jobs.append(self.pool.submit(np.copyto, dst[a:b], src[a:b]))
        for f in jobs:
            f.result()
        with t.cuda.stream(self.copy_stream):
            if self.used[slot] is not None:
                self.copy_stream.wait_event(self.used[slot])  # its last chunk is aggregated
            for dt, ns in self.groups.items():
                for j in range(len(ns)):
                    self.device[slot][dt][j, :m].copy_(self.pinned[slot][dt][j, :m],
                                                       non_blocking=True)
            ev = t.cuda.Event()
            ev.record(self.copy_stream)
            self.h2d[slot] = ev
        return m

    def frames(self, window_cache=None):
        t = self.t
        it = self.source.chunks(self.names, self.stitch)
        first = next(it, None)
        if first is None:
            return
        m = self._stage(0, first[0], first[1])
        plot_start = first[2]
        slot = 0
        while True:
            t.cuda.current_stream().wait_event(self.h2d[slot])
            cols = {}
            for dt, ns in self.groups.items():
                for j, n in enumerate(ns):
                    cols[n] = self.device[slot][dt][j, :m]
            frame = DeviceFrame({n: cols[n] for n in self.names}, self.cats)
            frame.plot_start = plot_start
            if window_cache is not None:
                frame.__dict__["_window_cache"] = window_cache  # one tuning state per stream
            yield frame
            ev = t.cuda.Event()
            ev.record(t.cuda.current_stream())
            self.used[slot] = ev
            nxt = next(it, None)
            if nxt is None:
                break
            slot = 1 - slot
            m = self._stage(slot, nxt[0], nxt[1])  # overlaps the aggregation just enqueued
            plot_start = nxt[2]

    def close(self):
        self.pool.shutdown(wait=False)


def _combine(lib, into, part):
    """into.planes (+)= part.planes on the device (dsx_combine)."""
    c = _abi.CombineDesc()
    c.ncells = int(next(iter(into.planes.values())).numel())
    c.key_bytes = into.keyb
    for name in ("count", "countv", "sum", "min", "max"):
        if name in into.planes:
            setattr(c.into, name, dev.ptr(into.planes[name]))
            setattr(c.part, name, dev.ptr(part.planes[name]))
    _abi.check(lib.dsx_combine(c, dev.stream_handle()), "dsx_combine")


def _add_plane(lib, into, part):
    """into += part for one float64 plane (the m2 planes)."""
    c = _abi.CombineDesc()
    c.ncells = int(into.numel())
    c.key_bytes = 8
    c.into.sum, c.part.sum = dev.ptr(into), dev.ptr(part)
    _abi.check(lib.dsx_combine(c, dev.stream_handle()), "dsx_combine")


def _axis_columns(glyph, axis):
    """Columns holding the glyph's coordinates on one axis (None: constant)."""
    if axis == "x" and isinstance(glyph, LinesAxis1XConstant):
        return None
    if axis == "y" and isinstance(glyph, LinesAxis1YConstant):
        return None
    v = getattr(glyph, axis)
    return [v] if isinstance(v, str) else list(v)


def _stream_bounds(source, glyph, axes):
    """NaN-skipping (min, max) of each axis over every chunk, streamed through the GPU
    (glyph.py:66-78 per chunk, merged with min / max)."""
    names = []
    for a in axes:
        names += [c for c in (_axis_columns(glyph, a) or []) if c not in names]
    lo = {a: np.inf for a in axes}
    hi = {a: -np.inf for a in axes}
    up = _Uploader(source, names)
    try:
        for frame in up.frames():
            for a in axes:
                for c in _axis_columns(glyph, a) or []:
                    b0, b1 = dev.bounds(frame[c])
                    lo[a] = np.fmin(lo[a], b0)
                    hi[a] = np.fmax(hi[a], b1)
    finally:
        up.close()
    return {a: dev.maybe_expand_bounds((lo[a], hi[a])) for a in axes}


def pipeline(source, schema, canvas, glyph, agg, *, antialias=False, to_host=True):
    """``bypixel.pipeline`` for a ``ChunkedFrame``: stream, aggregate, combine, finalize."""
    if antialias:
        raise NotImplementedError("antialiased lines are not supported by the MI355X backend")
    if not isinstance(glyph, (Point, LinesAxis1, LineAxis0, LineAxis0Multi)):
        raise NotImplementedError(
            f"{type(glyph).__name__} is not streamed in chunks; upload the table as a "
            "DeviceFrame")
    passes, leaves = backend.plan(agg)
    rowindex = [type(red.reduction if hasattr(red, "reduction") else red).__name__
                for (_, red, key) in leaves if key[0] == "rowindex"]
    if rowindex:
        raise NotImplementedError(
            f"row-index reductions ({', '.join(rowindex)}) are not streamed in chunks; "
            "upload the table as a DeviceFrame")
    stitch = ("axis0" if isinstance(glyph, LineAxis0) else
              "multi" if isinstance(glyph, LineAxis0Multi) else None)
    dev.require_gpu()
    lib = _abi.load()
    x_range, y_range = canvas.x_range, canvas.y_range
    missing = [a for a, r in (("x", x_range), ("y", y_range)) if r is None]
    if missing:
        b = _stream_bounds(source, glyph, [a for a in missing if _axis_columns(glyph, a)])
        if x_range is None:
            x_range = b["x"] if "x" in b else glyph.compute_x_bounds()
        if y_range is None:
            y_range = b["y"] if "y" in b else glyph.compute_y_bounds()
    canvas.validate_ranges(x_range, y_range)
    view = backend.make_view(canvas, x_range, y_range)
    H, W = view["H"], view["W"]
    names = [c for c in _cols_to_keep(source.columns, glyph, agg)]
    cache = OrderedDict()
    totals = {}
    up = _Uploader(source, names, stitch=stitch)
    try:
        for frame in up.frames(cache):
            for key, p in passes.items():
                r = backend.run_pass(lib, frame, schema, glyph, view, p, H, W, with_m2=False,
                                     fuse_mean=False)
                if key in totals:
                    _combine(lib, totals[key], r)
                else:
                    totals[key] = r
        m2_keys = [k for k, p in passes.items() if p.m2]
        if m2_keys:
            t = dev.torch()
            means = {}
            for k in m2_keys:
                res = totals[k]
                ncells = H * W * res.ncat
                means[k] = t.empty(ncells, dtype=t.float64, device="cuda")
                backend._finalize(lib, _abi.FIN_MEAN, res.keyb, ncells, res.planes, means[k])
            for frame in up.frames(cache):
                for k in m2_keys:
                    res = totals[k]
                    m2 = backend.m2_pass(lib, frame, schema, glyph, view, passes[k],
                                         H * W * res.ncat, means[k])
                    if "m2" in res.planes:
                        _add_plane(lib, res.planes["m2"], m2)
                    else:
                        res.planes["m2"] = m2
    finally:
        up.close()
    res = [(name, red, backend.finalize_leaf(lib, red, totals[key], H, W))
           for (name, red, key) in leaves]
    return backend.wrap_results(agg, res, canvas, glyph, view, x_range, y_range, schema,
                                to_host)


bypixel.pipeline.register(ChunkedFrame)(pipeline)

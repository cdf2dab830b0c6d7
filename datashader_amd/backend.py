"""The MI355X backend of ``bypixel`` (the analogue of ``data_libraries/pandas.py`` and
``data_libraries/cudf.py``).

``pipeline`` follows ``data_libraries/pandas.py:23-66`` step by step -- bounds, scale /
translate, coordinates, create, extend, finalize -- with every step that touches the
columns running as a libdsx kernel:

* reductions are grouped into *passes*: one kernel launch per (value column,
  categorizer) pair, producing the union of the accumulator planes its reductions need
  (``compiler.py:36-150`` dedups bases the same way);
* ``dsx_points_agg`` writes every cell of its planes; line planes are created with the
  reference's initial values by ``dsx_fill`` (``reductions.py:446-472``);
* ``dsx_points_agg`` / ``dsx_lines_agg`` is the ``extend`` call;
* var / std add a second ``dsx_points_agg`` pass that is given the per-bin means and
  accumulates squared deviations from them;
* ``dsx_finalize`` turns planes into the reference's output dtypes, and the result is
  wrapped with the same dims / coords / attrs as ``pandas.py:61-66``.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import pandas as pd

from . import _abi
from . import _device as dev
from . import reductions as rd
from .core import bypixel
from .frame import DSX_DTYPE, DeviceFrame
from .glyphs import Point, _LineLike

try:  # pragma: no cover - xarray is not part of this image
    import xarray as _xr
except Exception:  # noqa: BLE001
    _xr = None
from . import xrlite


# ---- instrumentation ------------------------------------------------------------------
# An optional timer with start() / stop() methods, called around every aggregation
# kernel launch (on the stream the kernel runs on).  None in normal use.
_kernel_timer = None


def set_kernel_timer(timer):
    global _kernel_timer
    _kernel_timer = timer


def _timed(call):
    t = _kernel_timer
    if t is None:
        return call()
    t.start()
    r = call()
    t.stop()
    return r


# ---- planning ---------------------------------------------------------------------

class Pass:
    """One kernel launch: a value column (or none) and a categorizer (or none)."""

    def __init__(self, val, cat):
        self.val = val
        self.cat = cat
        self.planes = set()
        self.m2 = False
        self.leaves = []  # the (inner) reductions finalized from this pass

    def __repr__(self):
        return f"Pass(val={self.val!r}, cat={self.cat!r}, planes={sorted(self.planes)}, m2={self.m2})"


def plan(agg, packed=frozenset()):
    """Group the leaves of ``agg`` into passes; returns (passes, [(name, red, pass)]).
    Row-index leaves (first, last, where, the row-index selectors) get the key
    ``("rowindex", i)``: they run their own kernel (``row_index_leaf``), after the passes
    (a where over max / min adds that plane to the pass of its selector's column, unless
    the column is in ``packed``: float32 columns, whose where takes one packed pass)."""
    passes = {}
    leaves = []
    for i, (name, red) in enumerate(rd.leaves(agg)):
        cat = red.categorizer if isinstance(red, rd.by) else None
        inner = red.reduction if isinstance(red, rd.by) else red
        if getattr(inner, "row_index", False):
            sel = inner.selector if isinstance(inner, rd.where) else inner
            if isinstance(sel, (rd.max, rd.min)) and sel.column not in packed:
                key = (sel.column, cat)
                p = passes.get(key)
                if p is None:
                    p = passes[key] = Pass(sel.column, cat)
                p.planes.add("max" if isinstance(sel, rd.max) else "min")
            leaves.append((name, red, ("rowindex", i)))
            continue
        key = (inner.column, cat)
        p = passes.get(key)
        if p is None:
            p = passes[key] = Pass(inner.column, cat)
        p.planes.update(inner.planes())
        p.m2 |= inner.needs_m2
        p.leaves.append(inner)
        leaves.append((name, red, key))
    # a pass with no value column rides along with a value pass of the same categorizer
    for key in [k for k in passes if k[0] is None and k in passes]:
        host = next((k for k in passes if k[0] is not None and k[1] == key[1]), None)
        if host is not None:
            gone = passes.pop(key)
            passes[host].planes.update(gone.planes)
            passes[host].leaves.extend(gone.leaves)
            leaves = [(n, r, host if k == key else k) for (n, r, k) in leaves]
    return passes, leaves


# ---- execution ----------------------------------------------------------------------

def _alloc(ncells, kind, device):
    t = dev.torch()
    dt = {"u32": t.int32, "f32": t.float32, "f64": t.float64, "u64": t.int64}[kind]
    return t.empty(ncells, dtype=dt, device=device)


def _fill(lib, plane, mode):
    _abi.check(lib.dsx_fill(dev.ptr(plane), mode, plane.numel(), dev.stream_handle()), "dsx_fill")


def _categorizer(cat, frame, schema):
    c = _abi.Categorizer()
    if cat is None:
        c.kind, c.ncat = _abi.CAT_NONE, 1
        return c, None, 1
    args = cat.kernel_args(schema)
    col = frame[cat.column]
    c.kind = args["kind"]
    c.ncat = args["ncat"]
    c.dtype = dev.dtype_code(col)
    c.col = dev.ptr(col)
    c.offset = args.get("offset", 0)
    c.modulo = args.get("modulo", 1)
    c.nbins = args.get("nbins", 0)
    c.bin_under = args.get("bin_under", 0)
    c.bin_over = args.get("bin_over", 0)
    c.bin0 = args.get("bin0", 0.0)
    c.binsize = args.get("binsize", 1.0)
    return c, col, args["ncat"]


class PassResult:
    def __init__(self, ncat, keyb, planes):
        self.ncat = ncat
        self.keyb = keyb          # min / max key width in bytes (4: float32 values)
        self.planes = planes


_WINDOW_CACHE_MAX = 64

# How points calls split their rows between passes: None lets each call pick (AUTO, then
# the strategy the first call on a resident frame picked); _abi.STRATEGY_WINDOW or
# _abi.STRATEGY_TILES forces one (tests and benchmarks compare them; results agree).
FORCE_STRATEGY = None


class _WindowSlot:
    """One cached window record: the device int32[8] {x0, y0, w, h, strategy, ...} a
    points call picks and later calls reuse, and the strategy it picked once read back."""
    __slots__ = ("rec", "host", "event", "strategy")

    def __init__(self, rec):
        self.rec, self.host, self.event, self.strategy = rec, None, None, None


def _window_slot(frame, key, device):
    """(slot, reuse flag, strategy) of the hot-window cache kept on a DeviceFrame.  The
    record only steers performance (which canvas cells are kept in LDS during the scan,
    and whether the rows are bucketed by tile instead); any stored record gives the same
    result.  A call copies the stored record into its own workspace before any of its
    kernels reads it, so calls on several streams may share a slot; the slot is marked as
    used by the current stream, so that evicting it cannot hand its memory to another
    allocation while a call on that stream still reads or writes it.  The strategy is
    AUTO until the one the first call picked has been read back (``_window_readback``):
    later calls then pass it and launch that strategy's kernels only."""
    t = dev.torch()
    cache = frame.__dict__.setdefault("_window_cache", OrderedDict())
    slot = cache.get(key)
    reuse = 1
    if slot is None:
        while len(cache) >= _WINDOW_CACHE_MAX:  # oldest first: a server sees many views
            cache.popitem(last=False)
        slot = cache[key] = _WindowSlot(
            t.zeros(_abi.WINDOW_RECORD_WORDS, dtype=t.int32, device=device))
        reuse = 0
    else:
        cache.move_to_end(key)
        if slot.strategy is None and slot.event is not None and slot.event.query():
            picked = int(slot.host[4])
            if picked in (_abi.STRATEGY_WINDOW, _abi.STRATEGY_TILES):
                slot.strategy = picked
    slot.rec.record_stream(t.cuda.current_stream(slot.rec.device))
    return slot, reuse, (_abi.STRATEGY_AUTO if slot.strategy is None else slot.strategy)


def _window_readback(slot):
    """After the call that picked a slot's record: start copying its strategy word to
    pinned host memory, without waiting (read by a later ``_window_slot``)."""
    if slot.event is not None:
        return
    t = dev.torch()
    slot.host = t.empty(_abi.WINDOW_RECORD_WORDS, dtype=t.int32, pin_memory=True)
    slot.host.copy_(slot.rec, non_blocking=True)
    slot.event = t.cuda.Event()
    slot.event.record()


def _grid_ref(frame, glyph, p, view):
    """The lattice of a frame sorted by a grid index (``serving.PointIndex``) when this
    pass can read it: the glyph's x / y are the indexed columns, no categorizer, linear
    axes.  None otherwise."""
    g = getattr(frame, "grid_index", None)
    if g is None or p.cat is not None or view["xlog"] or view["ylog"]:
        return None
    if (glyph.x, glyph.y) != (g["x"], g["y"]):
        return None
    r = _abi.GridRef()
    r.grid = g["grid"]
    r.x0, r.x1, r.y0, r.y1 = g["extent"]
    r.cell_start = dev.ptr(g["cell_start"])
    return r


def _points_call(lib, d, vw, planes_struct, mean, device, stream, grid=None):
    """dsx_points_agg (or, over an indexed frame, dsx_points_agg_indexed) with a
    workspace from torch's caching allocator."""
    t = dev.torch()
    ws = int(lib.dsx_points_indexed_workspace_size(d, vw, planes_struct, grid)) if grid else 0
    if grid is not None and ws > 0:  # a view the lattice tiles cannot take: plain call
        work = t.empty(ws, dtype=t.uint8, device=device)
        _abi.check(_timed(lambda: lib.dsx_points_agg_indexed(
            d, vw, planes_struct, dev.ptr(mean), grid, dev.ptr(work), ws, stream)),
            "dsx_points_agg_indexed")
        return
    ws = int(lib.dsx_points_workspace_size(d, vw, planes_struct))
    if ws == 0:
        raise _abi.DsxError("dsx_points_workspace_size rejected the call (invalid "
                            "descriptor or shape)")
    work = t.empty(ws, dtype=t.uint8, device=device)
    _abi.check(_timed(lambda: lib.dsx_points_agg(d, vw, planes_struct, dev.ptr(mean),
                                                 dev.ptr(work), ws, stream)),
               "dsx_points_agg")


def _pass_desc(frame, schema, glyph, view, p, device, catc):
    """The kernel descriptor of one pass over ``frame``: PointsDesc or LinesDesc, plus
    the (kept-alive) tensors it points to."""
    t = dev.torch()
    val = dev.as_float_column(frame[p.val]) if p.val is not None else None
    vdt = (_abi.F32 if val.dtype == t.float32 else _abi.F64) if val is not None else 0
    if isinstance(glyph, Point):
        x = dev.as_float_column(frame[glyph.x])
        y = dev.as_float_column(frame[glyph.y])
        if x.dtype != y.dtype:  # x and y share one dtype in the kernel
            x, y = x.to(t.float64), y.to(t.float64)
        d = _abi.PointsDesc()
        d.n = len(frame)
        d.x, d.y = dev.ptr(x), dev.ptr(y)
        d.xy_dtype = dev.dtype_code(x)
        d.val, d.val_dtype = dev.ptr(val), vdt
        d.cat = catc
        d.strategy = _abi.STRATEGY_AUTO if FORCE_STRATEGY is None else FORCE_STRATEGY
        return d, (x, y, val)
    lay = glyph.vertex_layout(frame)
    d = _abi.LinesDesc()
    d.nlines, d.nverts = lay["nlines"], lay["nverts"]
    d.x, d.y = dev.ptr(lay["x"]), dev.ptr(lay["y"])
    d.x_dtype, d.y_dtype = dev.dtype_code(lay["x"]), dev.dtype_code(lay["y"])
    d.x_line_stride, d.x_vert_stride = lay["xls"], lay["xvs"]
    d.y_line_stride, d.y_vert_stride = lay["yls"], lay["yvs"]
    d.axis = lay["axis"]
    # axis 0: a streamed chunk of a longer line carries its own plot_start
    d.plot_start = int(getattr(frame, "plot_start", 1))
    d.val, d.val_dtype = dev.ptr(val), vdt
    d.cat = catc
    return d, (lay, val)


def m2_pass(lib, frame, schema, glyph, view, p, ncells, mean, grid=None):
    """var / std: a second pass that accumulates every row's squared deviation from its
    cell's mean (``mean``: float64 plane); returns the m2 plane."""
    device = frame.device
    catc, catcol, _ = _categorizer(p.cat, frame, schema)
    d, keep = _pass_desc(frame, schema, glyph, view, p, device, catc)
    vw = _view_struct(view)
    stream = dev.stream_handle()
    m2 = _alloc(ncells, "f64", device)
    m2s = _abi.Planes()
    m2s.sum = dev.ptr(m2)
    if isinstance(glyph, Point):
        _points_call(lib, d, vw, m2s, mean, device, stream, grid)
    else:
        _fill(lib, m2, _abi.FILL_ZERO64)
        d.mean = dev.ptr(mean)
        _abi.check(lib.dsx_lines_agg(d, vw, m2s, stream), "dsx_lines_agg")
    return m2


def _fused_planes(p):
    """The planes a points pass writes when mean / var / std take the kernel's finalized
    mean plane: the sum and count(v) planes only where another leaf reads them."""
    if not any(isinstance(r, (rd.mean, rd.var, rd.std)) for r in p.leaves):
        return set(p.planes)
    need_sum = any(type(r) is rd.sum for r in p.leaves)
    need_countv = p.m2 or any(type(r) is rd.sum or (isinstance(r, (rd.count, rd.any)) and
                                                    r.column is not None) for r in p.leaves)
    planes = set(p.planes) | {"mean"}
    if not need_sum:
        planes.discard("sum")
    if not need_countv:
        planes.discard("countv")
    return planes


def run_pass(lib, frame, schema, glyph, view, p, H, W, with_m2=True, fuse_mean=True):
    """One kernel pass over ``frame``: its planes (and, for var / std, the m2 plane
    against the pass's own means unless ``with_m2`` is False).  Points passes with a
    mean / var / std leaf take the kernel's finalized mean plane (``fuse_mean``) instead
    of finalizing it from the sum and count(v) planes."""
    t = dev.torch()
    device = frame.device
    catc, catcol, ncat = _categorizer(p.cat, frame, schema)
    ncells = H * W * ncat
    val = dev.as_float_column(frame[p.val]) if p.val is not None else None
    f32 = val is not None and val.dtype == t.float32
    keyb = 4 if f32 else 8
    vw = _view_struct(view)
    stream = dev.stream_handle()
    points = isinstance(glyph, Point)
    if not points and not isinstance(glyph, _LineLike):
        raise NotImplementedError(f"glyph {type(glyph).__name__} is not supported")
    want = _fused_planes(p) if (points and fuse_mean) else set(p.planes)
    planes = {}
    ps = _abi.Planes()
    # points write every cell; lines accumulate into planes filled with the initial values
    for name, kind, init in (("count", "u32", _abi.FILL_ZERO32),
                             ("countv", "u32", _abi.FILL_ZERO32),
                             ("sum", "f64", _abi.FILL_ZERO64),
                             ("min", "u32" if f32 else "u64",
                              _abi.FILL_ONES32 if f32 else _abi.FILL_ONES64),
                             ("max", "u32" if f32 else "u64",
                              _abi.FILL_ZERO32 if f32 else _abi.FILL_ZERO64),
                             ("mean", "f64", None)):
        if name in want:
            planes[name] = _alloc(ncells, kind, device)
            if not points:
                _fill(lib, planes[name], init)
            setattr(ps, name, dev.ptr(planes[name]))
    d, keep = _pass_desc(frame, schema, glyph, view, p, device, catc)
    grid = None
    if points:
        grid = _grid_ref(frame, glyph, p, view)
        if grid is None and getattr(frame, "persistent", False):
            key = (glyph.x, glyph.y, p.val, None if p.cat is None else p.cat._key(),
                   tuple(sorted(p.planes)),
                   W, H, tuple(view[k] for k in ("xmin", "xmax", "ymin", "ymax")))
            slot, reuse, strategy = _window_slot(frame, key, device)
            d.window, d.reuse_window = dev.ptr(slot.rec), reuse
            if FORCE_STRATEGY is None:
                d.strategy = strategy
            _points_call(lib, d, vw, ps, None, device, stream, grid)
            _window_readback(slot)
        else:
            _points_call(lib, d, vw, ps, None, device, stream, grid)
    else:
        _abi.check(_timed(lambda: lib.dsx_lines_agg(d, vw, ps, stream)), "dsx_lines_agg")
    if p.m2 and with_m2:
        if "mean" in planes:
            mean = planes["mean"]
        else:
            mean = t.empty(ncells, dtype=t.float64, device=device)
            _finalize(lib, _abi.FIN_MEAN, keyb, ncells, planes, mean)
        planes["m2"] = m2_pass(lib, frame, schema, glyph, view, p, ncells, mean, grid)
    return PassResult(ncat, keyb, planes)


def _finalize(lib, mode, keyb, ncells, planes, out, key=None, counts=None):
    f = _abi.FinalizeDesc()
    f.mode = mode
    f.key_bytes = keyb
    f.ncells = ncells
    if mode == _abi.FIN_ANY:
        f.count = dev.ptr(counts)
    elif mode in (_abi.FIN_MIN, _abi.FIN_MAX):
        f.key = dev.ptr(key)
    else:
        f.countv = dev.ptr(planes["countv"])
        f.sum = dev.ptr(planes["m2"] if mode in (_abi.FIN_VAR, _abi.FIN_STD) else planes["sum"])
    f.out = dev.ptr(out)
    _abi.check(lib.dsx_finalize(f, dev.stream_handle()), "dsx_finalize")


def finalize_leaf(lib, red, res, H, W):
    """Device tensor holding the reference's finalized array of one reduction."""
    t = dev.torch()
    inner = red.reduction if isinstance(red, rd.by) else red
    shape = (H, W, res.ncat) if isinstance(red, rd.by) else (H, W)
    ncells = H * W * res.ncat
    device = next(iter(res.planes.values())).device
    if isinstance(inner, rd.count):
        plane = res.planes["count" if inner.column is None else "countv"]
        return plane.view(t.uint32).reshape(shape)
    if isinstance(inner, rd.any):
        counts = res.planes["count" if inner.column is None else "countv"]
        out = t.empty(ncells, dtype=t.uint8, device=device)
        _finalize(lib, _abi.FIN_ANY, res.keyb, ncells, res.planes, out, counts=counts)
        return out.view(t.bool).reshape(shape)
    mode = inner.finalize_mode
    if mode == _abi.FIN_MEAN and "mean" in res.planes:  # finalized by the kernel
        return res.planes["mean"].reshape(shape)
    out = t.empty(ncells, dtype=t.float64, device=device)
    key = res.planes.get("min") if mode == _abi.FIN_MIN else res.planes.get("max")
    _finalize(lib, mode, res.keyb, ncells, res.planes, out, key=key)
    return out.reshape(shape)


# ---- the pipeline (data_libraries/pandas.py:23-66) ---------------------------------

def _view_struct(view):
    v = _abi.View()
    for k in ("sx", "tx", "sy", "ty", "xmin", "xmax", "ymin", "ymax"):
        setattr(v, k, float(view[k]))
    v.xxmax, v.yymax = int(view["xxmax"]), int(view["yymax"])
    v.width, v.height = int(view["W"]), int(view["H"])
    v.xlog, v.ylog = int(view["xlog"]), int(view["ylog"])
    return v


def make_view(canvas, x_range, y_range):
    """Scale / translate, bounds and line-snapping limits of one call."""
    W, H = canvas.plot_width, canvas.plot_height
    x_st = canvas.x_axis.compute_scale_and_translate(x_range, W)
    y_st = canvas.y_axis.compute_scale_and_translate(y_range, H)
    sx, tx = x_st
    sy, ty = y_st
    # map_onto_pixel_snap (line.py:716-717): round(mapper(xmax)*sx + tx), the scalar
    # types of the operands kept (a float32 range stays float32), half to even.
    xxmax = int(np.rint(canvas.x_axis.mapper(x_range[1]) * sx + tx))
    yymax = int(np.rint(canvas.y_axis.mapper(y_range[1]) * sy + ty))
    return dict(sx=sx, tx=tx, sy=sy, ty=ty, xmin=x_range[0], xmax=x_range[1],
                ymin=y_range[0], ymax=y_range[1], xxmax=xxmax, yymax=yymax, W=W, H=H,
                xlog=canvas.x_axis.log, ylog=canvas.y_axis.log, x_st=x_st, y_st=y_st)


def _packed_columns(frame, agg):
    """The float32 columns of where(max / min) and the max / min row-index selectors: they
    take dsx_points_rowindex's packed pass (no max / min plane)."""
    t = dev.torch()
    out = set()
    for _, red in rd.leaves(agg):
        inner = red.reduction if isinstance(red, rd.by) else red
        if not getattr(inner, "row_index", False):
            continue
        sel = inner.selector if isinstance(inner, rd.where) else inner
        if isinstance(sel, (rd.max, rd.min)) and frame[sel.column].dtype == t.float32:
            out.add(sel.column)
    return frozenset(out)


def row_index_leaf(lib, frame, schema, glyph, view, red, passes, results, H, W):
    """The finalized array of a row-index leaf: the selected row of every cell
    (dsx_points_rowindex) and, with a lookup column, that row's value (dsx_take)."""
    t = dev.torch()
    if not isinstance(glyph, Point):
        raise NotImplementedError(f"{type(red).__name__} is only supported for points on "
                                  "the MI355X backend")
    cat = red.categorizer if isinstance(red, rd.by) else None
    inner = red.reduction if isinstance(red, rd.by) else red
    sel = inner.selector if isinstance(inner, rd.where) else inner
    if isinstance(inner, rd.where):
        lookup = inner.column
    elif isinstance(inner, (rd._first_or_last, rd.FloatingNReduction)):
        lookup = inner.column  # None for the row-index selectors: the index itself
    else:
        lookup = None  # the row-index selectors return the index itself
    device = frame.device
    catc, catcol, ncat = _categorizer(cat, frame, schema)
    ncells = H * W * ncat
    if isinstance(sel, rd.FloatingNReduction):
        return _n_select_leaf(lib, frame, schema, glyph, view, red, sel, lookup, cat, catc,
                              ncat, H, W)
    r = _abi.RowIndexDesc()
    key = None
    valcol = sel.column
    packed = False
    if isinstance(sel, (rd.max, rd.min)):
        r.mode = _abi.RI_ARGMAX if isinstance(sel, rd.max) else _abi.RI_ARGMIN
        r.nan_check = 1
        packed = frame[sel.column].dtype == t.float32
        if not packed:  # float64 values: the key plane of the max / min pass
            res = results[(sel.column, cat)]
            key = res.planes["max" if isinstance(sel, rd.max) else "min"]
            r.key = dev.ptr(key)
    else:
        r.mode = sel.ri_mode
        r.nan_check = 1 if valcol is not None else 0
    index = t.empty(ncells, dtype=t.int64, device=device)
    r.index = dev.ptr(index)
    p = Pass(valcol, cat)
    d, keep = _pass_desc(frame, schema, glyph, view, p, device, catc)
    if key is not None and d.val_dtype != (_abi.F32 if res.keyb == 4 else _abi.F64):
        raise _abi.DsxError("key plane width does not match the value column")
    vw = _view_struct(view)
    if key is None and getattr(frame, "persistent", False):
        # the LDS-window pass keeps its hot window per (frame, canvas, view) like the
        # points pass (a window of one 32-bit row per cell, or 64-bit packed cells)
        wkey = ("rowindex_packed" if packed else "rowindex", glyph.x, glyph.y, valcol,
                None if cat is None else cat._key(),
                W, H, tuple(view[k] for k in ("xmin", "xmax", "ymin", "ymax")))
        slot, reuse, _ = _window_slot(frame, wkey, device)
        d.window, d.reuse_window = dev.ptr(slot.rec), reuse
    ws = 0 if key is not None else int(lib.dsx_points_rowindex_workspace_size(d, vw, r))
    work = t.empty(ws, dtype=t.uint8, device=device) if ws else None
    _abi.check(lib.dsx_points_rowindex(d, vw, r, dev.ptr(work), ws, dev.stream_handle()),
               "dsx_points_rowindex")
    shape = (H, W, ncat) if isinstance(red, rd.by) else (H, W)
    if lookup is None:
        return index.reshape(shape)
    col = frame[lookup]
    if col.dtype == t.bool:
        col = col.to(t.uint8)
    tk = _abi.TakeDesc()
    tk.ncells = ncells
    tk.index = dev.ptr(index)
    tk.col = dev.ptr(col)
    tk.dtype = dev.dtype_code(col)
    out = t.empty(ncells, dtype=t.float64, device=device)
    tk.out = dev.ptr(out)
    _abi.check(lib.dsx_take(tk, dev.stream_handle()), "dsx_take")
    return out.reshape(shape)


def _take(lib, index, col, device):
    """float64 values of ``col`` at the rows of ``index`` (NaN where -1)."""
    t = dev.torch()
    if col.dtype == t.bool:
        col = col.to(t.uint8)
    tk = _abi.TakeDesc()
    tk.ncells = int(index.numel())
    tk.index = dev.ptr(index)
    tk.col = dev.ptr(col)
    tk.dtype = dev.dtype_code(col)
    out = t.empty(index.numel(), dtype=t.float64, device=device)
    tk.out = dev.ptr(out)
    _abi.check(lib.dsx_take(tk, dev.stream_handle()), "dsx_take")
    return out


def _n_select_leaf(lib, frame, schema, glyph, view, red, sel, lookup, cat, catc, ncat, H, W):
    """max_n / min_n / first_n / last_n / the n row-index selectors (and where over
    them): the rows of each cell's n-slot list (dsx_points_nselect), then the lookup."""
    t = dev.torch()
    device = frame.device
    n = sel.n
    r = _abi.NSelectDesc()
    r.mode = sel.ns_mode
    r.n = n
    r.nan_check = 1 if sel.column is not None else 0
    index = t.empty(H * W * ncat * n, dtype=t.int64, device=device)
    r.index = dev.ptr(index)
    p = Pass(sel.column, cat)
    d, keep = _pass_desc(frame, schema, glyph, view, p, device, catc)
    vw = _view_struct(view)
    ws = int(lib.dsx_points_nselect_workspace_size(d, vw, r))
    if ws == 0:
        raise _abi.DsxError("dsx_points_nselect_workspace_size rejected the call")
    work = t.empty(ws, dtype=t.uint8, device=device)
    _abi.check(lib.dsx_points_nselect(d, vw, r, dev.ptr(work), ws, dev.stream_handle()),
               "dsx_points_nselect")
    shape = (H, W, ncat, n) if isinstance(red, rd.by) else (H, W, n)
    if lookup is None:
        return index.reshape(shape)
    return _take(lib, index, frame[lookup], device).reshape(shape)


def aggregate(frame, schema, canvas, glyph, agg):
    """Run the aggregation on the device; returns (results, view, x_range, y_range)
    where ``results`` is ``[(name, reduction, device tensor)]``."""
    dev.require_gpu()
    lib = _abi.load()
    x_range = canvas.x_range or glyph.compute_x_bounds(frame)
    y_range = canvas.y_range or glyph.compute_y_bounds(frame)
    canvas.validate_ranges(x_range, y_range)
    view = make_view(canvas, x_range, y_range)
    H, W = view["H"], view["W"]
    passes, leaves = plan(agg, _packed_columns(frame, agg))
    results = {key: run_pass(lib, frame, schema, glyph, view, p, H, W)
               for key, p in passes.items()}
    out = []
    for name, red, key in leaves:
        if key[0] == "rowindex":
            arr = row_index_leaf(lib, frame, schema, glyph, view, red, passes, results, H, W)
        else:
            arr = finalize_leaf(lib, red, results[key], H, W)
        out.append((name, red, arr))
    return out, view, x_range, y_range


def _wrap(red, data, canvas, glyph, view, x_range, y_range, schema, to_host, name=None):
    x_axis = canvas.x_axis.compute_index(view["x_st"], view["W"])
    y_axis = canvas.y_axis.compute_index(view["y_st"], view["H"])
    coords = {glyph.x_label: x_axis, glyph.y_label: y_axis}
    dims = [glyph.y_label, glyph.x_label]
    if isinstance(red, rd.by):
        dims.append(red.cat_column)
        coords[red.cat_column] = red.categorizer.categories(schema)
    inner = red.reduction if isinstance(red, rd.by) else red
    sel = inner.selector if isinstance(inner, rd.where) else inner
    if isinstance(sel, rd.FloatingNReduction):
        # FloatingNReduction._add_finalize_kwargs (reductions.py:1467-1478)
        dims.append("n")
        coords["n"] = np.arange(sel.n)
    attrs = dict(x_range=x_range, y_range=y_range)
    if to_host:
        data = data.cpu().numpy()
        if _xr is not None:
            return _xr.DataArray(data, coords=coords, dims=dims, attrs=attrs, name=name)
    return xrlite.DataArray(data, coords=coords, dims=dims, attrs=attrs, name=name)


def pipeline(source, schema, canvas, glyph, agg, *, antialias=False, to_host=None):
    """``bypixel.pipeline`` implementation for pandas and device sources."""
    if antialias:
        raise NotImplementedError("antialiased lines are not supported by the MI355X backend")
    if isinstance(source, pd.DataFrame):
        dev.require_gpu()
        frame = DeviceFrame.from_pandas(source)
        frame.persistent = False  # uploaded for this call only: keep no tuning state
        to_host = True if to_host is None else to_host
    else:
        frame = source
        to_host = False if to_host is None else to_host
    res, view, x_range, y_range = aggregate(frame, schema, canvas, glyph, agg)
    return wrap_results(agg, res, canvas, glyph, view, x_range, y_range, schema, to_host)


def wrap_results(agg, res, canvas, glyph, view, x_range, y_range, schema, to_host):
    """The DataArray (or, for ``summary``, Dataset) of finalized device tensors
    (``pandas.py:61-66``, ``compiler.py:510-536``)."""
    if isinstance(agg, rd.summary):
        data = {name: _wrap(red, t_, canvas, glyph, view, x_range, y_range, schema, to_host,
                            name=name) for name, red, t_ in res}
        first = agg.keys[0]
        attrs = {a: data[first].attrs[a] for a in ("x_range", "y_range")}
        if to_host and _xr is not None:
            return _xr.Dataset(data, attrs=attrs)
        return xrlite.Dataset(data, attrs=attrs)
    (_, red, t_), = res
    return _wrap(red, t_, canvas, glyph, view, x_range, y_range, schema, to_host)


bypixel.pipeline.register(pd.DataFrame)(pipeline)
bypixel.pipeline.register(DeviceFrame)(pipeline)

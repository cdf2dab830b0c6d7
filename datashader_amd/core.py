"""Canvas API of the points / line path (datashader's ``core.py``).

``Canvas.points`` and ``Canvas.line`` keep the reference's signatures, argument meaning,
defaults and validation errors (``core.py:177-478``); ``bypixel`` validates and
dispatches on the source type through ``bypixel.pipeline`` exactly like
``core.py:1334-1446``.  Two source types are registered by ``backend``: pandas
DataFrames (columns uploaded per call, result returned as numpy) and ``DeviceFrame``
(columns resident in HBM, result left on the device).
"""
from __future__ import annotations

import math
import warnings
from inspect import getmro
from numbers import Number

import numpy as np

from . import reductions as rd
from .glyphs import (LineAxis0, LineAxis0Multi, LinesAxis1, LinesAxis1XConstant,
                     LinesAxis1YConstant, LinesAxis1Ragged, Point)


class Axis:
    """Linear transformation ``s * mapper(v) + t`` from data to pixel space."""

    log = False

    def compute_scale_and_translate(self, range, n):
        """``core.py:62-81``: ``s = n / (end - start)``, ``t = -start * s``."""
        start, end = map(self.mapper, range)
        s = n / (end - start)
        t = -start * s
        return s, t

    def compute_index(self, st, n):
        """``core.py:83-99``: pixel-centre coordinates."""
        px = np.arange(n) + 0.5
        s, t = st
        return self.inverse_mapper((px - t) / s)

    def validate(self, range):
        pass


class LinearAxis(Axis):
    @staticmethod
    def mapper(val):
        return val

    @staticmethod
    def inverse_mapper(val):
        return val


class LogAxis(Axis):
    """Base-10 logarithmic axis (``core.py:127-145``)."""

    log = True

    @staticmethod
    def mapper(val):
        return math.log10(float(val))

    @staticmethod
    def inverse_mapper(val):
        return 10 ** val

    def validate(self, range):
        if range is None:
            return
        if range[0] <= 0 or range[1] <= 0:
            raise ValueError("Range values must be >0 for logarithmic axes")


_axis_lookup = {"linear": LinearAxis(), "log": LogAxis()}


def validate_xy_or_geometry(glyph, x, y, geometry):
    """``core.py:151-160``."""
    if (geometry is None and (x is None or y is None) or
            geometry is not None and (x is not None or y is not None)):
        raise ValueError(f"""
{glyph} coordinates may be specified by providing both the x and y arguments, or by
providing the geometry argument. Received:
    x: {x!r}
    y: {y!r}
    geometry: {geometry!r}
""")


class Canvas:
    """The space in which to bin (``core.py:163-185``).

    Parameters
    ----------
    plot_width, plot_height : int
        Width and height of the output aggregate in pixels.
    x_range, y_range : tuple, optional
        Inclusive ``[min, max]`` bounds along each axis; computed from the data when
        omitted.
    x_axis_type, y_axis_type : 'linear' (default) or 'log'
    """

    def __init__(self, plot_width=600, plot_height=600, x_range=None, y_range=None,
                 x_axis_type="linear", y_axis_type="linear"):
        self.plot_width = plot_width
        self.plot_height = plot_height
        self.x_range = None if x_range is None else tuple(x_range)
        self.y_range = None if y_range is None else tuple(y_range)
        self.x_axis = _axis_lookup[x_axis_type]
        self.y_axis = _axis_lookup[y_axis_type]

    def points(self, source, x=None, y=None, agg=None, geometry=None):
        """Aggregate rows as points (``core.py:187-232``); default ``agg`` is ``count()``."""
        validate_xy_or_geometry("Point", x, y, geometry)
        if agg is None:
            agg = rd.count()
        if geometry is not None:
            raise NotImplementedError("geometry columns (spatialpandas / geopandas) are not "
                                      "supported by the MI355X backend")
        return bypixel(source, self, Point(x, y), agg)

    def line(self, source, x=None, y=None, agg=None, axis=0, geometry=None, line_width=0,
             antialias=False):
        """Aggregate rows as lines (``core.py:234-478``); default ``agg`` is ``any()``.

        Glyph selection follows the reference: axis 0 with scalar x / y -> one line;
        axis 0 with lists -> one line per column pair; axis 1 with lists -> one line per
        row; axis 1 with an ndarray x (or y) -> one line per row sharing that coordinate.
        Only zero-width lines (``line_width=0``, ``antialias=False``) are supported.
        """
        validate_xy_or_geometry("Line", x, y, geometry)
        if agg is None:
            agg = rd.any()
        if line_width is None:
            line_width = 0
        if antialias and line_width != 0:
            raise ValueError(
                "Do not specify values for both the line_width and \n"
                "antialias keyword arguments; use line_width instead.")
        if antialias:
            line_width = 1.0
        if geometry is not None:
            raise NotImplementedError("geometry columns (spatialpandas / geopandas) are not "
                                      "supported by the MI355X backend")
        orig_x, orig_y = x, y
        x, y = _broadcast_column_specifications(x, y)
        if axis == 0:
            if isinstance(x, (Number, str)) and isinstance(y, (Number, str)):
                glyph = LineAxis0(x, y)
            elif isinstance(x, (list, tuple)) and isinstance(y, (list, tuple)):
                glyph = LineAxis0Multi(tuple(x), tuple(y))
            else:
                raise ValueError(f"""
Invalid combination of x and y arguments to Canvas.line when axis=0.
    Received:
        x: {repr(orig_x)}
        y: {repr(orig_y)}
See docstring for more information on valid usage""")
        elif axis == 1:
            if isinstance(x, (list, tuple)) and isinstance(y, (list, tuple)):
                glyph = LinesAxis1(tuple(x), tuple(y))
            elif isinstance(x, np.ndarray) and isinstance(y, (list, tuple)):
                glyph = LinesAxis1XConstant(x, tuple(y))
            elif isinstance(x, (list, tuple)) and isinstance(y, np.ndarray):
                glyph = LinesAxis1YConstant(tuple(x), y)
            elif isinstance(x, (Number, str)) and isinstance(y, (Number, str)):
                glyph = LinesAxis1Ragged(x, y)
            else:
                raise ValueError(f"""
Invalid combination of x and y arguments to Canvas.line when axis=1.
    Received:
        x: {repr(orig_x)}
        y: {repr(orig_y)}
See docstring for more information on valid usage""")
        else:
            raise ValueError(f"""
The axis argument to Canvas.line must be 0 or 1
    Received: {axis}""")
        if line_width > 0:
            raise NotImplementedError("antialiased lines (line_width > 0) are not supported "
                                      "by the MI355X backend")
        return bypixel(source, self, glyph, agg)

    # -- validation (core.py:1272-1283) -----------------------------------------------
    def validate_ranges(self, x_range, y_range):
        self.x_axis.validate(x_range)
        self.y_axis.validate(y_range)

    def validate_size(self, width, height):
        if width <= 0 or height <= 0:
            raise ValueError("Invalid size: plot_width and plot_height must be bigger than 0")

    def validate(self):
        self.validate_ranges(self.x_range, self.y_range)
        self.validate_size(self.plot_width, self.plot_height)


class Dispatcher:
    """Single dispatch on the type of the first argument (``utils.py:95-122``)."""

    def __init__(self):
        self._lookup = {}

    def register(self, typ, func=None):
        if func is None:
            return lambda f: self.register(typ, f)
        if isinstance(typ, tuple):
            for t in typ:
                self.register(t, func)
        else:
            self._lookup[typ] = func
        return func

    def __call__(self, head, *rest, **kwargs):
        lk = self._lookup
        typ = type(head)
        if typ in lk:
            return lk[typ](head, *rest, **kwargs)
        for cls in getmro(typ)[1:]:
            if cls in lk:
                return lk[cls](head, *rest, **kwargs)
        raise TypeError(f"No dispatch for {typ} type")


def _cols_to_keep(columns, glyph, agg):
    """Columns the glyph and the reductions read (``core.py:1408-1430``)."""
    keep = {col: False for col in columns}
    for col in glyph.required_columns():
        keep[col] = True

    def recurse(a):
        if isinstance(a, rd.summary):
            for sub in a.values:
                recurse(sub)
        elif isinstance(a, rd.by):
            for c in a.columns:
                if c is not None:
                    keep[c] = True
            recurse(a.reduction)
        else:
            if a.column is not None:
                keep[a.column] = True
            if isinstance(a, rd.where):
                recurse(a.selector)

    recurse(agg)
    return [c for c, k in keep.items() if k]


def bypixel(source, canvas, glyph, agg, *, antialias=False):
    """Aggregate ``source`` into a canvas-shaped grid (``core.py:1334-1359``)."""
    from .frame import DeviceFrame, Schema
    import pandas as pd
    if isinstance(source, pd.DataFrame):
        cols = _cols_to_keep(source.columns, glyph, agg)
        if len(cols) < len(source.columns):
            source = source[cols]
        schema = Schema.from_pandas(source)
    elif isinstance(source, DeviceFrame) or getattr(source, "streams_in_chunks", False):
        schema = source.schema
    else:
        raise ValueError("source must be a pandas DataFrame, a datashader_amd.DeviceFrame "
                         "or a datashader_amd.ChunkedFrame")
    glyph.validate(schema)
    agg.validate(schema)
    canvas.validate()
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", r"All-NaN (slice|axis) encountered")
        return bypixel.pipeline(source, schema, canvas, glyph, agg, antialias=antialias)


def _broadcast_column_specifications(*args):
    """``core.py:1433-1443``."""
    lengths = {len(a) for a in args if isinstance(a, (list, tuple))}
    if len(lengths) != 1:
        return args
    n = lengths.pop()
    return tuple((arg,) * n if isinstance(arg, (Number, str)) else arg for arg in args)


bypixel.pipeline = Dispatcher()

"""Glyphs of the points / line path (datashader's ``glyphs/points.py`` and
``glyphs/line.py``): validation, axis labels, autorange bounds (on the device) and the
vertex layout each glyph hands to the line kernel."""
from __future__ import annotations

import numpy as np

from . import _device as dev


def _isreal(kind):
    return kind in ("f4", "f8", "i1", "i2", "i4", "i8", "u1", "u2", "u4", "u8")


class Glyph:
    maybe_expand_bounds = staticmethod(dev.maybe_expand_bounds)

    def __eq__(self, other):
        return type(self) is type(other) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())

    def _key(self):
        return (self.x if not isinstance(self.x, np.ndarray) else tuple(self.x.tolist()),
                self.y if not isinstance(self.y, np.ndarray) else tuple(self.y.tolist()))


class _PointLike(Glyph):
    """Shared by Point and LineAxis0 (``glyphs/points.py:114-167``)."""

    def __init__(self, x, y):
        self.x = x
        self.y = y

    def validate(self, schema):
        if not _isreal(schema[str(self.x)]):
            raise ValueError("x must be real")
        elif not _isreal(schema[str(self.y)]):
            raise ValueError("y must be real")

    @property
    def x_label(self):
        return self.x

    @property
    def y_label(self):
        return self.y

    def required_columns(self):
        return [self.x, self.y]

    def compute_x_bounds(self, frame):
        return self.maybe_expand_bounds(dev.bounds(frame[self.x]))

    def compute_y_bounds(self, frame):
        return self.maybe_expand_bounds(dev.bounds(frame[self.y]))


class Point(_PointLike):
    """A point at (x, y); rows on the upper bound map into the last bin
    (``glyphs/points.py:170-242``)."""


class _LineLike:
    """Line glyphs: ``vertex_layout(frame)`` returns the strided description the line
    kernel walks (nlines, nverts, x/y base tensors and element strides, axis)."""


class LineAxis0(_PointLike, _LineLike):
    """One line through all rows (``glyphs/line.py:59-103``)."""

    def vertex_layout(self, frame):
        x = dev.as_float_column(frame[self.x])
        y = dev.as_float_column(frame[self.y])
        return dict(nlines=1, nverts=len(frame), x=x, y=y, xls=0, xvs=x.stride(0), yls=0,
                    yvs=y.stride(0), axis=0)


def _bounds_of(frame, names):
    mins, maxs = zip(*[dev.bounds(frame[n]) for n in names])
    return dev.maybe_expand_bounds((min(mins), max(maxs)))


class LineAxis0Multi(_PointLike, _LineLike):
    """One line per (x, y) column pair, along the rows (``glyphs/line.py:106-196``)."""

    def validate(self, schema):
        if not all(_isreal(schema[str(c)]) for c in self.x):
            raise ValueError("x columns must be real")
        elif not all(_isreal(schema[str(c)]) for c in self.y):
            raise ValueError("y columns must be real")
        if len(self.x) != len(self.y):
            raise ValueError(
                f"x and y coordinate lengths do not match: {len(self.x)} != {len(self.y)}")

    @property
    def x_label(self):
        return "x"

    @property
    def y_label(self):
        return "y"

    def required_columns(self):
        return list(self.x) + list(self.y)

    def compute_x_bounds(self, frame):
        return _bounds_of(frame, self.x)

    def compute_y_bounds(self, frame):
        return _bounds_of(frame, self.y)

    def vertex_layout(self, frame):
        xb, xcs, xrs = dev.column_matrix(frame, self.x)
        yb, ycs, yrs = dev.column_matrix(frame, self.y)
        return dict(nlines=len(self.x), nverts=len(frame), x=xb, y=yb, xls=xcs, xvs=xrs,
                    yls=ycs, yvs=yrs, axis=0)


class LinesAxis1(_PointLike, _LineLike):
    """One line per row through the listed columns (``glyphs/line.py:198-306``)."""

    def validate(self, schema):
        if not all(_isreal(schema[str(c)]) for c in self.x):
            raise ValueError("x columns must be real")
        elif not all(_isreal(schema[str(c)]) for c in self.y):
            raise ValueError("y columns must be real")
        if len({schema[str(c)] for c in self.x}) > 1:
            raise ValueError("x columns must have the same data type")
        if len({schema[str(c)] for c in self.y}) > 1:
            raise ValueError("y columns must have the same data type")
        if len(self.x) != len(self.y):
            raise ValueError(
                f"x and y coordinate lengths do not match: {len(self.x)} != {len(self.y)}")

    @property
    def x_label(self):
        return "x"

    @property
    def y_label(self):
        return "y"

    def required_columns(self):
        return list(self.x) + list(self.y)

    def compute_x_bounds(self, frame):
        return _bounds_of(frame, self.x)

    def compute_y_bounds(self, frame):
        return _bounds_of(frame, self.y)

    def vertex_layout(self, frame):
        xb, xcs, xrs = dev.column_matrix(frame, self.x)
        yb, ycs, yrs = dev.column_matrix(frame, self.y)
        # line = row, vertex = column
        return dict(nlines=len(frame), nverts=len(self.x), x=xb, y=yb, xls=xrs, xvs=xcs,
                    yls=yrs, yvs=ycs, axis=1)


def _const_tensor(arr, device):
    t = dev.torch()
    a = np.asarray(arr)
    if a.dtype != np.float32:
        a = a.astype(np.float64)
    return t.from_numpy(np.ascontiguousarray(a)).to(device)


class LinesAxis1XConstant(LinesAxis1):
    """One line per row; every line shares the x coordinates ``x`` (an ndarray)
    (``glyphs/line.py:309-380``)."""

    def validate(self, schema):
        if not all(_isreal(schema[str(c)]) for c in self.y):
            raise ValueError("y columns must be real")
        if len({schema[str(c)] for c in self.y}) > 1:
            raise ValueError("y columns must have the same data type")
        if len(self.x) != len(self.y):
            raise ValueError(
                f"x and y coordinate lengths do not match: {len(self.x)} != {len(self.y)}")

    def required_columns(self):
        return list(self.y)

    def compute_x_bounds(self, *args):
        return self.maybe_expand_bounds((np.nanmin(self.x), np.nanmax(self.x)))

    def vertex_layout(self, frame):
        yb, ycs, yrs = dev.column_matrix(frame, self.y)
        xt = _const_tensor(self.x, yb.device)
        return dict(nlines=len(frame), nverts=len(self.y), x=xt, y=yb, xls=0, xvs=1,
                    yls=yrs, yvs=ycs, axis=1)


class LinesAxis1YConstant(LinesAxis1):
    """One line per row; every line shares the y coordinates ``y`` (an ndarray)
    (``glyphs/line.py:383-455``)."""

    def validate(self, schema):
        if not all(_isreal(schema[str(c)]) for c in self.x):
            raise ValueError("x columns must be real")
        if len({schema[str(c)] for c in self.x}) > 1:
            raise ValueError("x columns must have the same data type")
        if len(self.x) != len(self.y):
            raise ValueError(
                f"x and y coordinate lengths do not match: {len(self.x)} != {len(self.y)}")

    def required_columns(self):
        return list(self.x)

    def compute_y_bounds(self, *args):
        return self.maybe_expand_bounds((np.nanmin(self.y), np.nanmax(self.y)))

    def vertex_layout(self, frame):
        xb, xcs, xrs = dev.column_matrix(frame, self.x)
        yt = _const_tensor(self.y, xb.device)
        return dict(nlines=len(frame), nverts=len(self.x), x=xb, y=yt, xls=xrs, xvs=xcs,
                    yls=0, yvs=1, axis=1)


class LinesAxis1Ragged(_PointLike, _LineLike):
    """One line per row of two ragged columns (``glyphs/line.py:457-523``): row i joins
    the first ``min(len(x_i), len(y_i))`` vertices (``line.py:1569-1586``).  The rows are
    laid out as one NaN-padded ``[rows, longest]`` matrix per coordinate on the device; a
    NaN vertex ends a line's segments, so the padding draws nothing."""

    def validate(self, schema):
        if schema[str(self.x)] != "ragged":
            raise ValueError("x must be a RaggedArray")
        elif schema[str(self.y)] != "ragged":
            raise ValueError("y must be a RaggedArray")

    def required_columns(self):
        return [self.x, self.y]

    def compute_x_bounds(self, frame):
        # over every stored value, as the reference bounds the flat array
        return self.maybe_expand_bounds(dev.bounds(frame[self.x].flat))

    def compute_y_bounds(self, frame):
        return self.maybe_expand_bounds(dev.bounds(frame[self.y].flat))

    def vertex_layout(self, frame):
        t = dev.torch()
        xc, yc = frame[self.x], frame[self.y]
        lx = xc.starts[1:] - xc.starts[:-1]
        ly = yc.starts[1:] - yc.starts[:-1]
        lens = t.minimum(lx, ly)
        nrows = int(lens.numel())
        width = max(2, int(lens.max().item()) if nrows else 2)
        dtype = xc.flat.dtype if xc.flat.dtype == yc.flat.dtype == t.float32 else t.float64

        def padded(col, lengths):
            out = t.full((nrows, width), float("nan"), dtype=dtype, device=col.flat.device)
            if col.flat.numel():
                row = t.repeat_interleave(t.arange(nrows, device=col.flat.device), lengths)
                pos = t.arange(col.flat.numel(), device=col.flat.device) - col.starts[row]
                keep = pos < lens[row]
                out[row[keep], pos[keep]] = col.flat[keep].to(dtype)
            return out

        X, Y = padded(xc, lx), padded(yc, ly)
        return dict(nlines=nrows, nverts=width, x=X, y=Y, xls=width, xvs=1, yls=width, yvs=1,
                    axis=1)

"""Translate datashader's own glyph / reduction / canvas objects into this package's, so
that datashader can register this backend with ``bypixel.pipeline`` (see the
integration notes).  Matching is by class name and public attributes, so datashader
itself is not imported here.

The attributes read are the reference's:

* glyphs: ``x``, ``y`` (``glyphs/points.py:114-167``, ``glyphs/line.py:59-523``) and
  ``antialiased`` (``line.py:21-56``);
* reductions: ``column`` (``reductions.py:309-320``; the row-index special column
  ``SpecialColumn.RowIndex``, ``reductions.py:45-55``, becomes ``None``), ``n``
  (``FloatingNReduction``, ``1457-1470``), ``selector`` (``where``, ``1842-1880``),
  ``categorizer`` / ``reduction`` (``by``, ``696-823``), ``keys`` / ``values``
  (``summary``, ``2169-2233``);
* categorizers: ``column``, ``modulo``, ``offset``, ``bin0``, ``binsize``, ``nbins``,
  ``bin_under``, ``bin_over`` (``reductions.py:143-259``).
"""
from __future__ import annotations

from . import glyphs as g
from . import reductions as rd
from .core import Canvas

_GLYPHS = {"Point": g.Point, "LineAxis0": g.LineAxis0, "LineAxis0Multi": g.LineAxis0Multi,
           "LinesAxis1": g.LinesAxis1, "LinesAxis1XConstant": g.LinesAxis1XConstant,
           "LinesAxis1YConstant": g.LinesAxis1YConstant,
           "LinesAxis1Ragged": g.LinesAxis1Ragged}

# reductions built from (column); count / any / sum also take self_intersect
_COLUMN = {"count": rd.count, "any": rd.any, "sum": rd.sum, "mean": rd.mean, "min": rd.min,
           "max": rd.max, "var": rd.var, "std": rd.std, "first": rd.first, "last": rd.last}
# reductions built from (column, n)
_N = {"max_n": rd.max_n, "min_n": rd.min_n, "first_n": rd.first_n, "last_n": rd.last_n}
# row-index selectors: no column (SpecialColumn.RowIndex in the reference)
_ROW_INDEX = {"_min_row_index": rd._min_row_index, "_max_row_index": rd._max_row_index}
_ROW_INDEX_N = {"_min_n_row_index": rd._min_n_row_index,
                "_max_n_row_index": rd._max_n_row_index}


def glyph_from(glyph):
    name = type(glyph).__name__
    if name not in _GLYPHS:
        raise NotImplementedError(f"glyph {name} is not supported by the MI355X backend")
    if getattr(glyph, "antialiased", False):
        raise NotImplementedError("antialiased lines are not supported by the MI355X backend")
    return _GLYPHS[name](glyph.x, glyph.y)


def _categorizer(c):
    name = type(c).__name__
    if name == "category_codes":
        return rd.category_codes(c.column)
    if name == "category_binning":
        out = rd.category_binning(c.column, c.bin0, c.bin0 + c.binsize * c.nbins, c.nbins)
        out.bin_under, out.bin_over, out.binsize = c.bin_under, c.bin_over, c.binsize
        return out
    if name == "category_modulo":
        return rd.category_modulo(c.column, c.modulo, c.offset)
    raise NotImplementedError(f"categorizer {name} is not supported")


def _column(col):
    """A reduction's column, with the reference's row-index special column as None."""
    if col is not None and type(col).__name__ == "SpecialColumn":
        if getattr(col, "name", None) == "RowIndex":
            return None
        raise NotImplementedError(f"special column {col!r} is not supported")
    return col


def reduction_from(agg):
    name = type(agg).__name__
    if name == "summary":
        return rd.summary(**{k: reduction_from(v) for k, v in zip(agg.keys, agg.values)})
    if name in ("by", "count_cat"):
        return rd.by(_categorizer(agg.categorizer), reduction_from(agg.reduction))
    if name == "where":
        return rd.where(reduction_from(agg.selector), _column(agg.column))
    if name in _COLUMN:
        return _COLUMN[name](_column(agg.column))
    if name in _N:
        return _N[name](_column(agg.column), n=agg.n)
    if name in _ROW_INDEX:
        return _ROW_INDEX[name]()
    if name in _ROW_INDEX_N:
        return _ROW_INDEX_N[name](n=agg.n)
    raise NotImplementedError(f"reduction {name} is not supported by the MI355X backend")


def canvas_from(canvas):
    def kind(axis):
        return "log" if type(axis).__name__ == "LogAxis" else "linear"
    return Canvas(canvas.plot_width, canvas.plot_height, canvas.x_range, canvas.y_range,
                  kind(canvas.x_axis), kind(canvas.y_axis))


def from_datashader(glyph, agg, canvas):
    """(glyph, reduction, canvas) of this package for datashader's objects."""
    return glyph_from(glyph), reduction_from(agg), canvas_from(canvas)

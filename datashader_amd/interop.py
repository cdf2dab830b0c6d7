"""Translate datashader's own glyph / reduction / canvas objects into this package's, so
that datashader can register this backend with ``bypixel.pipeline`` (see the
integration notes).  Matching is by class name and public attributes, so datashader
itself is not imported here."""
from __future__ import annotations

from . import glyphs as g
from . import reductions as rd
from .core import Canvas

_GLYPHS = {"Point": g.Point, "LineAxis0": g.LineAxis0, "LineAxis0Multi": g.LineAxis0Multi,
           "LinesAxis1": g.LinesAxis1, "LinesAxis1XConstant": g.LinesAxis1XConstant,
           "LinesAxis1YConstant": g.LinesAxis1YConstant}

_SIMPLE = {"count": rd.count, "any": rd.any, "sum": rd.sum, "mean": rd.mean, "min": rd.min,
           "max": rd.max, "var": rd.var, "std": rd.std}


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


def reduction_from(agg):
    name = type(agg).__name__
    if name == "summary":
        return rd.summary(**{k: reduction_from(v) for k, v in zip(agg.keys, agg.values)})
    if name in ("by", "count_cat"):
        return rd.by(_categorizer(agg.categorizer), reduction_from(agg.reduction))
    if name in _SIMPLE:
        col = agg.column
        if name in ("count", "any"):
            return _SIMPLE[name](col)
        return _SIMPLE[name](col)
    raise NotImplementedError(f"reduction {name} is not supported by the MI355X backend")


def canvas_from(canvas):
    def kind(axis):
        return "log" if type(axis).__name__ == "LogAxis" else "linear"
    return Canvas(canvas.plot_width, canvas.plot_height, canvas.x_range, canvas.y_range,
                  kind(canvas.x_axis), kind(canvas.y_axis))


def from_datashader(glyph, agg, canvas):
    """(glyph, reduction, canvas) of this package for datashader's objects."""
    return glyph_from(glyph), reduction_from(agg), canvas_from(canvas)

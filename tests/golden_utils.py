"""Helpers shared by the tests: load the reference's golden cases and evaluate them with
the CPU oracle (the oracle is the checker; nothing here is product code)."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pandas as pd

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
_CASES = None


def cases():
    global _CASES
    if _CASES is None:
        with open(os.path.join(HERE, "golden", "reference_cases.json")) as f:
            _CASES = json.load(f)
    return _CASES


def dec(a, dtype=float):
    """Inverse of the generator's encoder: None -> NaN."""
    def fix(v):
        if isinstance(v, list):
            return [fix(u) for u in v]
        return np.nan if v is None else v
    return np.array(fix(a), dtype=dtype)


def frame(name):
    """The named golden frame as a pandas DataFrame (categoricals as ``category``)."""
    spec = cases()["frames"][name]
    cols = {}
    for k, c in spec.items():
        if c["dtype"] == "category":
            cols[k] = pd.Categorical(c["data"])
        else:
            cols[k] = dec(c["data"], float).astype(c["dtype"])
    return pd.DataFrame(cols)


def line_frame(spec):
    """The frame of a golden line case; ``{"ragged": rows, "dtype": d}`` columns become
    object columns of per-row arrays (one ragged value per row)."""
    cols = {}
    for k, v in spec.items():
        if isinstance(v, dict) and "ragged" in v:
            cols[k] = pd.Series([np.asarray(r, dtype=v["dtype"]) for r in v["ragged"]],
                                dtype=object)
        else:
            cols[k] = dec(v, float)
    return pd.DataFrame(cols)


def scale_translate(rng, n):
    """``Axis.compute_scale_and_translate`` for a linear axis (core.py:62-81)."""
    start, end = rng
    s = n / (end - start)
    t = -start * s
    return s, t


def maybe_expand(bounds):
    """``Glyph.maybe_expand_bounds`` (glyph.py:42-49)."""
    lo, hi = bounds
    if not (np.isfinite(lo) and np.isfinite(hi)):
        return -1.0, 1.0
    if lo == hi:
        return lo - 1, lo + 1
    return lo, hi


def view_for(W, H, x_range, y_range, xf32=False, yf32=False):
    sx, tx = scale_translate(x_range, W)
    sy, ty = scale_translate(y_range, H)
    return oracle.View(sx, tx, sy, ty, x_range[0], x_range[1], y_range[0], y_range[1],
                       xf32=xf32, yf32=yf32)


# ---- evaluating a reduction spec with the oracle ------------------------------------

def _cat_codes(df, cat):
    """Category codes and categories for ``by``'s categorizer (reductions.py:143-259)."""
    if isinstance(cat, str):
        c = df[cat]
        return np.asarray(c.cat.codes, np.int64), list(c.cat.categories)
    if cat["kind"] == "category_modulo":
        v = (df[cat["column"]] - cat.get("offset", 0)) % cat["modulo"]
        return np.asarray(v, np.int64), list(range(cat["modulo"]))
    if cat["kind"] == "category_binning":
        nb = cat["nbins"]
        bin0 = cat["lower"]
        binsize = (cat["upper"] - cat["lower"]) / float(nb)
        values = df[cat["column"]].to_numpy()
        nanv = np.isnan(values)
        idx = (values - bin0) / binsize
        idx[nanv] = 0
        idx = idx.astype(int)
        idx[idx < 0] = 0
        idx[idx >= nb] = nb - 1
        idx[nanv] = nb
        return idx.astype(np.int64), list(range(nb + 1))
    raise ValueError(cat)


def _leaf(spec, run):
    """Finalised array of one non-summary reduction, given ``run(bases, val, cat, ncat)``."""
    kind, col = spec["kind"], spec.get("column")
    if kind == "count_cat":
        spec = {"kind": "by", "cat": col, "reduction": {"kind": "count", "column": None}}
        kind = "by"
    if kind == "by":
        codes, cats = run.cat_codes(spec["cat"])
        sub = spec["reduction"]
        return _leaf_inner(sub, run, codes, len(cats))
    return _leaf_inner(spec, run, None, 1)


_ROW_MODES = {"first": "first", "last": "last", "_min_row_index": "first",
              "_max_row_index": "last", "max": "max", "min": "min"}


_N_MODES = {"max_n": "max", "min_n": "min", "first_n": "first", "last_n": "last",
            "_min_n_row_index": "first", "_max_n_row_index": "last"}


def _leaf_inner(spec, run, codes, ncat):
    kind, col = spec["kind"], spec.get("column")
    if kind in _ROW_MODES or kind in _N_MODES or kind == "where":
        if kind == "where":
            sel, lookup = spec["selector"], spec["lookup"]
        else:
            sel, lookup = spec, spec.get("column")
        if sel["kind"] in _N_MODES:
            return run.n_select(_N_MODES[sel["kind"]], sel.get("column"), lookup, sel["n"],
                                codes, ncat, codes is None)
        return run.row_index(_ROW_MODES[sel["kind"]], sel.get("column"), lookup, codes, ncat,
                             codes is None)
    val = None if col is None else run.column(col)
    squeeze = codes is None
    if kind == "count":
        r = run(("count",) if col is None else ("countv",), val, codes, ncat, squeeze)
        return r["count" if col is None else "countv"]
    if kind == "any":
        r = run(("any",) if col is None else ("anyv",), val, codes, ncat, squeeze)
        return r["any" if col is None else "anyv"]
    if kind in ("sum", "min", "max"):
        return run((kind,), val, codes, ncat, squeeze)[kind]
    if kind == "mean":
        r = run(("sumz", "countv"), val, codes, ncat, squeeze)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(r["countv"] > 0, r["sumz"] / r["countv"], np.nan)
    if kind in ("var", "std"):
        r = run(("countv", "m2"), val, codes, ncat, squeeze)
        with np.errstate(divide="ignore", invalid="ignore"):
            v = np.where(r["countv"] > 0, r["m2"] / r["countv"], np.nan)
        return np.sqrt(v) if kind == "std" else v
    raise ValueError(kind)


def evaluate(spec, run):
    if spec["kind"] == "summary":
        return {k: _leaf(v, run) for k, v in spec["items"].items()}
    return _leaf(spec, run)


class PointRun:
    def __init__(self, df, x, y, view, W, H):
        self.df, self.x, self.y, self.view, self.W, self.H = df, x, y, view, W, H

    def column(self, name):
        return self.df[name].to_numpy(dtype=np.float64)

    def cat_codes(self, cat):
        return _cat_codes(self.df, cat)

    def row_index(self, mode, selcol, lookup, codes, ncat, squeeze):
        """first / last / where / row-index selectors with the sequential oracle."""
        from oracle import closed_forms as cf
        from oracle import rowindex as ri
        bins = cf.point_bins(self.df[self.x].to_numpy(np.float64),
                             self.df[self.y].to_numpy(np.float64), self.view, self.W, self.H)
        cell = ri.cells(bins, codes, ncat)
        val = None if selcol is None else self.column(selcol)
        idx = ri.select_rows(cell, self.W * self.H * ncat, mode, val)
        out = idx if lookup is None else ri.take(idx, self.column(lookup))
        return out.reshape((self.H, self.W) if squeeze else (self.H, self.W, ncat))

    def n_select(self, mode, selcol, lookup, n, codes, ncat, squeeze):
        """The n-value reductions with the sequential oracle."""
        from oracle import closed_forms as cf
        from oracle import rowindex as ri
        bins = cf.point_bins(self.df[self.x].to_numpy(np.float64),
                             self.df[self.y].to_numpy(np.float64), self.view, self.W, self.H)
        cell = ri.cells(bins, codes, ncat)
        val = None if selcol is None else self.column(selcol)
        idx = ri.select_n(cell, self.W * self.H * ncat, n, mode, val)
        out = idx if lookup is None else ri.take(idx, self.column(lookup))
        return out.reshape((self.H, self.W, n) if squeeze else (self.H, self.W, ncat, n))

    def __call__(self, bases, val, codes, ncat, squeeze):
        return oracle.points(self.df[self.x].to_numpy(np.float64),
                             self.df[self.y].to_numpy(np.float64), self.view, self.W, self.H,
                             bases, val=val, cat=codes, ncat=ncat, squeeze=squeeze)


def point_case_oracle(case):
    """Evaluate one golden points case with the oracle (autorange as the reference)."""
    df = frame(case["frame"])
    cv = case["canvas"]
    W, H = cv["W"], cv["H"]
    xr_ = cv["x_range"] or maybe_expand(oracle.bounds(df[case["x"]].to_numpy(np.float64)))
    yr_ = cv["y_range"] or maybe_expand(oracle.bounds(df[case["y"]].to_numpy(np.float64)))
    view = view_for(W, H, xr_, yr_)
    res = evaluate(case["agg"], PointRun(df, case["x"], case["y"], view, W, H))
    if isinstance(res, dict):
        res = res[case["pick"]]
    return res, (xr_, yr_)


def expected_array(case):
    exp = case["expected"]
    return dec(exp, float)


def assert_same(got, exp):
    """Exact equality with NaN == NaN (xarray.testing.assert_equal semantics)."""
    got = np.asarray(got, dtype=float)
    exp = np.asarray(exp, dtype=float)
    assert got.shape == exp.shape, (got.shape, exp.shape)
    both_nan = np.isnan(got) & np.isnan(exp)
    bad = ~(both_nan | (got == exp))
    assert not bad.any(), f"mismatch at {np.argwhere(bad)[:5].tolist()}: got {got[bad][:5]}, exp {exp[bad][:5]}"


# ---- lines ---------------------------------------------------------------------------

def line_polylines(case_kwargs, df):
    """Expand a Canvas.line call of the golden cases into (xs_2d, ys_2d, layout) where
    row k of xs_2d/ys_2d is one line.  ``layout`` is 'axis0' (value index = vertex) or
    'axis1' (value index = line)."""
    x, y, axis = case_kwargs["x"], case_kwargs["y"], case_kwargs.get("axis", 0)
    if isinstance(x, list) and isinstance(y, str):
        y = [y] * len(x)
    if isinstance(y, list) and isinstance(x, str):
        x = [x] * len(y)
    if axis == 1 and isinstance(x, str) and isinstance(y, str):
        # ragged rows: row i joins its first min(len x_i, len y_i) vertices
        xs, ys = [], []
        for xr_, yr_ in zip(df[x], df[y]):
            m = min(len(xr_), len(yr_))
            xs.append(np.asarray(xr_[:m], float))
            ys.append(np.asarray(yr_[:m], float))
        return xs, ys, "ragged"
    if axis == 0:
        if isinstance(x, str):
            return [df[x].to_numpy(float)], [df[y].to_numpy(float)], "axis0"
        return ([df[c].to_numpy(float) for c in x], [df[c].to_numpy(float) for c in y], "axis0")
    if isinstance(x, dict):
        xs = np.asarray(x["ndarray"], float)
        ys = df[list(y)].to_numpy(float)
        return [xs] * ys.shape[0], list(ys), "axis1"
    if isinstance(y, dict):
        ys = np.asarray(y["ndarray"], float)
        xs = df[list(x)].to_numpy(float)
        return list(xs), [ys] * xs.shape[0], "axis1"
    return list(df[list(x)].to_numpy(float)), list(df[list(y)].to_numpy(float)), "axis1"


def line_case_oracle(case):
    df = line_frame(case["frame"])
    kw = case["kwargs"]
    xs, ys, layout = line_polylines(kw, df)
    cv = case["canvas"]
    W, H = cv["W"], cv["H"]

    def bnd(arrs):
        mins, maxs = zip(*[oracle.bounds(a) for a in arrs])
        return maybe_expand((min(mins), max(maxs)))

    if layout == "ragged":
        xs_all = [np.concatenate([np.asarray(r, float) for r in df[kw["x"]]])]
        ys_all = [np.concatenate([np.asarray(r, float) for r in df[kw["y"]]])]
    else:
        xs_all, ys_all = xs, ys
    xr_ = cv["x_range"] or (maybe_expand((np.nanmin(kw["x"]["ndarray"]),
                                         np.nanmax(kw["x"]["ndarray"])))
                            if isinstance(kw["x"], dict) else bnd(xs_all))
    yr_ = cv["y_range"] or (maybe_expand((np.nanmin(kw["y"]["ndarray"]),
                                         np.nanmax(kw["y"]["ndarray"])))
                            if isinstance(kw["y"], dict) else bnd(ys_all))
    view = view_for(W, H, xr_, yr_)
    agg = kw["agg"]
    col = agg.get("column")
    kind = agg["kind"]
    base = {"count": "count" if col is None else "countv", "sum": "sum", "max": "max",
            "min": "min", "any": "any" if col is None else "anyv"}[kind]
    val = None if col is None else df[col].to_numpy(float)
    if layout == "axis0":
        total = None
        for xl, yl in zip(xs, ys):
            r = oracle.line_axis0(xl, yl, view, W, H, (base,), plot_start=True, val=val)[base]
            total = r.astype(float) if total is None else _merge(kind, total, r)
        return total, (xr_, yr_)
    if layout == "ragged":
        total = None
        for i, (xl, yl) in enumerate(zip(xs, ys)):
            v1 = None if val is None else val[i:i + 1]
            r = oracle.lines_axis1(xl[None, :], yl[None, :], view, W, H, (base,), val=v1)[base]
            total = r.astype(float) if total is None else _merge(kind, total, r)
        return total, (xr_, yr_)
    r = oracle.lines_axis1(np.array(xs), np.array(ys), view, W, H, (base,), val=val)[base]
    return r, (xr_, yr_)


def _merge(kind, a, b):
    b = b.astype(float)
    if kind in ("count", "any"):
        return a + b
    if kind == "sum":
        return np.where(np.isnan(a), b, np.where(np.isnan(b), a, a + b))
    if kind == "max":
        return np.fmax(a, b)
    return np.fmin(a, b)


# ---- tf.shade cases (golden/shade_cases.json) ---------------------------------------------
_SHADE = None


def shade_cases():
    global _SHADE
    if _SHADE is None:
        with open(os.path.join(HERE, "golden", "shade_cases.json")) as f:
            _SHADE = json.load(f)["cases"]
    return _SHADE


def _colour(c):
    return tuple(c["tuple"]) if isinstance(c, dict) else c


def shade_case(case):
    """(plane, shade kwargs, expected uint32 image) of one golden shade case; colours are
    decoded to names / RGB tuples and dict colour keys to dicts, as the test passed them."""
    spec = case["data"]
    vals = [np.nan if v is None else v for v in spec["values"]]
    data = np.array(vals, dtype=spec["dtype"]).reshape(spec["shape"])
    kw = dict(case["kwargs"])
    if "cmap" in kw:
        cm = kw["cmap"]
        kw["cmap"] = [_colour(c) for c in cm] if isinstance(cm, list) else _colour(cm)
    if "color_key" in kw:
        ck = kw["color_key"]
        kw["color_key"] = ({k: _colour(v) for k, v in ck["dict"]} if isinstance(ck, dict)
                           else [_colour(c) for c in ck])
    expected = np.array(case["expected"], dtype=np.uint32).reshape(data.shape[:2])
    return data, kw, expected


def shade_oracle(data, kw):
    """The oracle's image for shade(plane, **kw) (names resolved with the reference's
    colour table as datashader_amd.colors restates it)."""
    from datashader_amd.colors import Sets1to3, rgb
    from oracle import shade as osh
    kw = dict(kw)
    how = kw.pop("how", "eq_hist")
    alpha = kw.pop("alpha", 255)
    min_alpha = kw.pop("min_alpha", 40)
    span = kw.pop("span", None)
    rescale = kw.pop("rescale_discrete_levels", False) and how == "eq_hist"
    baseline = kw.pop("color_baseline", None)
    if data.ndim == 3:
        ck = kw.pop("color_key", Sets1to3)
        ncat = data.shape[2]
        cols = list(ck.values())[:ncat] if isinstance(ck, dict) else list(ck)[:ncat]
        return osh.colorize(data, [rgb(c) for c in cols], how=how, alpha=alpha, span=span,
                            min_alpha=min_alpha, color_baseline=baseline,
                            rescale_discrete_levels=rescale)
    ck = kw.pop("color_key", None)
    if isinstance(ck, dict):
        return osh.discrete(data, list(ck.keys()), [rgb(c) for c in ck.values()], alpha=alpha,
                            color_baseline=baseline)
    cmap = kw.pop("cmap", ["lightblue", "darkblue"])
    if isinstance(cmap, (str, tuple)):
        return osh.interpolate(data, rgb(cmap), how=how, alpha=alpha, span=span,
                               min_alpha=min_alpha, rescale_discrete_levels=rescale,
                               single=True)
    return osh.interpolate(data, [rgb(c) for c in cmap], how=how, alpha=alpha, span=span,
                           min_alpha=min_alpha, rescale_discrete_levels=rescale)


def channels_close(a, b, tol):
    """Per-channel |a - b| <= tol for two uint32 RGBA images."""
    x = np.asarray(a, dtype=np.uint32).view(np.uint8).astype(int)
    y = np.asarray(b, dtype=np.uint32).view(np.uint8).astype(int)
    return x.shape == y.shape and bool(np.all(np.abs(x - y) <= tol))


# ---- float32 segments ending exactly on a view bound --------------------------------

def f32_bound_cases():
    """Two-segment float32 polylines whose last vertex lies exactly on one of the four
    view bounds, with a float32 segment delta that rounds above the float64 distance q
    from the segment's start to that bound.

    Under the reference's typing (``dx1 = x1 - x0`` in the column's float32,
    ``q = x0 - xmin`` / ``xmax - x0`` in float64, ``line.py:752-759``), ``_clipt``
    takes its ``q < p`` branch (``line.py:792-798``), sets ``t1 = q / p < 1`` and the
    second segment is clipped; both its ends snap to one pixel, so it draws nothing
    (``line.py:1003-1008``: a clipped segment without ``segment_start`` appends only
    while stepping).  With float64 deltas the same segment is unclipped and zero-length
    and appends its pixel once.  The first segment lies inside the view and, by
    Bresenham, touches max(|dx|, |dy|) + 1 = 4 pixels either way.  So every case draws
    4 pixels in float32 mode and 5 in float64 mode.

    The base case puts the last vertex on ymax on a 34 x 7 canvas; the others mirror it
    (y -> -y) and swap the axes.  The premise p > q is checked with numpy's float32 and
    float64 arithmetic for each case."""
    f = np.float32
    bx = np.array([0.30, 0.3397695, 0.33822325], f)
    by = np.array([-0.5, -0.027073061, 0.043172628], f)
    top = float(by[2])
    base = [
        ("ymax", bx, by, 34, 7, (0.0, 1.0), (-1.0, top), "y", +1),
        ("ymin", bx, -by, 34, 7, (0.0, 1.0), (-top, 1.0), "y", -1),
        ("xmax", by, bx, 7, 34, (-1.0, top), (0.0, 1.0), "x", +1),
        ("xmin", -by, bx, 7, 34, (-top, 1.0), (0.0, 1.0), "x", -1),
    ]
    out = []
    for name, xs, ys, W, H, xr, yr, axis, side in base:
        c = ys if axis == "y" else xs
        lo, hi = (yr if axis == "y" else xr)
        p = float(f(c[2]) - f(c[1])) * side          # float32 delta toward the bound
        q = (hi - float(c[1])) if side > 0 else (float(c[1]) - lo)  # float64 distance
        assert c[2] == (hi if side > 0 else lo) and p > q > 0, name
        out.append(dict(name=name, xs=xs, ys=ys, W=W, H=H, x_range=xr, y_range=yr,
                        total_f32=4, total_f64=5))
    return out

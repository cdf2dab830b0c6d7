"""Write ``reference_cases.json``: inputs and expected outputs of the reference's own tests
for the points / line path, as data.

Every case below restates one assertion of datashader 0.19.1's test-suite: the inputs the
test builds and the expected array it asserts, computed the way that test computes it
(e.g. ``values(df.i32).reshape((2, 2, 5)).sum(axis=2, dtype='f8').T``).  The citation of
each case names the test (file:line, relative to the reference checkout).

The touched-pixel masks of the reference's line images
(``datashader/tests/data/test_00{1..7}_noaa_{normal,clipped}.png``, produced by
``tests/test_antialias.py:302-370``) are read with Pillow when the reference checkout is
present; a pixel is "touched" when it is not the white background, after undoing the
image's vertical flip.  The generated JSON is committed; this script only needs to run
again when a case is added.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_cases.json")
TP = "datashader/tests/test_pandas.py"
TG = "datashader/tests/test_glyphs.py"
TA = "datashader/tests/test_antialias.py"


def enc(a):
    """JSON-encode an array: NaN -> None, bool/int/float kept."""
    a = np.asarray(a)
    if a.dtype.kind == "f":
        return [enc(r) for r in a] if a.ndim > 1 else [None if math.isnan(v) else float(v) for v in a]
    if a.dtype.kind == "b":
        return a.astype(int).tolist()
    return a.tolist()


def pandas_frame():
    """The 20-row frame of ``_pandas()`` (test_pandas.py:27-59), column by column."""
    nan = float("nan")
    f32 = np.arange(20, dtype="f4").astype(float)
    f32[2] = nan
    f64 = np.arange(20, dtype="f8")
    f64[2] = nan
    reverse = np.arange(20, 0, -1).astype("f8")
    reverse[6] = nan
    plusminus = np.arange(20, dtype="f8") * ([1, -1] * 10)
    plusminus[2] = nan
    return {
        "x": {"dtype": "f8", "data": [0.0] * 10 + [1.0] * 10},
        "y": {"dtype": "f8", "data": [0.0] * 5 + [1.0] * 5 + [0.0] * 5 + [1.0] * 5},
        "i32": {"dtype": "i4", "data": list(range(20))},
        "i64": {"dtype": "i8", "data": list(range(20))},
        "f32": {"dtype": "f4", "data": enc(f32)},
        "f64": {"dtype": "f8", "data": enc(f64)},
        "reverse": {"dtype": "f8", "data": enc(reverse)},
        "plusminus": {"dtype": "f8", "data": enc(plusminus)},
        "empty_bin": {"dtype": "f8", "data": [0.0] * 15 + [None] * 5},
        "cat": {"dtype": "category", "data": ["a"] * 5 + ["b"] * 5 + ["c"] * 5 + ["d"] * 5},
        "cat2": {"dtype": "category", "data": ["a", "b", "c", "d"] * 5},
        "onecat": {"dtype": "category", "data": ["one"] * 20},
        "cat_int": {"dtype": "i8", "data": [10] * 5 + [11] * 5 + [12] * 5 + [13] * 5},
    }


def col(frame, name):
    c = frame[name]
    return np.array([np.nan if v is None else v for v in c["data"]], dtype=float)


def point_cases(frame):
    cases = []
    C = {"W": 2, "H": 2, "x_range": [0, 1], "y_range": [0, 1]}

    def add(name, cite, agg, expected, canvas=C, df="_pandas", x="x", y="y", **kw):
        cases.append(dict(name=name, cite=cite, canvas=canvas, frame=df, x=x, y=y, agg=agg,
                          expected=enc(expected), **kw))

    i32 = col(frame, "i32")
    f64 = col(frame, "f64")
    cite = TP + ":194-203"
    for c in ("i32", "i64", None):
        add(f"test_count/count({c})", cite, {"kind": "count", "column": c}, [[5, 5], [5, 5]])
    for c in ("f32", "f64"):
        add(f"test_count/count({c})", cite, {"kind": "count", "column": c}, [[4, 5], [5, 5]])
    cite = TP + ":206-214"
    for c in ("i64", "f64", None):
        add(f"test_any/any({c})", cite, {"kind": "any", "column": c}, [[True, True], [True, True]])
    add("test_any/any(empty_bin)", cite, {"kind": "any", "column": "empty_bin"},
        [[True, True], [True, False]])
    cite = TP + ":217-225"
    exp = i32.reshape((2, 2, 5)).sum(axis=2, dtype="f8").T
    for c in ("i32", "i64"):
        add(f"test_sum/sum({c})", cite, {"kind": "sum", "column": c}, exp)
    exp = np.nansum(f64.reshape((2, 2, 5)), axis=2).T
    for c in ("f32", "f64"):
        add(f"test_sum/sum({c})", cite, {"kind": "sum", "column": c}, exp)
    cite = TP + ":228-235"
    exp = np.arange(20).reshape((2, 2, 5)).min(axis=2).astype("f8").T
    for c in ("i32", "i64", "f32", "f64"):
        add(f"test_min/min({c})", cite, {"kind": "min", "column": c}, exp)
    cite = TP + ":237-243"
    exp = np.arange(20).reshape((2, 2, 5)).max(axis=2).astype("f8").T
    for c in ("i32", "i64", "f32", "f64"):
        add(f"test_max/max({c})", cite, {"kind": "max", "column": c}, exp)
    cite = TP + ":698-706"
    exp = i32.reshape((2, 2, 5)).mean(axis=2, dtype="f8").T
    for c in ("i32", "i64"):
        add(f"test_mean/mean({c})", cite, {"kind": "mean", "column": c}, exp)
    exp = np.nanmean(f64.reshape((2, 2, 5)), axis=2).T
    for c in ("f32", "f64"):
        add(f"test_mean/mean({c})", cite, {"kind": "mean", "column": c}, exp)
    for red, fn, fnn, cite in (("var", np.var, np.nanvar, TP + ":709-717"),
                               ("std", np.std, np.nanstd, TP + ":720-728")):
        exp = fn(i32.reshape((2, 2, 5)), axis=2, dtype="f8").T
        for c in ("i32", "i64"):
            add(f"test_{red}/{red}({c})", cite, {"kind": red, "column": c}, exp)
        exp = fnn(f64.reshape((2, 2, 5)), axis=2).T
        for c in ("f32", "f64"):
            add(f"test_{red}/{red}({c})", cite, {"kind": red, "column": c}, exp)

    sol = np.array([[[5, 0, 0, 0], [0, 0, 5, 0]], [[0, 5, 0, 0], [0, 0, 0, 5]]])
    add("test_count_cat/count_cat(cat)", TP + ":731-740", {"kind": "count_cat", "column": "cat"},
        sol, cat_coords=["a", "b", "c", "d"], cat_dim="cat")
    add("test_categorical_count/by(cat,count(i32))", TP + ":743-754",
        {"kind": "by", "cat": "cat", "reduction": {"kind": "count", "column": "i32"}},
        sol, cat_coords=["a", "b", "c", "d"], cat_dim="cat")
    add("test_categorical_count/summary(name=by(cat,count(i32)))", TP + ":756-758",
        {"kind": "summary", "items": {"name": {
            "kind": "by", "cat": "cat", "reduction": {"kind": "count", "column": "i32"}}}},
        sol, pick="name", cat_coords=["a", "b", "c", "d"], cat_dim="cat")
    add("test_categorical_count/by(category_modulo(cat_int,4,10),count())", TP + ":760-764",
        {"kind": "by", "cat": {"kind": "category_modulo", "column": "cat_int", "modulo": 4,
                               "offset": 10}, "reduction": {"kind": "count", "column": None}},
        sol, cat_coords=[0, 1, 2, 3], cat_dim="cat_int")
    add("test_one_category/by(onecat,count(i32))", TP + ":763-770",
        {"kind": "by", "cat": "onecat", "reduction": {"kind": "count", "column": "i32"}},
        np.array([[[5], [5]], [[5], [5]]]), cat_coords=["one"], cat_dim="onecat")
    solb = np.append(sol, [[[0], [0]], [[0], [0]]], axis=2)
    for c in ("i32", "i64"):
        add(f"test_categorical_count_binning/{c}", TP + ":773-788",
            {"kind": "by", "cat": {"kind": "category_binning", "column": c, "lower": 0,
                                   "upper": 20, "nbins": 4},
             "reduction": {"kind": "count", "column": None}},
            solb, cat_coords=[0, 1, 2, 3, 4], cat_dim=c)
    solf = solb.copy()
    solf[0, 0, 0] = 4
    solf[0, 0, 4] = 1
    for c in ("f32", "f64"):
        add(f"test_categorical_count_binning/{c}", TP + ":790-797",
            {"kind": "by", "cat": {"kind": "category_binning", "column": c, "lower": 0,
                                   "upper": 20, "nbins": 4},
             "reduction": {"kind": "count", "column": None}},
            solf, cat_coords=[0, 1, 2, 3, 4], cat_dim=c)
    cite = TP + ":1067-1077"
    summ = {"kind": "summary", "items": {
        "f64_mean": {"kind": "mean", "column": "f64"},
        "i32_sum": {"kind": "sum", "column": "i32"},
        "i32_count": {"kind": "count", "column": "i32"}}}
    add("test_multiple_aggregates/f64_mean", cite, summ,
        np.nanmean(f64.reshape((2, 2, 5)), axis=2).T, pick="f64_mean")
    add("test_multiple_aggregates/i32_sum", cite, summ,
        i32.reshape((2, 2, 5)).sum(axis=2, dtype="f8").T, pick="i32_sum")
    add("test_multiple_aggregates/i32_count", cite, summ, [[5, 5], [5, 5]], pick="i32_count")

    # row-index reductions: first / last / where / _min_row_index / _max_row_index
    nan = float("nan")
    lo_rows, hi_rows = [[0, 10], [5, 15]], [[4, 14], [9, 19]]
    for c in ("i32", "i64", "f32", "f64"):
        add(f"test_first/first({c})", TP + ":991-996", {"kind": "first", "column": c}, lo_rows)
        add(f"test_last/last({c})", TP + ":999-1004", {"kind": "last", "column": c}, hi_rows)
        for sel, rev, rows, cite in (("first", [[20, 10], [15, 5]], lo_rows, ":448-462"),
                                     ("last", [[16, 6], [11, 1]], hi_rows, ":465-479"),
                                     ("max", [[16, 6], [11, 1]], hi_rows, ":482-494"),
                                     ("min", [[20, 10], [15, 5]], lo_rows, ":497-509")):
            s_ = {"kind": sel, "column": c}
            add(f"test_where_{sel}/where({sel}({c}),reverse)", TP + cite,
                {"kind": "where", "selector": s_, "lookup": "reverse"}, rev)
            add(f"test_where_{sel}/where({sel}({c}))", TP + cite,
                {"kind": "where", "selector": s_, "lookup": None}, rows, dtype="int64")
    add("test_min_row_index", TP + ":1029-1033", {"kind": "_min_row_index"}, lo_rows,
        dtype="int64")
    add("test_max_row_index", TP + ":1036-1040", {"kind": "_max_row_index"}, hi_rows,
        dtype="int64")
    add("test_where_min_row_index", TP + ":404-406",
        {"kind": "where", "selector": {"kind": "_min_row_index"}, "lookup": "plusminus"},
        [[0, 10], [-5, -15]])
    add("test_where_max_row_index", TP + ":409-411",
        {"kind": "where", "selector": {"kind": "_max_row_index"}, "lookup": "plusminus"},
        [[4, 14], [-9, -19]])
    add("test_categorical_min_row_index", TP + ":316-319",
        {"kind": "by", "cat": "cat2", "reduction": {"kind": "_min_row_index"}},
        [[[0, 1, 2, 3], [12, 13, 10, 11]], [[8, 5, 6, 7], [16, 17, 18, 15]]],
        cat_coords=["a", "b", "c", "d"], cat_dim="cat2", dtype="int64")
    add("test_categorical_max_row_index", TP + ":322-325",
        {"kind": "by", "cat": "cat2", "reduction": {"kind": "_max_row_index"}},
        [[[4, 1, 2, 3], [12, 13, 14, 11]], [[8, 9, 6, 7], [16, 17, 18, 19]]],
        cat_coords=["a", "b", "c", "d"], cat_dim="cat2", dtype="int64")
    add("test_categorical_first", TP + ":356-363",
        {"kind": "by", "cat": "cat2", "reduction": {"kind": "first", "column": "plusminus"}},
        [[[0, -1, nan, -3], [12, -13, 10, -11]], [[8, -5, 6, -7], [16, -17, 18, -15]]],
        cat_coords=["a", "b", "c", "d"], cat_dim="cat2")
    add("test_categorical_last", TP + ":366-373",
        {"kind": "by", "cat": "cat2", "reduction": {"kind": "last", "column": "plusminus"}},
        [[[4, -1, nan, -3], [12, -13, 14, -11]], [[8, -9, 6, -7], [16, -17, 18, -19]]],
        cat_coords=["a", "b", "c", "d"], cat_dim="cat2")

    # n-value reductions (max_n / min_n / first_n / last_n / row-index forms / where)
    pm_min = np.array([[[-3, -1, 0, 4, nan, nan], [-13, -11, 10, 12, 14, nan]],
                       [[-9, -7, -5, 6, 8, nan], [-19, -17, -15, 16, 18, nan]]])
    pm_max = np.array([[[4, 0, -1, -3, nan, nan], [14, 12, 10, -11, -13, nan]],
                       [[8, 6, -5, -7, -9, nan], [18, 16, -15, -17, -19, nan]]])
    pm_first = np.array([[[0, -1, -3, 4, nan, nan], [10, -11, 12, -13, 14, nan]],
                         [[-5, 6, -7, 8, -9, nan], [-15, 16, -17, 18, -19, nan]]])
    pm_last = np.array([[[4, -3, -1, 0, nan, nan], [14, -13, 12, -11, 10, nan]],
                        [[-9, 8, -7, 6, -5, nan], [-19, 18, -17, 16, -15, nan]]])
    ri_min = np.array([[[0, 1, 2, 3, 4, -1], [10, 11, 12, 13, 14, -1]],
                       [[5, 6, 7, 8, 9, -1], [15, 16, 17, 18, 19, -1]]])
    ri_max = np.array([[[4, 3, 2, 1, 0, -1], [14, 13, 12, 11, 10, -1]],
                       [[9, 8, 7, 6, 5, -1], [19, 18, 17, 16, 15, -1]]])
    w_min_ri = np.array([[[0, -1, nan, -3, 4, nan], [10, -11, 12, -13, 14, nan]],
                         [[-5, 6, -7, 8, -9, nan], [-15, 16, -17, 18, -19, nan]]])
    w_max_ri = np.array([[[4, -3, nan, -1, 0, nan], [14, -13, 12, -11, 10, nan]],
                         [[-9, 8, -7, 6, -5, nan], [-19, 18, -17, 16, -15, nan]]])
    sel_rows = {  # test_where_{first,last,max,min}_n (512-621): selected row indices
        "first_n": np.array([[[0, 1, 3, 4, -1, -1], [10, 11, 12, 13, 14, -1]],
                             [[5, 6, 7, 8, 9, -1], [15, 16, 17, 18, 19, -1]]]),
        "last_n": np.array([[[4, 3, 1, 0, -1, -1], [14, 13, 12, 11, 10, -1]],
                            [[9, 8, 7, 6, 5, -1], [19, 18, 17, 16, 15, -1]]]),
        "max_n": np.array([[[4, 0, 1, 3, -1, -1], [14, 12, 10, 11, 13, -1]],
                           [[8, 6, 5, 7, 9, -1], [18, 16, 15, 17, 19, -1]]]),
        "min_n": np.array([[[3, 1, 0, 4, -1, -1], [13, 11, 10, 12, 14, -1]],
                           [[9, 7, 5, 6, 8, -1], [19, 17, 15, 16, 18, -1]]]),
    }
    cites = {"first_n": ":512-536", "last_n": ":539-563", "max_n": ":566-589",
             "min_n": ":592-615"}
    for n in (1, 3, 6):
        for kind, sol, cite in (("min_n", pm_min, ":246-254"), ("max_n", pm_max, ":257-265"),
                                ("first_n", pm_first, ":1007-1015"),
                                ("last_n", pm_last, ":1018-1026")):
            add(f"test_{kind}/n={n}", TP + cite, {"kind": kind, "column": "plusminus", "n": n},
                sol[:, :, :n], n_dim=True)
        add(f"test_min_n_row_index/n={n}", TP + ":1043-1052",
            {"kind": "_min_n_row_index", "n": n}, ri_min[:, :, :n], dtype="int64", n_dim=True)
        add(f"test_max_n_row_index/n={n}", TP + ":1055-1064",
            {"kind": "_max_n_row_index", "n": n}, ri_max[:, :, :n], dtype="int64", n_dim=True)
        add(f"test_where_min_n_row_index/n={n}", TP + ":414-428",
            {"kind": "where", "selector": {"kind": "_min_n_row_index", "n": n},
             "lookup": "plusminus"}, w_min_ri[:, :, :n], n_dim=True)
        add(f"test_where_max_n_row_index/n={n}", TP + ":431-445",
            {"kind": "where", "selector": {"kind": "_max_n_row_index", "n": n},
             "lookup": "plusminus"}, w_max_ri[:, :, :n], n_dim=True)
        for kind, rows in sel_rows.items():
            sel = {"kind": kind, "column": "plusminus", "n": n}
            add(f"test_where_{kind}/rows/n={n}", TP + cites[kind],
                {"kind": "where", "selector": sel, "lookup": None}, rows[:, :, :n],
                dtype="int64", n_dim=True)
            rev = np.where((rows < 0) | (rows == 6), nan, 20 - rows)
            add(f"test_where_{kind}/reverse/n={n}", TP + cites[kind],
                {"kind": "where", "selector": sel, "lookup": "reverse"}, rev[:, :, :n],
                n_dim=True)
    cat_sols = {
        "min_n": ("f32", [[[[0, 4], [1, nan], [nan, nan], [3, nan]],
                           [[12, nan], [13, nan], [10, 14], [11, nan]]],
                          [[[8, nan], [5, 9], [6, nan], [7, nan]],
                           [[16, nan], [17, nan], [18, nan], [15, 19]]]], ":288-299"),
        "max_n": ("f32", [[[[4, 0], [1, nan], [nan, nan], [3, nan]],
                           [[12, nan], [13, nan], [14, 10], [11, nan]]],
                          [[[8, nan], [9, 5], [6, nan], [7, nan]],
                           [[16, nan], [17, nan], [18, nan], [19, 15]]]], ":302-313"),
        "first_n": ("plusminus", [[[[0, 4], [-1, nan], [nan, nan], [-3, nan]],
                                   [[12, nan], [-13, nan], [10, 14], [-11, nan]]],
                                  [[[8, nan], [-5, -9], [6, nan], [-7, nan]],
                                   [[16, nan], [-17, nan], [18, nan], [-15, -19]]]], ":376-387"),
        "last_n": ("plusminus", [[[[4, 0], [-1, nan], [nan, nan], [-3, nan]],
                                  [[12, nan], [-13, nan], [14, 10], [-11, nan]]],
                                 [[[8, nan], [-9, -5], [6, nan], [-7, nan]],
                                  [[16, nan], [-17, nan], [18, nan], [-19, -15]]]], ":390-401"),
    }
    cat_ri = {
        "_min_n_row_index": ([[[[0, 4], [1, -1], [2, -1], [3, -1]],
                               [[12, -1], [13, -1], [10, 14], [11, -1]]],
                              [[[8, -1], [5, 9], [6, -1], [7, -1]],
                               [[16, -1], [17, -1], [18, -1], [15, 19]]]], ":328-339"),
        "_max_n_row_index": ([[[[4, 0], [1, -1], [2, -1], [3, -1]],
                               [[12, -1], [13, -1], [14, 10], [11, -1]]],
                              [[[8, -1], [9, 5], [6, -1], [7, -1]],
                               [[16, -1], [17, -1], [18, -1], [19, 15]]]], ":342-353"),
    }
    for n in (1, 2):
        for kind, (column, sol, cite) in cat_sols.items():
            add(f"test_categorical_{kind}/n={n}", TP + cite,
                {"kind": "by", "cat": "cat2",
                 "reduction": {"kind": kind, "column": column, "n": n}},
                np.array(sol)[..., :n], cat_coords=["a", "b", "c", "d"], cat_dim="cat2",
                n_dim=True)
        for kind, (sol, cite) in cat_ri.items():
            add(f"test_categorical{kind}/n={n}", TP + cite,
                {"kind": "by", "cat": "cat2", "reduction": {"kind": kind, "n": n}},
                np.array(sol)[..., :n], cat_coords=["a", "b", "c", "d"], cat_dim="cat2",
                dtype="int64", n_dim=True)
    summ_n = {"kind": "summary", "items": {
        "count": {"kind": "count", "column": None},
        "min_n": {"kind": "where", "selector": {"kind": "min_n", "column": "plusminus", "n": 5},
                  "lookup": None},
        "max_n": {"kind": "where", "selector": {"kind": "max_n", "column": "plusminus", "n": 5},
                  "lookup": "reverse"}}}
    add("test_summary_where_n/min_n", TP + ":648-676", summ_n, sel_rows["min_n"][:, :, :5],
        pick="min_n", dtype="int64", n_dim=True)
    mx = sel_rows["max_n"][:, :, :5]
    add("test_summary_where_n/max_n", TP + ":648-676", summ_n,
        np.where((mx < 0) | (mx == 6), nan, 20 - mx), pick="max_n", n_dim=True)
    add("test_summary_where_n/count", TP + ":648-676", summ_n, [[5, 5], [5, 5]], pick="count")

    # test_auto_range_points, test_pandas.py:1080-1129 (integer x, y; autorange)
    frames = {}
    for n, mult, extra, cite in ((10, 1, 0, ":1080-1092"), (10, 1, 1, ":1094-1101"),
                                 (4, 2, 0, ":1103-1118"), (4, 2, 1, ":1120-1129")):
        fname = f"_auto_range_{n}"
        frames[fname] = {"time": {"dtype": "i8", "data": list(range(n))},
                         "x": {"dtype": "i4", "data": list(range(n))},
                         "y": {"dtype": "i4", "data": list(range(n))}}
        w = mult * n + extra
        sol = np.zeros((w, w), int)
        if (n, extra) == (10, 0):
            np.fill_diagonal(sol, 1)
        elif (n, extra) == (10, 1):
            np.fill_diagonal(sol, 1)
            sol[5, 5] = 0
        elif (n, extra) == (4, 0):
            np.fill_diagonal(sol, 1)
            sol[np.array([tuple(range(1, 4, 2))])] = 0
            sol[np.array([tuple(range(4, 8, 2))])] = 0
        else:
            sol[0, 0] = sol[3, 3] = sol[6, 6] = sol[8, 8] = 1
        add(f"test_auto_range_points/n={n},w={w}", TP + cite,
            {"kind": "count", "column": "time"}, sol,
            canvas={"W": w, "H": w, "x_range": None, "y_range": None}, df=fname,
            x_range=[0, n - 1], y_range=[0, n - 1])
    # test_uniform_points, test_pandas.py:1132-1144
    n = 101
    frames["_uniform"] = {
        "time": {"dtype": "i4", "data": [1] * (2 * n)},
        "x": {"dtype": "f8", "data": [float(v) for v in np.concatenate((np.arange(n), np.arange(n)))]},
        "y": {"dtype": "f8", "data": [0.0] * n + [1.0] * n}}
    add("test_uniform_points", TP + ":1132-1144", {"kind": "count", "column": "time"},
        np.array([[10] * 9 + [11], [10] * 9 + [11]]),
        canvas={"W": 10, "H": 2, "x_range": None, "y_range": [0, 1]}, df="_uniform",
        x_range=[0, 100], y_range=[0, 1])
    # test_points_on_edge, test_pandas.py:1279-1296
    frames["_edge"] = {"x": {"dtype": "f8", "data": [0, 0.5, 1.1, 1.5, 2.2, 3, 3, 0]},
                       "y": {"dtype": "f8", "data": [0, 0, 0, 0, 0, 0, 3, 3]}}
    add("test_points_on_edge", TP + ":1279-1296", {"kind": "count", "column": None},
        np.array([[2, 2, 2], [0, 0, 0], [1, 0, 1]]),
        canvas={"W": 3, "H": 3, "x_range": [0, 3], "y_range": [0, 3]}, df="_edge",
        x_coords=[0.5, 1.5, 2.5], y_coords=[0.5, 1.5, 2.5])
    return cases, frames


def glyph_line_cases():
    """Kernel-level line cases of test_glyphs.py (draw_segment / extend_line helpers)."""
    cases = []
    seg_view = dict(sx=1, tx=0, sy=1, ty=0, xmin=0, xmax=5, ymin=0, ymax=5)

    def seg(name, cite, x0, y0, x1, y1, start, expected=None, total=None, at=None):
        cases.append(dict(name=name, cite=cite, kind="segment", view=seg_view, W=5, H=5,
                          p0=[x0, y0], p1=[x1, y1], segment_start=start,
                          expected=None if expected is None else enc(expected),
                          total=total, at=at))

    cite = TG + ":96-142"
    out = np.array([[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, 0, 1, 0, 0], [0, 0, 0, 1, 0],
                    [0, 0, 0, 0, 0]])
    seg("test_draw_line/diag", cite, 0, 0, 3, 3, True, out)
    seg("test_draw_line/diag_rev", cite, 3, 3, 0, 0, True, out)
    o2 = out.copy(); o2[0, 0] = 0
    seg("test_draw_line/diag_nostart", cite, 0, 0, 3, 3, False, o2)
    o3 = o2.copy(); o3[0, 0] = 1; o3[3, 3] = 0
    seg("test_draw_line/diag_rev_nostart", cite, 3, 3, 0, 0, False, o3)
    out = np.array([[0, 0, 0, 0, 0], [0, 0, 0, 1, 0], [0, 0, 1, 0, 0], [0, 1, 0, 0, 0],
                    [1, 0, 0, 0, 0]])
    seg("test_draw_line/flip", cite, 0, 4, 3, 1, True, out)
    seg("test_draw_line/flip_rev", cite, 3, 1, 0, 4, True, out)
    o2 = out.copy(); o2[4, 0] = 0
    seg("test_draw_line/flip_nostart", cite, 0, 4, 3, 1, False, o2)
    cite = TG + ":145-166"
    seg("test_draw_line_same_point/start", cite, 4, 4, 4, 4, True, total=1, at=[4, 4, 1])
    seg("test_draw_line_same_point/nostart", cite, 4, 4, 4, 4, False, total=1, at=[4, 4, 1])
    seg("test_draw_line_same_point/clipped_start", cite, 4, 4, 10, 10, True, total=1, at=[4, 4, 1])
    seg("test_draw_line_same_point/clipped_nostart", cite, 4, 4, 10, 10, False, total=0, at=[4, 4, 0])
    cite = TG + ":169-183"
    out = np.zeros((5, 5), int); out[:4, 3] = 1
    seg("test_draw_line_vertical_horizontal/vertical", cite, 3, 3, 3, 0, True, out)
    out = np.zeros((5, 5), int); out[3, :4] = 1
    seg("test_draw_line_vertical_horizontal/horizontal", cite, 3, 3, 0, 3, True, out)

    ext_view = dict(sx=1.0, tx=3.0, sy=1.0, ty=3.0, xmin=-3, xmax=1, ymin=-3, ymax=1)

    def ext(name, cite, xs, ys, start, W, expected=None, total=None):
        cases.append(dict(name=name, cite=cite, kind="axis0", view=ext_view, W=W, H=W,
                          xs=enc(np.asarray(xs, float)), ys=enc(np.asarray(ys, float)),
                          plot_start=start, expected=None if expected is None else enc(expected),
                          total=total))

    cite = TG + ":186-215"
    out = np.array([[0, 0, 0, 0, 0], [0, 0, 0, 0, 0], [0, 1, 1, 1, 0], [0, 1, 0, 1, 0],
                    [0, 0, 0, 0, 0]])
    xs = [0, -2, -2, 0, 0]; ys = [-1, -1, 1.1, 1.1, -1]
    ext("test_extend_lines/box_nostart", cite, xs, ys, False, 5, out)
    o2 = out.copy(); o2[2, 3] += 1
    ext("test_extend_lines/box_start", cite, xs, ys, True, 5, o2)
    out = np.array([[0, 1, 0, 1, 0], [1, 0, 0, 1, 0], [0, 0, 0, 0, 0], [1, 1, 0, 1, 0],
                    [0, 0, 0, 0, 0]])
    ext("test_extend_lines/clipped", cite, [2, 1, 0, -1, -4, -1, -100, -1, 2],
        [-1, -2, -3, -4, -1, 2, 100, 2, -1], True, 5, out)
    ext("test_extend_lines_all_out_of_bounds", TG + ":218-226", [-100, -200, -100], [0, 0, 1],
        True, 5, total=0)
    nan = float("nan")
    ext("test_extend_lines_nan", TG + ":229-238", [-3, -2, nan, 0, 1], [-3, -2, nan, 0, 1],
        True, 5, np.diag([1, 1, 0, 1, 0]))
    cite = TG + ":241-262"
    xs = [-3, 1, 1, -3, -3]; ys = [-3, -3, 1, 1, -3]
    ext("test_extend_lines_exact_bounds/start", cite, xs, ys, True, 4,
        np.array([[2, 1, 1, 1], [1, 0, 0, 1], [1, 0, 0, 1], [1, 1, 1, 1]]))
    ext("test_extend_lines_exact_bounds/nostart", cite, xs, ys, False, 4,
        np.array([[1, 1, 1, 1], [1, 0, 0, 1], [1, 0, 0, 1], [1, 1, 1, 1]]))
    return cases


def api_line_cases():
    """Canvas.line cases of test_pandas.py."""
    cases = []
    sol7 = np.array([[0, 0, 1, 0, 1, 0, 0], [0, 1, 0, 0, 0, 1, 0], [1, 0, 0, 0, 0, 0, 1],
                     [0, 0, 0, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0, 1], [0, 1, 0, 0, 0, 1, 0],
                     [0, 0, 1, 0, 1, 0, 0]])
    c7 = {"W": 7, "H": 7, "x_range": [-3, 3], "y_range": [-3, 3]}

    def add(name, cite, frame, kwargs, canvas, expected, **kw):
        cases.append(dict(name=name, cite=cite, frame=frame, kwargs=kwargs, canvas=canvas,
                          expected=None if expected is None else enc(expected), **kw))

    sol = sol7.copy(); sol[5, 1] = 2
    add("test_line", TP + ":1256-1276",
        {"x": [4, 0, -4, -3, -2, -1.9, 0, 10, 10, 0, 4], "y": [0, -4, 0, 1, 2, 2.1, 4, 20, 30, 4, 0]},
        {"x": "x", "y": "y", "agg": {"kind": "count", "column": None}}, c7, sol,
        x_range=[-3, 3], y_range=[-3, 3])
    add("test_lines_on_edge", TP + ":1299-1316", {"x": [0, 3, 3, 0], "y": [0, 0, 3, 3]},
        {"x": "x", "y": "y", "agg": {"kind": "count", "column": None}},
        {"W": 3, "H": 3, "x_range": [0, 3], "y_range": [0, 3]},
        np.array([[1, 1, 1], [0, 0, 1], [1, 1, 1]]))
    add("test_auto_range_line", TP + ":1358-1375",
        {"x": [-10, 0, 10, 0, -10], "y": [0, 10, 0, -10, 0]},
        {"x": "x", "y": "y", "agg": {"kind": "count", "column": None}},
        {"W": 5, "H": 5, "x_range": None, "y_range": None},
        np.array([[0, 0, 1, 0, 0], [0, 1, 0, 1, 0], [2, 0, 0, 0, 1], [0, 1, 0, 1, 0],
                  [0, 0, 1, 0, 0]]), x_range=[-10, 10], y_range=[-10, 10])
    # test_bug_570: only pixels (73..299, 590) are 1.
    b = [[int(r), 590] for r in range(73, 300)]
    add("test_bug_570", TP + ":1635-1652",
        {"Time": [1456353642.2053893, 1456353642.2917893],
         "data": [-59.4948743433377, 506.4847376716022]},
        {"x": "Time", "y": "data", "agg": {"kind": "count", "column": None}},
        {"W": 1000, "H": 300, "x_range": [1456323293.9859753, 1456374687.0009754],
         "y_range": [-228.56721300380943, 460.4042291124646]}, None, ones_at=b)
    nan = None
    cite = TP + ":1656-1751"
    add("test_line_manual_range/axis1_none_constant", cite,
        {"x0": [4, -4], "x1": [0, 0], "x2": [-4, 4], "y0": [0, 0], "y1": [-4, 4], "y2": [0, 0]},
        {"x": ["x0", "x1", "x2"], "y": ["y0", "y1", "y2"], "axis": 1,
         "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_manual_range/axis1_x_constant", cite,
        {"y0": [0, 0], "y1": [-4, 4], "y2": [0, 0]},
        {"x": {"ndarray": [-4, 0, 4]}, "y": ["y0", "y1", "y2"], "axis": 1,
         "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_manual_range/axis1_y_constant", cite,
        {"x0": [0, 0], "x1": [-4, 4], "x2": [0, 0]},
        {"x": ["x0", "x1", "x2"], "y": {"ndarray": [-4, 0, 4]}, "axis": 1,
         "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_manual_range/axis0_single", cite,
        {"x": [0, -4, 0, nan, 0, 4, 0], "y": [-4, 0, 4, nan, -4, 0, 4]},
        {"x": "x", "y": "y", "axis": 0, "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_manual_range/axis0_multi", cite,
        {"x0": [0, -4, 0], "x1": [0, 4, 0], "y0": [-4, 0, 4], "y1": [-4, 0, 4]},
        {"x": ["x0", "x1"], "y": ["y0", "y1"], "axis": 0,
         "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_manual_range/axis0_multi_str", cite,
        {"x0": [0, -4, 0], "x1": [0, 4, 0], "y0": [-4, 0, 4], "y1": [-4, 0, 4]},
        {"x": ["x0", "x1"], "y": "y0", "axis": 0,
         "agg": {"kind": "count", "column": None}}, c7, sol7)
    sol9 = np.array([[0, 0, 0, 0, 2, 0, 0, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                     [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                     [1, 0, 0, 0, 0, 0, 0, 0, 1], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                     [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                     [0, 0, 0, 0, 2, 0, 0, 0, 0]])
    c9 = {"W": 9, "H": 9, "x_range": None, "y_range": None}
    cite = TP + ":1754-1860"
    add("test_line_autorange/axis1_none_constant", cite,
        {"x0": [0, 0], "x1": [-4, 4], "x2": [0, 0], "y0": [-4, -4], "y1": [0, 0], "y2": [4, 4]},
        {"x": ["x0", "x1", "x2"], "y": ["y0", "y1", "y2"], "axis": 1,
         "agg": {"kind": "count", "column": None}}, c9, sol9, x_range=[-4, 4], y_range=[-4, 4])
    add("test_line_autorange/axis1_y_constant", cite,
        {"x0": [0, 0], "x1": [-4, 4], "x2": [0, 0]},
        {"x": ["x0", "x1", "x2"], "y": {"ndarray": [-4, 0, 4]}, "axis": 1,
         "agg": {"kind": "count", "column": None}}, c9, sol9, x_range=[-4, 4], y_range=[-4, 4])
    add("test_line_autorange/axis0_single", cite,
        {"x": [0, -4, 0, nan, 0, 4, 0], "y": [-4, 0, 4, nan, -4, 0, 4]},
        {"x": "x", "y": "y", "axis": 0, "agg": {"kind": "count", "column": None}}, c9, sol9,
        x_range=[-4, 4], y_range=[-4, 4])
    add("test_line_autorange/axis0_multi", cite,
        {"x0": [0, -4, 0], "x1": [0, 4, 0], "y0": [-4, 0, 4], "y1": [-4, 0, 4]},
        {"x": ["x0", "x1"], "y": ["y0", "y1"], "axis": 0,
         "agg": {"kind": "count", "column": None}}, c9, sol9, x_range=[-4, 4], y_range=[-4, 4])
    add("test_line_autorange/axis0_multi_str", cite,
        {"x0": [0, -4, 0], "x1": [0, 4, 0], "y0": [-4, 0, 4], "y1": [-4, 0, 4]},
        {"x": ["x0", "x1"], "y": "y0", "axis": 0,
         "agg": {"kind": "count", "column": None}}, c9, sol9, x_range=[-4, 4], y_range=[-4, 4])
    solx = np.array([[0, 0, 0, 0, 1, 0, 0, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                     [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                     [2, 0, 0, 0, 0, 0, 0, 0, 2], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                     [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                     [0, 0, 0, 0, 1, 0, 0, 0, 0]])
    add("test_line_manual_range/axis1_ragged", TP + ":1703-1707",
        {"x": {"ragged": [[4, 0], [0, -4, 0, 4]], "dtype": "f4"},
         "y": {"ragged": [[0, -4], [-4, 0, 4, 0]], "dtype": "f4"}},
        {"x": "x", "y": "y", "axis": 1, "agg": {"kind": "count", "column": None}}, c7, sol7)
    add("test_line_autorange_axis1_ragged", TP + ":1935-1967",
        {"x": {"ragged": [[4, 0], [0, -4, 0, 4]], "dtype": "f4"},
         "y": {"ragged": [[0, -4], [-4, 0, 4, 0]], "dtype": "f4"}},
        {"x": "x", "y": "y", "axis": 1, "agg": {"kind": "count", "column": None}}, c9,
        np.array([[0, 0, 0, 0, 2, 0, 0, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                  [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                  [1, 0, 0, 0, 0, 0, 0, 0, 2], [0, 1, 0, 0, 0, 0, 0, 1, 0],
                  [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 0, 0, 1, 0, 1, 0, 0, 0],
                  [0, 0, 0, 0, 1, 0, 0, 0, 0]]), x_range=[-4, 4], y_range=[-4, 4])
    add("test_line_autorange_axis1_x_constant", TP + ":1863-1895",
        {"y0": [0, 0], "y1": [-4, 4], "y2": [0, 0]},
        {"x": {"ndarray": [-4, 0, 4]}, "y": ["y0", "y1", "y2"], "axis": 1,
         "agg": {"kind": "count", "column": None}}, c9, solx)
    n = float("nan")
    sols = np.array([[n, n, 7, n, 7, n, n], [n, 7, n, n, n, 7, n], [7, n, n, n, n, n, 7],
                     [n, n, n, n, n, n, n], [9, n, n, n, n, n, 9], [n, 9, n, n, n, 9, n],
                     [n, n, 9, n, 9, n, n]])
    add("test_line_agg_sum_axis1_none_constant", TP + ":1899-1932",
        {"x0": [4, -4], "x1": [0, 0], "x2": [-4, 4], "y0": [0, 0], "y1": [-4, 4], "y2": [0, 0],
         "v": [7, 9]},
        {"x": ["x0", "x1", "x2"], "y": ["y0", "y1", "y2"], "axis": 1,
         "agg": {"kind": "sum", "column": "v"}}, c7, sols)
    return cases


def noaa_cases(ref):
    """Touched-pixel masks of the reference's non-antialiased line images."""
    gens = {
        "test_001": [((1, 1), (49, a)) for a in range(1, 55, 6)],
        "test_002": [((1, 1), (a, 49)) for a in range(1, 55, 6)],
        "test_003": [((49, 49), (1, a)) for a in range(1, 55, 6)],
        "test_004": [((49, 49), (a, 1)) for a in range(1, 55, 6)],
        "test_005": [((1, 1), (1, 49)), ((1, 49), (2, 1)), ((2, 1), (2, 49)), ((2, 49), (4, 1)),
                     ((4, 1), (4, 49)), ((4, 49), (8, 1)), ((8, 1), (8, 49)), ((8, 49), (16, 1)),
                     ((16, 1), (16, 49)), ((16, 49), (32, 1)), ((32, 1), (32, 49)),
                     ((32, 49), (49, 1)), ((49, 1), (49, 49))],
        "test_007": [((0.5, 0.5), (1.5, 48.5)), ((0.5, 0.5), (48.5, 1.5)),
                     ((48.5, 48.5), (47.5, 0.5)), ((48.5, 48.5), (0.5, 47.5))],
    }
    multi = {"test_006": [(1, 1), (1, 49), (2, 1), (2, 49), (4, 1), (4, 49), (8, 1), (8, 49),
                          (16, 1), (16, 49), (32, 1), (32, 49), (49, 1), (49, 49)]}
    canvases = {"normal": {"W": 50, "H": 50, "x_range": [0, 49], "y_range": [0, 49]},
                "clipped": {"W": 20, "H": 20, "x_range": [14, 34], "y_range": [14, 34]}}
    cases = []
    if ref is None:
        return cases
    from PIL import Image
    for cname, canvas in canvases.items():
        for name, lines in list(gens.items()) + list(multi.items()):
            png = os.path.join(ref, "datashader", "tests", "data", f"{name}_noaa_{cname}.png")
            img = np.asarray(Image.open(png).convert("RGB"))
            if (img == img[0, 0]).all():
                # A uniform image is an all-zero aggregate (no line enters the canvas):
                # shading maps every pixel to one colour, so no pixel is touched.
                touched = np.zeros(img.shape[:2], bool)
            else:
                touched = np.flipud(~np.all(img == 255, axis=2))
            if name in multi:
                polylines = [[list(p) for p in multi[name]]]
            else:
                polylines = [[list(p0), list(p1)] for p0, p1 in lines]
            cases.append(dict(name=f"{name}_noaa_{cname}", cite=TA + ":302-370",
                              canvas=canvas, polylines=polylines, expected=enc(touched)))
    return cases


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    if not os.path.isdir(ref):
        ref = None
    frame = pandas_frame()
    pcases, frames = point_cases(frame)
    frames["_pandas"] = frame
    doc = {
        "source": "holoviz/datashader 0.19.1 test-suite (datashader/tests)",
        "frames": frames,
        "points": pcases,
        "glyph_lines": glyph_line_cases(),
        "lines": api_line_cases(),
        "noaa_masks": noaa_cases(ref),
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(pcases)} point, {len(doc['glyph_lines'])} glyph-line, "
          f"{len(doc['lines'])} line, {len(doc['noaa_masks'])} mask cases")


if __name__ == "__main__":
    main()

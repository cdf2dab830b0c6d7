"""Host-side logic of the Canvas API (CPU only): validation errors, glyph selection,
reduction planning, coordinates and the xarray stand-in."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import datashader_amd as dsa
from datashader_amd import backend, glyphs, xrlite


def test_points_validation_errors():
    df = pd.DataFrame({"x": [0.0], "y": [0.0], "s": ["a"], "v": [1.0],
                       "c": pd.Categorical(["a"])})
    c = dsa.Canvas(2, 2, x_range=(0, 1), y_range=(0, 1))
    with pytest.raises(ValueError, match="x must be real"):
        c.points(df, "s", "y")
    with pytest.raises(ValueError, match="y must be real"):
        c.points(df, "x", "s")
    # with columns to prune, the reference's df[cols_to_keep] raises KeyError first
    # (core.py:1384-1392); without, the reduction's validation raises ValueError
    with pytest.raises(KeyError):
        c.points(df, "x", "y", dsa.sum("nope"))
    with pytest.raises(ValueError, match="specified column not found"):
        c.points(df[["x", "y"]], "x", "y", dsa.sum("nope"))
    with pytest.raises(ValueError, match="input must be numeric"):
        c.points(df, "x", "y", dsa.sum("s"))
    with pytest.raises(ValueError, match="input must be categorical"):
        c.points(df, "x", "y", dsa.by("v", dsa.count()))
    with pytest.raises(ValueError, match="input must be an integer column"):
        c.points(df, "x", "y", dsa.by(dsa.category_modulo("v", 3), dsa.count()))
    with pytest.raises(TypeError, match="first argument must be a column name"):
        dsa.by(3, dsa.count())
    with pytest.raises(ValueError, match="Invalid size"):
        dsa.Canvas(0, 2).points(df, "x", "y")
    with pytest.raises(ValueError, match="Range values must be >0"):
        dsa.Canvas(2, 2, x_range=(0, 1), x_axis_type="log").points(df, "x", "y")
    with pytest.raises(ValueError, match="coordinates may be specified"):
        c.points(df, "x")
    with pytest.raises(NotImplementedError):
        c.points(df, geometry="g")


def test_line_validation_and_glyph_selection(monkeypatch):
    seen = {}

    def fake(source, schema, canvas, glyph, agg, **kw):
        seen["glyph"] = glyph
        return glyph

    monkeypatch.setitem(dsa.bypixel.pipeline._lookup, pd.DataFrame, fake)
    df = pd.DataFrame({"a": [0.0, 1.0], "b": [0.0, 1.0], "c": [1.0, 2.0]})
    c = dsa.Canvas(4, 4, x_range=(0, 1), y_range=(0, 1))
    assert type(c.line(df, "a", "b")) is glyphs.LineAxis0
    assert type(c.line(df, ["a", "b"], ["b", "c"])) is glyphs.LineAxis0Multi
    assert type(c.line(df, ["a", "b"], "c")) is glyphs.LineAxis0Multi
    assert type(c.line(df, ["a", "b"], ["b", "c"], axis=1)) is glyphs.LinesAxis1
    assert type(c.line(df, np.array([0.0, 1.0]), ["b", "c"], axis=1)) is glyphs.LinesAxis1XConstant
    assert type(c.line(df, ["a", "b"], np.array([0.0, 1.0]), axis=1)) is glyphs.LinesAxis1YConstant
    with pytest.raises(ValueError, match="x and y coordinate lengths do not match"):
        c.line(df, x=np.array([0.0, 1.0, 2.0]), y=["a", "b"], axis=1)
    with pytest.raises(ValueError, match="x columns must have the same data type"):
        d2 = df.assign(i=np.array([1, 2], dtype="i4"))
        c.line(d2, ["a", "i"], ["b", "c"], axis=1)
    with pytest.raises(ValueError, match="The axis argument"):
        c.line(df, "a", "b", axis=2)
    with pytest.raises(ValueError, match="Invalid combination"):
        c.line(df, np.array([0.0]), "b", axis=0)
    with pytest.raises(ValueError, match="Do not specify values for both"):
        c.line(df, "a", "b", line_width=2, antialias=True)
    with pytest.raises(NotImplementedError):
        c.line(df, "a", "b", line_width=1)


def test_plan_groups_reductions_per_value_column():
    agg = dsa.summary(n=dsa.count(), m=dsa.mean("v"), s=dsa.std("v"), w=dsa.max("w"),
                      k=dsa.by("cat", dsa.count()))
    passes, leaves = backend.plan(agg)
    keys = {(k[0], None if k[1] is None else k[1].column): p for k, p in passes.items()}
    assert set(keys) == {("v", None), ("w", None), (None, "cat")}
    assert keys[("v", None)].planes == {"count", "countv", "sum"}
    assert keys[("v", None)].m2
    assert keys[("w", None)].planes == {"max"}
    assert [n for n, _, _ in leaves] == ["k", "m", "n", "s", "w"]   # summary keys sorted


def test_coordinates_and_view():
    c = dsa.Canvas(7, 5, x_range=(-3, 3), y_range=(0.5, 2.5))
    v = backend.make_view(c, c.x_range, c.y_range)
    assert v["sx"] == 7 / 6 and v["tx"] == 3 * (7 / 6)
    assert v["xxmax"] == 7 and v["yymax"] == 5
    idx = c.x_axis.compute_index(v["x_st"], 7)
    np.testing.assert_array_equal(idx, (np.arange(7) + 0.5 - v["tx"]) / v["sx"])
    # float32 ranges keep float32 scalar arithmetic (numpy 2 promotion rules)
    r32 = (np.float32(0.0), np.float32(1.0))
    v32 = backend.make_view(dsa.Canvas(2048, 4), r32, (0.0, 1.0))
    assert isinstance(v32["sx"], np.float32) and v32["xxmax"] == 2048


def test_xrlite():
    a = xrlite.DataArray(np.arange(6).reshape(2, 3), coords=[[0.5, 1.5], [1, 2, 3]],
                         dims=["y", "x"], attrs={"x_range": (0, 1)})
    assert a.x_range == (0, 1) and a.dims == ("y", "x") and a.shape == (2, 3)
    assert a.coords["x"].values.tolist() == [1, 2, 3]
    b = xrlite.DataArray(np.arange(6).reshape(2, 3), coords=[[0.5, 1.5], [1, 2, 3]],
                         dims=["y", "x"])
    assert a.equals(b)
    sub = a[:, 1]
    assert sub.dims == ("y",) and sub.values.tolist() == [1, 4]
    ds = xrlite.Dataset({"a": a}, attrs={"x_range": (0, 1)})
    assert ds["a"] is a and ds.a is a and ds.x_range == (0, 1)


def test_missing_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    df = pd.DataFrame({"x": [0.5], "y": [0.5]})
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        dsa.Canvas(2, 2, x_range=(0, 1), y_range=(0, 1)).points(df, "x", "y")

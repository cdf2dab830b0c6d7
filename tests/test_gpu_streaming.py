"""Chunked (out-of-core) aggregation on the GPU: a ChunkedFrame streamed chunk by chunk
gives the resident result -- counts, min and max exactly, float64 sums and means within
1e-12 relative, var / std within 1e-9 relative (the chunks change only the order of the
float64 additions) -- and its counts equal the oracle's."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import datashader_amd as dsa
import oracle

pytestmark = pytest.mark.gpu


def _table(n, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(0.3, 0.05, n // 2), rng.uniform(-1, 1, n - n // 2)])
    y = np.concatenate([rng.normal(-0.2, 0.08, n // 2), rng.uniform(-1, 1, n - n // 2)])
    perm = rng.permutation(n)
    x, y = x[perm].astype("f4"), y[perm].astype("f4")
    x[rng.random(n) < 0.001] = np.nan
    v = rng.lognormal(0, 1, n)
    v[rng.random(n) < 0.01] = np.nan
    cat = pd.Categorical(rng.choice(list("abcde"), n), categories=list("abcde"))
    return pd.DataFrame({"x": x, "y": y, "v": v, "cat": cat})


def _close(a, b, rtol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    np.testing.assert_allclose(a[ok], b[ok], rtol=rtol, atol=0)


AGG = dict(n=dsa.count(), nv=dsa.count("v"), s=dsa.sum("v"), m=dsa.mean("v"), lo=dsa.min("v"),
           hi=dsa.max("v"), var=dsa.var("v"), sd=dsa.std("v"), any=dsa.any(),
           bc=dsa.by("cat", dsa.count()), bm=dsa.by("cat", dsa.max("v")))


@pytest.mark.parametrize("ranges", [None, ((-1, 1), (-1, 1))], ids=["autorange", "ranges"])
@pytest.mark.parametrize("chunk_rows", [77_777, 1 << 20], ids=["many-chunks", "one-chunk"])
def test_points_stream_equals_resident(chunk_rows, ranges):
    df = _table(600_000, 3)
    kw = {} if ranges is None else dict(x_range=ranges[0], y_range=ranges[1])
    cvs = dsa.Canvas(233, 171, **kw)
    resident = cvs.points(df, "x", "y", dsa.summary(**AGG))
    parts = [df.iloc[:250_000], df.iloc[250_000:]]
    streamed = cvs.points(dsa.ChunkedFrame(parts, chunk_rows=chunk_rows), "x", "y",
                          dsa.summary(**AGG))
    assert streamed.attrs == resident.attrs
    for k in ("n", "nv", "lo", "hi", "any", "bc", "bm"):
        np.testing.assert_array_equal(streamed[k].values, resident[k].values, err_msg=k)
    for k in ("s", "m"):
        _close(streamed[k].values, resident[k].values, 1e-12)
    for k in ("var", "sd"):
        _close(streamed[k].values, resident[k].values, 1e-9)
    for k in ("x", "y"):
        np.testing.assert_array_equal(streamed["n"].coords[k].values,
                                      resident["n"].coords[k].values)
    xr, yr = resident.attrs["x_range"], resident.attrs["y_range"]
    sx, sy = 233 / (xr[1] - xr[0]), 171 / (yr[1] - yr[0])
    view = oracle.View(sx, -xr[0] * sx, sy, -yr[0] * sy, xr[0], xr[1], yr[0], yr[1])
    ref = oracle.points(df["x"].to_numpy(), df["y"].to_numpy(), view, 233, 171, ("count",))
    np.testing.assert_array_equal(streamed["n"].values, ref["count"])


def test_lines_axis1_stream_equals_resident():
    rng = np.random.default_rng(5)
    ys = np.cumsum(rng.normal(0, 0.05, (20_000, 48)), axis=1).astype("f4")
    names = [f"y{j}" for j in range(48)]
    df = pd.DataFrame(ys, columns=names)
    df["v"] = rng.uniform(0, 1, len(df))
    xs = np.linspace(0, 1, 48)
    cvs = dsa.Canvas(300, 200)
    for agg in (dsa.count(), dsa.sum("v"), dsa.max("v")):
        resident = cvs.line(df, x=xs, y=names, agg=agg, axis=1)
        streamed = cvs.line(dsa.ChunkedFrame(df, chunk_rows=3_001), x=xs, y=names, agg=agg,
                            axis=1)
        assert streamed.attrs == resident.attrs
        _close(streamed.values, resident.values, 1e-12)


def _walk(n, seed, nan_every=0):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.normal(0, 0.05, n))
    y = np.cumsum(rng.normal(0, 0.05, n))
    if nan_every:
        x[rng.random(n) < 1.0 / nan_every] = np.nan
    return pd.DataFrame({"x": x, "y": y, "v": rng.uniform(0, 1, n),
                         "x2": x[::-1].copy(), "y2": y + 0.3})


@pytest.mark.parametrize("chunk_rows", [1, 2, 3, 997, 40_000])
def test_axis0_line_stream_equals_resident(chunk_rows):
    """LineAxis0 over one partition: a chunk boundary is invisible (the joining segment is
    drawn once, with the segment_start the unchunked line gives it), NaN gaps included."""
    df = _walk(3000 if chunk_rows < 4 else 30_000, 8, nan_every=50)
    cvs = dsa.Canvas(157, 113)
    for agg in (dsa.count(), dsa.max("v"), dsa.sum("v")):
        resident = cvs.line(df, "x", "y", agg=agg)
        streamed = cvs.line(dsa.ChunkedFrame(df, chunk_rows=chunk_rows), "x", "y", agg=agg)
        assert streamed.attrs == resident.attrs
        _close(streamed.values, resident.values, 1e-12)


def test_axis0_line_partitions_follow_dask_stitching():
    """LineAxis0 over partitions: partition i > 0 is drawn as the previous partition's
    last row followed by its own rows with plot_start False (dask.py:242-248); the
    expected canvas is the oracle's sum of those drawings."""
    df = _walk(20_000, 9, nan_every=40)
    cuts = [0, 5_000, 5_001, 12_345, 20_000]
    parts = [df.iloc[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    W, H = 140, 90
    cvs = dsa.Canvas(W, H)
    got = cvs.line(dsa.ChunkedFrame(parts, chunk_rows=1_000), "x", "y", agg=dsa.count())
    xr, yr = got.attrs["x_range"], got.attrs["y_range"]
    sx, sy = W / (xr[1] - xr[0]), H / (yr[1] - yr[0])
    view = oracle.View(sx, -xr[0] * sx, sy, -yr[0] * sy, xr[0], xr[1], yr[0], yr[1])
    want = np.zeros((H, W), np.int64)
    for i, p in enumerate(parts):
        rows = p if i == 0 else pd.concat([parts[i - 1].iloc[-1:], p])
        want += oracle.line_axis0(rows["x"].to_numpy(), rows["y"].to_numpy(), view, W, H,
                                  plot_start=(i == 0))["count"]
    np.testing.assert_array_equal(got.values, want)


def test_axis0_multi_stream():
    """LinesAxis0Multi: chunks of one partition equal the resident lines; partitions are
    drawn one by one (dask's default path)."""
    df = _walk(12_000, 10, nan_every=30)
    cvs = dsa.Canvas(120, 80, x_range=(-3, 3), y_range=(-3, 3))
    kw = dict(x=["x", "x2"], y=["y", "y2"], axis=0)
    resident = cvs.line(df, agg=dsa.count(), **kw)
    streamed = cvs.line(dsa.ChunkedFrame(df, chunk_rows=777), agg=dsa.count(), **kw)
    np.testing.assert_array_equal(streamed.values, resident.values)
    parts = [df.iloc[:4_000], df.iloc[4_000:]]
    got = cvs.line(dsa.ChunkedFrame(parts, chunk_rows=999), agg=dsa.count(), **kw)
    want = sum(cvs.line(p, agg=dsa.count(), **kw).values.astype(np.int64) for p in parts)
    np.testing.assert_array_equal(got.values, want)

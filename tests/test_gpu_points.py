"""GPU parity of the points path: libdsx (through Canvas.points and through the C ABI) vs
the reference's expectations and the CPU oracle.

Tolerances (float32 value columns, against the float64 oracle):
  count, count(v), any, by(...) counts, min, max   bit-exact
  sum, mean                                         max per-bin relative error <= 1e-6
  var, std                                          max per-bin relative error <= 1e-5
float64 value columns accumulate in float64 and are held to 1e-12.
"""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import oracle
from oracle import closed_forms as cf

import golden_utils as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import datashader_amd as dsa  # noqa: E402

GC = G.cases()


def _red(spec):
    """Build a datashader_amd reduction from a golden-case spec."""
    k = spec["kind"]
    if k == "summary":
        return dsa.summary(**{n: _red(v) for n, v in spec["items"].items()})
    if k == "count_cat":
        return dsa.count_cat(spec["column"])
    if k == "by":
        cat = spec["cat"]
        if isinstance(cat, dict):
            kind = cat.pop("kind") if "kind" in cat else None
            cat = dict(cat)
            if kind == "category_modulo":
                cat = dsa.category_modulo(cat["column"], cat["modulo"], cat.get("offset", 0))
            else:
                cat = dsa.category_binning(cat["column"], cat["lower"], cat["upper"],
                                           cat["nbins"])
        return dsa.by(cat, _red(spec["reduction"]))
    if k == "where":
        return dsa.where(_red(spec["selector"]), spec["lookup"])
    cls = getattr(dsa, k)
    col = spec.get("column")
    if k in ("_min_row_index", "_max_row_index"):
        return cls()
    if k in ("_min_n_row_index", "_max_n_row_index"):
        return cls(n=spec["n"])
    if "n" in spec:
        return cls(col, n=spec["n"])
    return cls() if col is None and k in ("count", "any") else cls(col)


def _copy(spec):
    import copy
    return copy.deepcopy(spec)


def _canvas(cv):
    return dsa.Canvas(plot_width=cv["W"], plot_height=cv["H"], x_range=cv["x_range"],
                      y_range=cv["y_range"])


@pytest.mark.parametrize("source", ["pandas", "device"])
@pytest.mark.parametrize("case", GC["points"], ids=[c["name"] for c in GC["points"]])
def test_points_golden(case, source):
    df = G.frame(case["frame"])
    src = df if source == "pandas" else dsa.DeviceFrame.from_pandas(df)
    agg = _canvas(case["canvas"]).points(src, case["x"], case["y"], _red(_copy(case["agg"])))
    if "pick" in case:
        agg = agg[case["pick"]]
    G.assert_same(agg.values, G.expected_array(case))
    if "x_range" in case:
        assert np.allclose(agg.attrs["x_range"], case["x_range"])
        assert np.allclose(agg.attrs["y_range"], case["y_range"])
    ndim = ["n"] if case.get("n_dim") else []
    if "cat_dim" in case:
        assert list(agg.dims) == [case["y"], case["x"], case["cat_dim"]] + ndim
        assert list(agg.coords[case["cat_dim"]].values) == case["cat_coords"]
    else:
        assert list(agg.dims) == [case["y"], case["x"]] + ndim
    if ndim:
        assert agg.coords["n"].values.tolist() == list(range(agg.shape[-1]))
    if "x_coords" in case:
        assert agg.coords["x"].values.tolist() == case["x_coords"]
        assert agg.coords["y"].values.tolist() == case["y_coords"]
    spec = case["agg"] if "pick" not in case else case["agg"]["items"][case["pick"]]
    k = spec["kind"]
    if k == "by":
        k = spec["reduction"]["kind"]
    exp_dtype = {"count": np.uint32, "any": np.bool_, "count_cat": np.uint32}.get(k, np.float64)
    if "dtype" in case:
        exp_dtype = np.dtype(case["dtype"])
    assert agg.values.dtype == exp_dtype


def test_coords_match_reference_formula():
    """core.py:83-99 coordinates, compared with the reference's own lincoords."""
    df = G.frame("_pandas")
    agg = dsa.Canvas(2, 2, x_range=(0, 1), y_range=(0, 1)).points(df, "x", "y")
    ax = dsa.LinearAxis()
    lin = ax.compute_index(ax.compute_scale_and_translate((0, 1), 2), 2)
    assert agg.coords["x"].values.tolist() == lin.tolist() == [0.25, 0.75]
    assert agg.x_range == (0, 1)


def test_log_axis_points():
    """test_pandas.py:1171-1190 (log_x / log_y columns on a 2x2 canvas)."""
    df = G.frame("_pandas")
    df["log_x"] = np.array([1.0] * 10 + [10] * 10)
    df["log_y"] = np.array([1.0] * 5 + [10] * 5 + [1] * 5 + [10] * 5)
    sol = np.array([[5, 5], [5, 5]])
    c = dsa.Canvas(2, 2, x_range=(1, 10), y_range=(0, 1), x_axis_type="log")
    agg = c.points(df, "log_x", "y", dsa.count("i32"))
    np.testing.assert_array_equal(agg.values, sol)
    ax = dsa.LogAxis()
    logc = ax.compute_index(ax.compute_scale_and_translate((1, 10), 2), 2)
    np.testing.assert_array_equal(agg.coords["log_x"].values, logc)
    c = dsa.Canvas(2, 2, x_range=(1, 10), y_range=(1, 10), x_axis_type="log",
                   y_axis_type="log")
    agg = c.points(df, "log_x", "log_y", dsa.count("i32"))
    np.testing.assert_array_equal(agg.values, sol)


@pytest.mark.parametrize("high", [9, 10, 99, 100])
def test_uniform_diagonal_points(high):
    """test_pandas.py:1147-1168."""
    n = high * high
    df = pd.DataFrame({"time": np.ones(n, dtype="i4"),
                       "x": np.array([np.arange(0, high, dtype="f8")] * high).flatten(),
                       "y": np.array([np.arange(0, high, dtype="f8")] * high).flatten()})
    agg = dsa.Canvas(2, 2, x_range=(0, high), y_range=(0, high)).points(df, "x", "y",
                                                                        dsa.count("time"))
    diagonal = agg.values.diagonal(0)
    assert sum(diagonal) == n
    assert abs(high) % 2 == abs(diagonal[1] / high - diagonal[0] / high)


# ---- random inputs vs the oracle ------------------------------------------------------

def _rel_err(got, exp):
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    m = ~np.isnan(exp)
    denom = np.maximum(np.abs(exp[m]), 1e-300)
    return float(np.max(np.abs(got[m] - exp[m]) / denom)) if m.any() else 0.0


def _all_aggs():
    return dsa.summary(cnt=dsa.count(), cntv=dsa.count("v"), any=dsa.any(), s=dsa.sum("v"),
                       mean=dsa.mean("v"), mn=dsa.min("v"), mx=dsa.max("v"),
                       var=dsa.var("v"), std=dsa.std("v"))


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("shape", [(97, 61), (800, 600)])
def test_random_points_vs_oracle(dtype, shape):
    W, H = shape
    rng = np.random.default_rng(7)
    n = 1_000_003
    x = rng.normal(0.0, 3.0, n).astype(dtype)
    y = rng.normal(0.5, 2.0, n).astype(dtype)
    v = rng.lognormal(0.0, 2.0, n).astype(dtype)
    for a in (x, y, v):
        a[rng.choice(n, n // 200, replace=False)] = np.nan
    df = pd.DataFrame({"x": x, "y": y, "v": v})
    xr_, yr_ = (-6.0, 7.5), (-3.0, 4.0)
    ds = dsa.Canvas(W, H, x_range=xr_, y_range=yr_).points(df, "x", "y", _all_aggs())
    view = G.view_for(W, H, xr_, yr_)
    r = oracle.points(x, y, view, W, H, ("count", "countv", "sumz", "min", "max", "m2"),
                      val=v)
    np.testing.assert_array_equal(ds["cnt"].values, r["count"])
    np.testing.assert_array_equal(ds["cntv"].values, r["countv"])
    np.testing.assert_array_equal(ds["any"].values, r["count"] > 0)
    G.assert_same(ds["mn"].values, np.where(r["countv"] > 0, r["min"], np.nan))
    G.assert_same(ds["mx"].values, np.where(r["countv"] > 0, r["max"], np.nan))
    with np.errstate(divide="ignore", invalid="ignore"):
        s_exp = np.where(r["countv"] > 0, r["sumz"], np.nan)
        m_exp = np.where(r["countv"] > 0, r["sumz"] / r["countv"], np.nan)
        v_exp = np.where(r["countv"] > 0, r["m2"] / r["countv"], np.nan)
    tol = 1e-6 if dtype == "f4" else 1e-12
    assert _rel_err(ds["s"].values, s_exp) <= tol
    assert _rel_err(ds["mean"].values, m_exp) <= tol
    vt = 1e-5 if dtype == "f4" else 1e-9
    big = r["countv"] >= 2
    ve = _rel_err(np.where(big, ds["var"].values, np.nan), np.where(big, v_exp, np.nan))
    assert ve <= vt, ve
    se = _rel_err(np.where(big, ds["std"].values, np.nan), np.where(big, np.sqrt(v_exp), np.nan))
    assert se <= vt, se


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_canvas_beyond_the_tiles_vs_oracle(dtype):
    """5000 x 4000 = 20M cells of count / countv / sum / min / max / m2 planes: more than
    4096 LDS tiles hold, so the call takes the global-atomic pass; same results."""
    W, H = 5000, 4000
    rng = np.random.default_rng(21)
    n = 2_000_003
    x = rng.normal(0.0, 2.0, n).astype(dtype)
    y = rng.normal(0.0, 2.0, n).astype(dtype)
    v = rng.lognormal(0.0, 1.0, n).astype(dtype)
    v[rng.choice(n, n // 100, replace=False)] = np.nan
    df = pd.DataFrame({"x": x, "y": y, "v": v})
    xr_, yr_ = (-5.0, 5.0), (-4.0, 4.0)
    agg = dsa.summary(cnt=dsa.count(), cntv=dsa.count("v"), mean=dsa.mean("v"),
                      mn=dsa.min("v"), mx=dsa.max("v"), var=dsa.var("v"))
    ds = dsa.Canvas(W, H, x_range=xr_, y_range=yr_).points(df, "x", "y", agg)
    view = G.view_for(W, H, xr_, yr_)
    r = oracle.points(x, y, view, W, H, ("count", "countv", "sumz", "min", "max", "m2"), val=v)
    np.testing.assert_array_equal(ds["cnt"].values, r["count"])
    np.testing.assert_array_equal(ds["cntv"].values, r["countv"])
    G.assert_same(ds["mn"].values, np.where(r["countv"] > 0, r["min"], np.nan))
    G.assert_same(ds["mx"].values, np.where(r["countv"] > 0, r["max"], np.nan))
    with np.errstate(divide="ignore", invalid="ignore"):
        m_exp = np.where(r["countv"] > 0, r["sumz"] / r["countv"], np.nan)
        v_exp = np.where(r["countv"] > 0, r["m2"] / r["countv"], np.nan)
    assert _rel_err(ds["mean"].values, m_exp) <= (1e-6 if dtype == "f4" else 1e-12)
    big = r["countv"] >= 2
    ve = _rel_err(np.where(big, ds["var"].values, np.nan), np.where(big, v_exp, np.nan))
    assert ve <= (1e-5 if dtype == "f4" else 1e-9), ve


def test_skewed_hot_bins_compensated_sum():
    """90% of the rows in a 0.01-sigma blob: thousands of rows per hot bin.  The float32
    compensated sum stays within 1e-6 of the float64 row-ordered sum."""
    from datashader_amd import synthetic
    n = 4_000_000
    x, y, v = synthetic.points_skewed_np(n, seed=2)
    W, H = 1920, 1080
    xr_ = (float(np.nanmin(x)), float(np.nanmax(x)))
    yr_ = (float(np.nanmin(y)), float(np.nanmax(y)))
    df = pd.DataFrame({"x": x, "y": y, "v": v})
    ds = dsa.Canvas(W, H, x_range=xr_, y_range=yr_).points(
        df, "x", "y", dsa.summary(c=dsa.count(), m=dsa.mean("v"), s=dsa.sum("v")))
    view = G.view_for(W, H, xr_, yr_)
    bins = cf.point_bins(x, y, view, W, H)
    np.testing.assert_array_equal(ds["c"].values, cf.count(bins, W, H))
    assert ds["c"].values.max() > 1000
    assert _rel_err(ds["m"].values, cf.mean(bins, v, W, H)) <= 1e-6
    assert _rel_err(ds["s"].values, cf.sum_nan(bins, v, W, H)) <= 1e-6


@pytest.mark.parametrize("n,offset", [(4, 0), (8, 0), (4100, 0), (1_000_004, 0), (5, 0),
                                      (1_000_003, 0), (1_000_004, 1)])
def test_count_mean_row_counts_and_alignment(n, offset):
    """count + mean over linear axes take 16-byte loads of four rows when the float32
    columns are 16-byte aligned and n is a multiple of 4 (a partial last chunk and the
    prefetch past the end included), else 4-byte loads: both equal the closed forms.
    offset 1 starts the columns one element into a larger tensor (misaligned)."""
    import torch
    rng = np.random.default_rng(n + offset)
    m = n + offset
    hot = rng.random(m) < 0.8
    x = np.where(hot, rng.normal(0.5, 0.01, m), rng.uniform(0, 1, m)).astype(np.float32)
    y = np.where(hot, rng.normal(0.5, 0.01, m), rng.uniform(0, 1, m)).astype(np.float32)
    v = rng.lognormal(0, 1, m).astype(np.float32)
    v[rng.random(m) < 0.03] = np.nan
    cols = {k: torch.from_numpy(a).cuda()[offset:] for k, a in (("x", x), ("y", y), ("v", v))}
    x, y, v = x[offset:], y[offset:], v[offset:]
    W, H = 1920, 1080
    ds = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(
        dsa.DeviceFrame(cols), "x", "y", dsa.summary(c=dsa.count(), m=dsa.mean("v")))
    view = G.view_for(W, H, (0, 1), (0, 1))
    bins = cf.point_bins(x, y, view, W, H)
    np.testing.assert_array_equal(ds["c"].values, cf.count(bins, W, H))
    assert _rel_err(ds["m"].values, cf.mean(bins, v, W, H)) <= 1e-6


def test_by_category_vs_oracle():
    from datashader_amd import synthetic
    n = 2_000_000
    x, y, cat = synthetic.points_ridges_np(n, seed=3)
    cats = [f"c{k}" for k in range(16)]
    df = pd.DataFrame({"x": x, "y": y, "cat": pd.Categorical.from_codes(cat, cats)})
    W = H = 256
    agg = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(
        df, "x", "y", dsa.by("cat", dsa.count()))
    view = G.view_for(W, H, (0, 1), (0, 1))
    exp = oracle.points(x, y, view, W, H, ("count",), cat=cat.astype(np.int64), ncat=16,
                        squeeze=False)["count"]
    np.testing.assert_array_equal(agg.values, exp)
    assert list(agg.coords["cat"].values) == cats
    plain = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(df, "x", "y")
    np.testing.assert_array_equal(agg.values.sum(axis=2), plain.values)


def test_full_size_count_800x600_f64_bitexact():
    """10M float64 rows of the 5-Gaussian mixture into 800x600, bit-exact vs the oracle."""
    from datashader_amd import synthetic
    x, y = synthetic.points_mixture_np(10_000_000, seed=1)
    W, H = 800, 600
    df = pd.DataFrame({"x": x, "y": y})
    agg = dsa.Canvas(W, H, x_range=(-10, 10), y_range=(-10, 10)).points(df, "x", "y",
                                                                        dsa.count())
    view = G.view_for(W, H, (-10, 10), (-10, 10))
    exp = oracle.points(x, y, view, W, H, ("count",))["count"]
    np.testing.assert_array_equal(agg.values, exp)
    assert agg.values.dtype == np.uint32


def test_autorange_on_device_matches_reference_bounds():
    rng = np.random.default_rng(11)
    x = rng.normal(size=100_001).astype(np.float32)
    y = rng.normal(size=100_001)
    x[5] = np.nan
    df = pd.DataFrame({"x": x, "y": y})
    agg = dsa.Canvas(50, 40).points(df, "x", "y")
    assert agg.x_range == G.maybe_expand(oracle.bounds(x))
    assert agg.y_range == G.maybe_expand(oracle.bounds(y))
    # all-NaN and constant columns (glyph.py:42-49)
    df = pd.DataFrame({"x": np.full(10, np.nan), "y": np.full(10, 3.0)})
    agg = dsa.Canvas(4, 4).points(df, "x", "y")
    assert agg.x_range == (-1.0, 1.0) and agg.y_range == (2.0, 4.0)
    assert agg.values.sum() == 0


@pytest.mark.parametrize("n", [4_194_305, 12_000_007])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bounds_of_long_columns(n, dtype):
    """dsx_bounds over columns long enough for the eight-loads-per-thread loop and its
    tail: NaN-skipping min / max equal numpy's nanmin / nanmax, with the extremes planted
    at the start, the middle and the last row."""
    from datashader_amd import _device as dev
    g = torch.Generator(device="cuda").manual_seed(n)
    col = torch.randn(n, generator=g, device="cuda", dtype=dtype)
    col[torch.randint(0, n, (1000,), generator=g, device="cuda")] = float("nan")
    for lo_at, hi_at in ((0, n - 1), (n // 2, 3), (n - 1, n // 2 + 1)):
        c = col.clone()
        c[lo_at] = -50.0
        c[hi_at] = 60.0
        host = c.cpu().numpy()
        assert dev.bounds(c) == (float(np.nanmin(host)), float(np.nanmax(host)))


def test_empty_and_tiny_inputs():
    df = pd.DataFrame({"x": np.zeros(0), "y": np.zeros(0), "v": np.zeros(0)})
    ds = dsa.Canvas(3, 2, x_range=(0, 1), y_range=(0, 1)).points(
        df, "x", "y", dsa.summary(c=dsa.count(), m=dsa.mean("v"), mx=dsa.max("v")))
    assert ds["c"].values.sum() == 0
    assert np.isnan(ds["m"].values).all() and np.isnan(ds["mx"].values).all()
    df = pd.DataFrame({"x": [0.5], "y": [0.5], "v": [np.inf]})
    ds = dsa.Canvas(1, 1, x_range=(0, 1), y_range=(0, 1)).points(
        df, "x", "y", dsa.summary(s=dsa.sum("v"), mn=dsa.min("v")))
    assert ds["s"].values[0, 0] == np.inf and ds["mn"].values[0, 0] == np.inf


def test_device_frame_result_stays_on_device():
    x = torch.tensor([0.1, 0.9, 0.9], dtype=torch.float32, device="cuda")
    y = torch.tensor([0.1, 0.9, 0.9], dtype=torch.float32, device="cuda")
    f = dsa.DeviceFrame({"x": x, "y": y})
    agg = dsa.Canvas(2, 2, x_range=(0, 1), y_range=(0, 1)).points(f, "x", "y")
    assert agg.data.is_cuda and agg.data.dtype == torch.uint32
    np.testing.assert_array_equal(agg.values, [[1, 0], [0, 2]])


def test_window_cache_reuse_and_stale_hint():
    """The hot window only steers performance: a resident frame reuses the window picked
    on its first call, and a stale or nonsensical stored window (sanitised to empty by
    the kernels) still gives the same aggregate."""
    from datashader_amd import synthetic
    x, y, v = synthetic.points_skewed_np(500_000, seed=9)
    f = dsa.DeviceFrame({k: torch.from_numpy(a).cuda() for k, a in (("x", x), ("y", y), ("v", v))})
    c = dsa.Canvas(640, 480, x_range=(0, 1), y_range=(0, 1))
    agg = dsa.summary(n=dsa.count(), m=dsa.mean("v"), hi=dsa.max("v"))
    first = c.points(f, "x", "y", agg)
    (slot,) = f._window_cache.values()
    rec = slot.rec
    assert int(rec[2]) > 0 and int(rec[3]) > 0   # a window was picked and stored
    assert int(rec[4]) == 1                      # ... with the window strategy
    second = c.points(f, "x", "y", agg)
    for bogus in ([-5, 10**6, 7, 9, 1], [600, 400, 100, 100, 2], [0, 0, 0, 0, 99],
                  [10, 10, 50, 50, 2]):
        rec.copy_(torch.tensor(bogus + [0, 0, 0], dtype=torch.int32))
        slot.strategy = slot.event = None  # the stored record's own strategy word runs
        third = c.points(f, "x", "y", agg)
        for k in ("n", "hi"):
            np.testing.assert_array_equal(third[k].values, first[k].values)
        np.testing.assert_allclose(third["m"].values, first["m"].values, rtol=1e-12)
    for k in ("n", "m", "hi"):
        np.testing.assert_array_equal(second[k].values, first[k].values)


def test_by_category_1024x1024x16_vs_oracle():
    """by('cat', count()) at the categorical workload's own shape: 1024 x 1024 x 16
    cells (64 MB of planes, about a thousand LDS tiles) from 4M ridge rows with Zipf
    categories, ranges = the data extent; bit-exact vs the oracle."""
    from datashader_amd import synthetic
    n = 4_000_000
    x, y, cat = synthetic.points_ridges_np(n, seed=3)
    cats = [f"c{k}" for k in range(16)]
    frame = dsa.DeviceFrame({"x": torch.as_tensor(x, device="cuda"),
                             "y": torch.as_tensor(y, device="cuda"),
                             "cat": torch.as_tensor(cat, device="cuda")},
                            categories={"cat": cats})
    W = H = 1024
    agg = dsa.Canvas(W, H).points(frame, "x", "y", dsa.by("cat", dsa.count()))
    xr_, yr_ = agg.attrs["x_range"], agg.attrs["y_range"]
    assert xr_ == G.maybe_expand(oracle.bounds(x)) and yr_ == G.maybe_expand(oracle.bounds(y))
    view = G.view_for(W, H, xr_, yr_)
    exp = oracle.points(x, y, view, W, H, ("count",), cat=cat.astype(np.int64), ncat=16,
                        squeeze=False)["count"]
    got = agg.values
    assert got.shape == (H, W, 16)
    np.testing.assert_array_equal(got, exp)
    assert int(exp.sum()) == n and (exp > 0).sum() > 1_000_000


def test_skewed_min_max_std_vs_oracle():
    """The float-reduction workload's shape (1920 x 1080, 90% of the rows in a 0.01-sigma
    blob, lognormal float32 values) at 4M rows: min and max bit-exact, var / std within
    1e-5 relative of the two-pass float64 closed form, sum / mean within 1e-6."""
    from datashader_amd import synthetic
    n = 4_000_000
    x, y, v = synthetic.points_skewed_np(n, seed=2)
    W, H = 1920, 1080
    xr_ = (float(np.nanmin(x)), float(np.nanmax(x)))
    yr_ = (float(np.nanmin(y)), float(np.nanmax(y)))
    frame = dsa.DeviceFrame({k: torch.as_tensor(a, device="cuda")
                             for k, a in (("x", x), ("y", y), ("v", v))})
    ds = dsa.Canvas(W, H, x_range=xr_, y_range=yr_).points(
        frame, "x", "y", dsa.summary(c=dsa.count(), s=dsa.sum("v"), m=dsa.mean("v"),
                                     lo=dsa.min("v"), hi=dsa.max("v"), sd=dsa.std("v"),
                                     var=dsa.var("v")))
    view = G.view_for(W, H, xr_, yr_)
    bins = cf.point_bins(x, y, view, W, H)
    np.testing.assert_array_equal(ds["c"].values, cf.count(bins, W, H))
    mn, mx = cf.minmax(bins, v, W, H)
    G.assert_same(ds["lo"].values, mn)
    G.assert_same(ds["hi"].values, mx)
    assert _rel_err(ds["s"].values, cf.sum_nan(bins, v, W, H)) <= 1e-6
    assert _rel_err(ds["m"].values, cf.mean(bins, v, W, H)) <= 1e-6
    var = cf.var2pass(bins, v, W, H)
    assert _rel_err(ds["var"].values, var) <= 1e-5
    assert _rel_err(ds["sd"].values, np.sqrt(var)) <= 1e-5
    # and the reference's own sequential m2 (reductions.py:1128-1137) via the oracle
    r = oracle.points(x, y, view, W, H, ("countv", "m2"), val=v)
    with np.errstate(divide="ignore", invalid="ignore"):
        seq = np.where(r["countv"] > 0, np.sqrt(r["m2"] / r["countv"]), np.nan)
    assert _rel_err(ds["sd"].values, seq) <= 1e-5
    assert ds["c"].values.max() > 1000

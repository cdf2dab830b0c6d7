"""GPU parity of the row-index reductions (first, last, where, _min_row_index,
_max_row_index) against the oracle, bit-exact: the selected row of every cell, and the
looked-up value at it.  Random skewed points (hot cells hold thousands of rows), values
with ties and NaNs, float32 and float64 coordinates and values, plain and by-category."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import datashader_amd as dsa
from oracle import closed_forms as cf
from oracle import rowindex as ri
import golden_utils as G

pytestmark = pytest.mark.gpu


def _frame(n, xdt, vdt, seed):
    rng = np.random.default_rng(seed)
    hot = rng.random(n) < 0.7
    x = np.where(hot, rng.normal(0.5, 0.02, n), rng.uniform(0, 1, n)).astype(xdt)
    y = np.where(hot, rng.normal(0.4, 0.02, n), rng.uniform(0, 1, n)).astype(xdt)
    v = rng.integers(0, 50, n).astype(vdt)  # many ties per cell
    v[rng.random(n) < 0.05] = np.nan
    w = rng.normal(size=n)
    cat = pd.Categorical(rng.choice(list("pqrs"), n), categories=list("pqrs"))
    return pd.DataFrame({"x": x, "y": y, "v": v, "w": w, "cat": cat})


MODES = {"first": "first", "last": "last", "max": "max", "min": "min"}


@pytest.mark.parametrize("by", [False, True], ids=["plain", "by-cat"])
@pytest.mark.parametrize("vdt", ["f4", "f8"])
@pytest.mark.parametrize("xdt", ["f4", "f8"])
def test_row_index_reductions_vs_oracle(xdt, vdt, by):
    df = _frame(300_000, xdt, vdt, 3)
    W, H = 123, 77
    cvs = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1))
    src = dsa.DeviceFrame.from_pandas(df)
    view = G.view_for(W, H, (0, 1), (0, 1))
    bins = cf.point_bins(df["x"].to_numpy(np.float64), df["y"].to_numpy(np.float64), view, W, H)
    ncat = 4 if by else 1
    cell = ri.cells(bins, df["cat"].cat.codes.to_numpy() if by else None, ncat)
    v, w = df["v"].to_numpy(np.float64), df["w"].to_numpy()
    shape = (H, W, ncat) if by else (H, W)

    def wrap(r):
        return dsa.by("cat", r) if by else r

    aggs = {}
    for name, sel in (("first", dsa.first("v")), ("last", dsa.last("v")),
                      ("max", dsa.max("v")), ("min", dsa.min("v"))):
        aggs[f"{name}_w"] = wrap(dsa.where(sel, "w"))
        aggs[f"{name}_i"] = wrap(dsa.where(sel))
    aggs["first"] = wrap(dsa.first("v"))
    aggs["last"] = wrap(dsa.last("v"))
    aggs["rmin"] = wrap(dsa._min_row_index())
    aggs["rmax"] = wrap(dsa._max_row_index())
    aggs["count"] = wrap(dsa.count())
    got = cvs.points(src, "x", "y", dsa.summary(**aggs))
    for name in MODES:
        idx = ri.selected_rows_closed_form(cell, W * H * ncat, MODES[name], v).reshape(shape)
        np.testing.assert_array_equal(got[f"{name}_i"].values, idx, err_msg=name)
        assert got[f"{name}_i"].values.dtype == np.int64
        G.assert_same(got[f"{name}_w"].values, ri.take(idx, w))
    for name in ("first", "last"):
        idx = ri.selected_rows_closed_form(cell, W * H * ncat, name, v).reshape(shape)
        G.assert_same(got[name].values, ri.take(idx, v))
    for name, mode in (("rmin", "first"), ("rmax", "last")):
        idx = ri.selected_rows_closed_form(cell, W * H * ncat, mode).reshape(shape)
        np.testing.assert_array_equal(got[name].values, idx)
    assert int(got["count"].values.sum()) == int((cell >= 0).sum())


def test_where_validation():
    with pytest.raises(TypeError, match="selector can only be"):
        dsa.where(dsa.sum("v"), "w")
    df = _frame(100, "f8", "f8", 1)
    with pytest.raises(ValueError, match="same column"):
        dsa.Canvas(4, 4).points(df, "x", "y", dsa.where(dsa.max("v"), "v"))


def test_row_index_lines_not_supported():
    df = pd.DataFrame({"x": np.arange(5.0), "y": np.arange(5.0), "v": np.arange(5.0)})
    with pytest.raises(NotImplementedError):
        dsa.Canvas(4, 4).line(df, "x", "y", agg=dsa.first("v"))


N_MODES = {"max_n": "max", "min_n": "min", "first_n": "first", "last_n": "last"}


# n <= 8 takes the n-pass form (float32 values for max_n / min_n), n = 12 the sort
@pytest.mark.parametrize("n", [1, 4, 12])
@pytest.mark.parametrize("by", [False, True], ids=["plain", "by-cat"])
@pytest.mark.parametrize("vdt", ["f4", "f8"])
def test_n_reductions_vs_oracle(vdt, by, n):
    df = _frame(400_000, "f4", vdt, 9)
    W, H = 97, 61
    cvs = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1))
    src = dsa.DeviceFrame.from_pandas(df)
    view = G.view_for(W, H, (0, 1), (0, 1))
    bins = cf.point_bins(df["x"].to_numpy(np.float64), df["y"].to_numpy(np.float64), view, W, H)
    ncat = 4 if by else 1
    cell = ri.cells(bins, df["cat"].cat.codes.to_numpy() if by else None, ncat)
    v, w = df["v"].to_numpy(np.float64), df["w"].to_numpy()
    shape = (H, W, ncat, n) if by else (H, W, n)

    def wrap(r):
        return dsa.by("cat", r) if by else r

    aggs = {}
    for name in N_MODES:
        red = getattr(dsa, name)
        aggs[name] = wrap(red("v", n=n))
        aggs[name + "_i"] = wrap(dsa.where(red("v", n=n)))
        aggs[name + "_w"] = wrap(dsa.where(red("v", n=n), "w"))
    aggs["rmin"] = wrap(dsa._min_n_row_index(n=n))
    aggs["rmax"] = wrap(dsa._max_n_row_index(n=n))
    got = cvs.points(src, "x", "y", dsa.summary(**aggs))
    ncells = W * H * ncat
    for name, mode in N_MODES.items():
        idx = ri.select_n_closed_form(cell, ncells, n, mode, v).reshape(shape)
        np.testing.assert_array_equal(got[name + "_i"].values, idx, err_msg=name)
        G.assert_same(got[name].values, ri.take(idx, v))
        G.assert_same(got[name + "_w"].values, ri.take(idx, w))
        assert got[name].dims[-1] == "n"
    for name, mode in (("rmin", "first"), ("rmax", "last")):
        idx = ri.select_n_closed_form(cell, ncells, n, mode).reshape(shape)
        np.testing.assert_array_equal(got[name].values, idx)
        assert got[name].values.dtype == np.int64


@pytest.mark.parametrize("by", [False, True], ids=["plain", "by-cat"])
@pytest.mark.parametrize("xdt", ["f4", "f8"])
def test_windowed_first_last_vs_global_pass_and_oracle(xdt, by):
    """first / last / the row-index selectors on a canvas much larger than the LDS window
    (1920 x 1080, a hot 0.01-sigma blob plus uniform rows): the LDS-window pass (the
    default, with a workspace) and the global-atomic pass (no workspace) both equal the
    oracle's closed form."""
    import ctypes

    import torch

    from datashader_amd import _abi, backend
    from datashader_amd import _device as dev
    from datashader_amd.glyphs import Point

    rng = np.random.default_rng(17)
    n = 3_000_000
    hot = rng.random(n) < 0.9
    x = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(xdt)
    y = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(xdt)
    v = rng.lognormal(0, 2, n).astype(xdt)
    v[rng.random(n) < 0.02] = np.nan
    cat = pd.Categorical(rng.choice(list("pqrs"), n), categories=list("pqrs"))
    df = pd.DataFrame({"x": x, "y": y, "v": v, "cat": cat})
    W, H = 1920, 1080
    xr = (float(np.nanmin(x)), float(np.nanmax(x)))
    yr = (float(np.nanmin(y)), float(np.nanmax(y)))
    cvs = dsa.Canvas(W, H, x_range=xr, y_range=yr)
    src = dsa.DeviceFrame.from_pandas(df)
    view = G.view_for(W, H, xr, yr)
    bins = cf.point_bins(x.astype(np.float64), y.astype(np.float64), view, W, H)
    ncat = 4 if by else 1
    cell = ri.cells(bins, df["cat"].cat.codes.to_numpy() if by else None, ncat)
    ncells = W * H * ncat
    shape = (H, W, ncat) if by else (H, W)

    def wrap(r):
        return dsa.by("cat", r) if by else r

    got = cvs.points(src, "x", "y", dsa.summary(
        f=wrap(dsa.first("v")), l=wrap(dsa.last("v")), rmin=wrap(dsa._min_row_index()),
        rmax=wrap(dsa._max_row_index())))
    vv = v.astype(np.float64)
    want_f = ri.selected_rows_closed_form(cell, ncells, "first", vv).reshape(shape)
    want_l = ri.selected_rows_closed_form(cell, ncells, "last", vv).reshape(shape)
    G.assert_same(got["f"].values, ri.take(want_f, vv))
    G.assert_same(got["l"].values, ri.take(want_l, vv))
    np.testing.assert_array_equal(
        got["rmin"].values, ri.selected_rows_closed_form(cell, ncells, "first").reshape(shape))
    np.testing.assert_array_equal(
        got["rmax"].values, ri.selected_rows_closed_form(cell, ncells, "last").reshape(shape))
    if by:
        return
    # the same selections through the C ABI with and without the workspace
    lib = _abi.load()
    p = backend.Pass("v", None)
    catc, _, _ = backend._categorizer(None, src, None)
    mview = backend.make_view(cvs, xr, yr)
    d, keep = backend._pass_desc(src, None, Point("x", "y"), mview, p, src.device, catc)
    vw = backend._view_struct(mview)
    for mode, want in ((_abi.RI_FIRST, want_f), (_abi.RI_LAST, want_l)):
        outs = []
        for windowed in (True, False):
            r = _abi.RowIndexDesc()
            r.mode, r.nan_check = mode, 1
            idx = torch.empty(ncells, dtype=torch.int64, device="cuda")
            r.index = dev.ptr(idx)
            ws = int(lib.dsx_points_rowindex_workspace_size(d, vw, r)) if windowed else 0
            work = torch.empty(max(ws, 1), dtype=torch.uint8, device="cuda")
            print(f"rowindex mode {mode} windowed {windowed} ws {ws}", flush=True)
            _abi.check(lib.dsx_points_rowindex(d, vw, r, dev.ptr(work) if windowed else None, ws,
                                               dev.stream_handle()), "dsx_points_rowindex")
            outs.append(idx.cpu().numpy().reshape(shape))
        np.testing.assert_array_equal(outs[0], want)
        np.testing.assert_array_equal(outs[1], want)


@pytest.mark.parametrize("by", [False, True], ids=["plain", "by-cat"])
def test_packed_where_max_min_vs_oracle(by):
    """where(max / min) over a float32 value column on a canvas much larger than the LDS
    window (1920 x 1080, a hot 0.01-sigma blob plus uniform rows): one packed pass (64-bit
    cells key(v) << 32 | row, or | ~row for max) keeps the extreme value's first row, in
    the window and through the global atomics.  Integer values give many ties per cell;
    NaN rows are skipped; bit-exact vs the oracle's closed form, also through the C ABI
    (no key plane, with a workspace) and with an empty input."""
    import torch

    from datashader_amd import _abi, backend
    from datashader_amd import _device as dev
    from datashader_amd.glyphs import Point

    rng = np.random.default_rng(23)
    n = 3_000_000
    hot = rng.random(n) < 0.9
    x = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(np.float32)
    y = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(np.float32)
    v = rng.integers(-20, 20, n).astype(np.float32)
    v[rng.random(n) < 0.02] = np.nan
    w = rng.normal(size=n)
    cat = pd.Categorical(rng.choice(list("pqrs"), n), categories=list("pqrs"))
    df = pd.DataFrame({"x": x, "y": y, "v": v, "w": w, "cat": cat})
    W, H = 1920, 1080
    cvs = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1))
    src = dsa.DeviceFrame.from_pandas(df)
    view = G.view_for(W, H, (0, 1), (0, 1))
    bins = cf.point_bins(x.astype(np.float64), y.astype(np.float64), view, W, H)
    ncat = 4 if by else 1
    cell = ri.cells(bins, df["cat"].cat.codes.to_numpy() if by else None, ncat)
    ncells = W * H * ncat
    shape = (H, W, ncat) if by else (H, W)

    def wrap(r):
        return dsa.by("cat", r) if by else r

    got = cvs.points(src, "x", "y", dsa.summary(
        mx=wrap(dsa.where(dsa.max("v"))), mn=wrap(dsa.where(dsa.min("v"), "w"))))
    vv = v.astype(np.float64)
    want_mx = ri.selected_rows_closed_form(cell, ncells, "max", vv).reshape(shape)
    want_mn = ri.selected_rows_closed_form(cell, ncells, "min", vv).reshape(shape)
    np.testing.assert_array_equal(got["mx"].values, want_mx)
    G.assert_same(got["mn"].values, ri.take(want_mn, w))
    assert (want_mx >= 0).sum() > 10_000  # cells outside the window are covered
    if by:
        return
    lib = _abi.load()
    p = backend.Pass("v", None)
    catc, _, _ = backend._categorizer(None, src, None)
    mview = backend.make_view(cvs, (0, 1), (0, 1))
    d, keep = backend._pass_desc(src, None, Point("x", "y"), mview, p, src.device, catc)
    vw = backend._view_struct(mview)
    for mode, want in ((_abi.RI_ARGMAX, want_mx), (_abi.RI_ARGMIN, want_mn)):
        for rows in (n, 0):
            d.n = rows
            r = _abi.RowIndexDesc()
            r.mode, r.nan_check = mode, 1
            idx = torch.empty(ncells, dtype=torch.int64, device="cuda")
            r.index = dev.ptr(idx)
            ws = int(lib.dsx_points_rowindex_workspace_size(d, vw, r))
            assert ws > 0
            work = torch.empty(ws, dtype=torch.uint8, device="cuda")
            print(f"packed rowindex mode {mode} rows {rows} ws {ws}", flush=True)
            _abi.check(lib.dsx_points_rowindex(d, vw, r, dev.ptr(work), ws, dev.stream_handle()),
                       "dsx_points_rowindex")
            out = idx.cpu().numpy().reshape(shape)
            np.testing.assert_array_equal(out, want if rows else np.full(shape, -1))
        # no key and no workspace: the key plane is required
        r = _abi.RowIndexDesc()
        r.mode, r.nan_check = mode, 1
        r.index = dev.ptr(idx)
        assert lib.dsx_points_rowindex(d, vw, r, None, 0, dev.stream_handle()) == -1  # DSX_ERR_NULL
    d.n = n


@pytest.mark.parametrize("by", [False, True], ids=["plain", "by-cat"])
def test_n_pass_on_a_large_canvas_vs_oracle(by):
    """max_n / min_n / first_n / last_n (n = 3, float32 values) on a canvas much larger
    than the LDS window (1920 x 1080, a hot 0.01-sigma blob plus uniform rows): the
    n-pass form keeps slot k of every cell in pass k, in the window and through the
    global atomics with the previous slot's key as the bound.  Integer values give many
    ties per cell; NaN rows are skipped; bit-exact vs the oracle's closed form."""
    rng = np.random.default_rng(31)
    n = 3_000_000
    hot = rng.random(n) < 0.9
    x = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(np.float32)
    y = np.where(hot, rng.normal(0.5, 0.01, n), rng.uniform(0, 1, n)).astype(np.float32)
    v = rng.integers(1, 40, n).astype(np.float32)
    v[rng.random(n) < 0.02] = np.nan
    cat = pd.Categorical(rng.choice(list("pqrs"), n), categories=list("pqrs"))
    df = pd.DataFrame({"x": x, "y": y, "v": v, "cat": cat})
    W, H = 1920, 1080
    cvs = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1))
    src = dsa.DeviceFrame.from_pandas(df)
    view = G.view_for(W, H, (0, 1), (0, 1))
    bins = cf.point_bins(x.astype(np.float64), y.astype(np.float64), view, W, H)
    ncat = 4 if by else 1
    cell = ri.cells(bins, df["cat"].cat.codes.to_numpy() if by else None, ncat)
    ncells = W * H * ncat
    k = 3
    shape = (H, W, ncat, k) if by else (H, W, k)

    def wrap(r):
        return dsa.by("cat", r) if by else r

    got = cvs.points(src, "x", "y", dsa.summary(
        mx=wrap(dsa.where(dsa.max_n("v", n=k))), mn=wrap(dsa.where(dsa.min_n("v", n=k))),
        fi=wrap(dsa.where(dsa.first_n("v", n=k))), la=wrap(dsa._max_n_row_index(n=k))))
    vv = v.astype(np.float64)
    for name, mode, val in (("mx", "max", vv), ("mn", "min", vv), ("fi", "first", vv),
                            ("la", "last", None)):
        want = ri.select_n_closed_form(cell, ncells, k, mode, val).reshape(shape)
        np.testing.assert_array_equal(got[name].values, want, err_msg=name)
    # second slots of cells outside the window (the blob holds ~7.5k cells)
    assert (got["mx"].values[..., 1] >= 0).sum() > 20_000

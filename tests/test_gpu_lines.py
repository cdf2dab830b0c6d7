"""GPU parity of the line path: libdsx's Bresenham rasteriser (through the C ABI and
through Canvas.line) vs the reference's expectations, the touched-pixel masks of the
reference's line images, and the CPU oracle.  Every line result is integer: bit-exact."""
from __future__ import annotations

import ctypes

import numpy as np
import pandas as pd
import pytest

import oracle

import golden_utils as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import datashader_amd as dsa  # noqa: E402
from datashader_amd import _abi, _device  # noqa: E402

GC = G.cases()


def abi_lines(xs, ys, view, W, H, axis=0, plot_start=True, x_ls=0, x_vs=1, y_ls=0, y_vs=1,
              nlines=1, nverts=None, xdtype=np.float64, ydtype=np.float64):
    """One dsx_lines_agg call (count plane) on explicit device buffers."""
    lib = _abi.load()
    xt = torch.as_tensor(np.ascontiguousarray(xs, xdtype), device="cuda")
    yt = torch.as_tensor(np.ascontiguousarray(ys, ydtype), device="cuda")
    cnt = torch.zeros(H * W, dtype=torch.int32, device="cuda")
    d = _abi.LinesDesc()
    d.nlines, d.nverts = nlines, (len(xs) if nverts is None else nverts)
    d.x, d.y = _device.ptr(xt), _device.ptr(yt)
    d.x_dtype = _abi.F32 if xdtype == np.float32 else _abi.F64
    d.y_dtype = _abi.F32 if ydtype == np.float32 else _abi.F64
    d.x_line_stride, d.x_vert_stride, d.y_line_stride, d.y_vert_stride = x_ls, x_vs, y_ls, y_vs
    d.axis, d.plot_start = axis, int(plot_start)
    v = _abi.View()
    v.sx, v.tx, v.sy, v.ty = view.sx, view.tx, view.sy, view.ty
    v.xmin, v.xmax, v.ymin, v.ymax = view.xmin, view.xmax, view.ymin, view.ymax
    v.xxmax, v.yymax, v.width, v.height = view.xxmax, view.yymax, W, H
    acc = _abi.Planes()
    acc.count = _device.ptr(cnt)
    _abi.check(lib.dsx_lines_agg(d, v, acc, _device.stream_handle()), "dsx_lines_agg")
    return cnt.view(torch.uint32).cpu().numpy().reshape(H, W)


@pytest.mark.parametrize("case", GC["glyph_lines"], ids=[c["name"] for c in GC["glyph_lines"]])
def test_glyph_lines_golden_abi(case):
    vv = case["view"]
    view = oracle.View(vv["sx"], vv["tx"], vv["sy"], vv["ty"], vv["xmin"], vv["xmax"],
                       vv["ymin"], vv["ymax"])
    W, H = case["W"], case["H"]
    if case["kind"] == "segment":
        (x0, y0), (x1, y1) = case["p0"], case["p1"]
        agg = abi_lines([x0, x1], [y0, y1], view, W, H, plot_start=case["segment_start"])
    else:
        agg = abi_lines(G.dec(case["xs"]), G.dec(case["ys"]), view, W, H,
                        plot_start=case["plot_start"])
    if case["expected"] is not None:
        np.testing.assert_array_equal(agg, np.array(case["expected"]))
    if case.get("total") is not None:
        assert agg.sum() == case["total"]
    if case.get("at"):
        r, c_, val = case["at"]
        assert agg[r, c_] == val


def _line_kwargs(kw):
    kw = dict(kw)
    spec = kw.pop("agg")
    col = spec.get("column")
    agg = getattr(dsa, spec["kind"])() if col is None else getattr(dsa, spec["kind"])(col)
    for k in ("x", "y"):
        if isinstance(kw[k], dict):
            kw[k] = np.asarray(kw[k]["ndarray"])
    return kw, agg


@pytest.mark.parametrize("source", ["pandas", "device"])
@pytest.mark.parametrize("case", GC["lines"], ids=[c["name"] for c in GC["lines"]])
def test_lines_golden(case, source):
    df = G.line_frame(case["frame"])
    src = df if source == "pandas" else dsa.DeviceFrame.from_pandas(df)
    kw, agg_ = _line_kwargs(case["kwargs"])
    cv = case["canvas"]
    agg = dsa.Canvas(cv["W"], cv["H"], x_range=cv["x_range"], y_range=cv["y_range"]).line(
        src, agg=agg_, **kw)
    got = agg.values
    if case.get("ones_at") is not None:
        assert np.argwhere(got == 1).tolist() == case["ones_at"]
        assert got.sum() == len(case["ones_at"])
    else:
        G.assert_same(got, G.expected_array(case))
    if "x_range" in case:
        assert np.allclose(agg.attrs["x_range"], case["x_range"])
        assert np.allclose(agg.attrs["y_range"], case["y_range"])


@pytest.mark.parametrize("case", GC["noaa_masks"], ids=[c["name"] for c in GC["noaa_masks"]])
def test_noaa_masks(case):
    """The reference's line images (test_antialias.py:302-370, max('val') lines): the
    touched pixels of each Canvas.line call, unioned, equal the image's non-white mask."""
    cv = case["canvas"]
    c = dsa.Canvas(cv["W"], cv["H"], x_range=cv["x_range"], y_range=cv["y_range"])
    touched = np.zeros((cv["H"], cv["W"]), bool)
    for pl in case["polylines"]:
        pl = np.asarray(pl, float)
        df = pd.DataFrame({"x": pl[:, 0], "y": pl[:, 1], "val": 5.0})
        agg = c.line(df, "x", "y", agg=dsa.max("val"))
        touched |= ~np.isnan(agg.values)
    np.testing.assert_array_equal(touched, np.array(case["expected"], bool))


# ---- random lines vs the oracle --------------------------------------------------------

@pytest.mark.parametrize("ydtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout", ["pandas", "row_major", "block"])
def test_random_xconst_lines_vs_oracle(ydtype, layout):
    """Random walks with NaN gaps, partly outside the view (clipped), x-constant axis 1."""
    rng = np.random.default_rng(5)
    nlines, nverts, W, H = 3000, 64, 300, 200
    xs = np.linspace(-0.2, 1.1, nverts)
    ys = (rng.uniform(-1.2, 1.2, (nlines, 1)) +
          np.cumsum(rng.normal(0, 0.08, (nlines, nverts)), axis=1)).astype(ydtype)
    ys[rng.random(ys.shape) < 0.01] = np.nan
    names = [f"y{j}" for j in range(nverts)]
    df = pd.DataFrame(ys, columns=names)
    if layout == "block":
        src = dsa.DeviceFrame.from_pandas(df)
    elif layout == "row_major":  # columns are strided views of one [nlines, nverts] tensor
        t = torch.as_tensor(ys, device="cuda")
        src = dsa.DeviceFrame({n: t[:, j] for j, n in enumerate(names)})
    else:
        src = df
    xr_, yr_ = (0.0, 1.0), (-1.0, 1.0)
    agg = dsa.Canvas(W, H, x_range=xr_, y_range=yr_).line(src, x=xs, y=names,
                                                         agg=dsa.count(), axis=1)
    view = G.view_for(W, H, xr_, yr_, yf32=(ydtype == np.float32))
    exp = oracle.lines_xconst(xs, ys.astype(np.float64), view, W, H)["count"]
    np.testing.assert_array_equal(agg.values, exp)
    assert exp.sum() > 0


def test_random_axis0_and_axis1_vs_oracle():
    rng = np.random.default_rng(6)
    n, W, H = 20000, 128, 96
    x = np.cumsum(rng.normal(0, 0.05, n))
    y = np.cumsum(rng.normal(0, 0.05, n))
    x[rng.random(n) < 0.002] = np.nan
    df = pd.DataFrame({"x": x, "y": y, "v": rng.uniform(0, 1, n)})
    xr_, yr_ = (float(np.nanpercentile(x, 10)), float(np.nanpercentile(x, 90))), \
        (float(np.nanpercentile(y, 10)), float(np.nanpercentile(y, 90)))
    c = dsa.Canvas(W, H, x_range=xr_, y_range=yr_)
    view = G.view_for(W, H, xr_, yr_)
    agg = c.line(df, "x", "y", agg=dsa.count())
    exp = oracle.line_axis0(x, y, view, W, H)["count"]
    np.testing.assert_array_equal(agg.values, exp)
    # value reductions along a line take the value of the segment's first row
    agg = c.line(df, "x", "y", agg=dsa.max("v"))
    expm = oracle.line_axis0(x, y, view, W, H, ("max",), val=df["v"].to_numpy())["max"]
    G.assert_same(agg.values, expm)
    # axis 1: each row of a 2-D block is one line
    X = rng.uniform(xr_[0] - 1, xr_[1] + 1, (500, 7))
    Y = rng.uniform(yr_[0] - 1, yr_[1] + 1, (500, 7))
    cols = {f"x{j}": X[:, j] for j in range(7)} | {f"y{j}": Y[:, j] for j in range(7)}
    df1 = pd.DataFrame(cols)
    agg = c.line(df1, x=[f"x{j}" for j in range(7)], y=[f"y{j}" for j in range(7)],
                 agg=dsa.count(), axis=1)
    np.testing.assert_array_equal(agg.values, oracle.lines_axis1(X, Y, view, W, H)["count"])


def test_random_walks_2048x1024_vs_oracle():
    """The 2048x1024 random-walk workload at 20k lines, bit-exact vs the oracle."""
    from datashader_amd import synthetic
    xs, ys = synthetic.random_walks_np(20000, 256, seed=4)
    names = [f"y{j}" for j in range(256)]
    f = dsa.DeviceFrame.from_pandas(pd.DataFrame(ys, columns=names))
    agg = dsa.Canvas(2048, 1024).line(f, x=xs, y=names, agg=dsa.count(), axis=1)
    xr_, yr_ = agg.x_range, agg.y_range
    assert yr_ == G.maybe_expand(oracle.bounds(ys))
    view = G.view_for(2048, 1024, xr_, yr_, yf32=True)
    exp = oracle.lines_xconst(xs, ys.astype(np.float64), view, 2048, 1024)["count"]
    np.testing.assert_array_equal(agg.values, exp)


@pytest.mark.parametrize("case", ["linspace", "shuffled_x", "zoomed", "tall", "nan_x", "few_lines"])
def test_xconst_strips_vs_oracle_and_general_path(case):
    """x-constant lines take the LDS-strip kernel; pixels outside a strip's band (x not
    monotone, zoomed-in views, wraparound) take its global fallback.  The same lines
    given with one x row per line take the per-segment kernel.  Both match the oracle."""
    rng = np.random.default_rng(11)
    nlines, nverts, W, H = 2000, 40, 256, 1000
    xs = np.linspace(0, 1, nverts)
    xr_, yr_ = (0.0, 1.0), (-1.0, 1.0)
    if case == "shuffled_x":
        rng.shuffle(xs)
    elif case == "zoomed":
        xr_ = (0.3, 0.4)
    elif case == "tall":
        H = 3000
    elif case == "nan_x":
        xs[[0, 17, 18]] = np.nan
    elif case == "few_lines":
        nlines = 3
    ys = rng.uniform(-1.2, 1.2, (nlines, 1)) + np.cumsum(rng.normal(0, 0.1, (nlines, nverts)),
                                                         axis=1)
    ys[rng.random(ys.shape) < 0.01] = np.nan
    view = G.view_for(W, H, xr_, yr_)
    exp = oracle.lines_xconst(xs, ys, view, W, H)["count"]
    assert exp.sum() > 0
    # one shared x row, y as a [nverts, nlines] block (column j = vertex j of every line)
    strip = abi_lines(xs, ys.T, view, W, H, axis=1, x_ls=0, x_vs=1, y_ls=1, y_vs=nlines,
                      nlines=nlines, nverts=nverts)
    np.testing.assert_array_equal(strip, exp)
    # the same lines with an x row per line
    general = abi_lines(np.tile(xs, (nlines, 1)), ys, view, W, H, axis=1, x_ls=nverts, x_vs=1,
                        y_ls=nverts, y_vs=1, nlines=nlines, nverts=nverts)
    np.testing.assert_array_equal(general, exp)


def test_line_value_reductions_vs_oracle():
    """sum / mean / min / max / var / std along lines (axis 0): each pixel of a segment takes
    the value of the segment's first row; var / std are two-pass on the GPU and the
    reference's sequential m2 in the oracle (float64 values: relative 1e-9)."""
    rng = np.random.default_rng(9)
    n, W, H = 30000, 96, 64
    x = np.cumsum(rng.normal(0, 0.05, n))
    y = np.cumsum(rng.normal(0, 0.05, n))
    v = rng.normal(3.0, 2.0, n)
    v[rng.random(n) < 0.01] = np.nan
    df = pd.DataFrame({"x": x, "y": y, "v": v})
    xr_ = (float(np.percentile(x, 5)), float(np.percentile(x, 95)))
    yr_ = (float(np.percentile(y, 5)), float(np.percentile(y, 95)))
    c = dsa.Canvas(W, H, x_range=xr_, y_range=yr_)
    view = G.view_for(W, H, xr_, yr_)
    r = oracle.line_axis0(x, y, view, W, H, ("countv", "sumz", "min", "max", "m2"), val=v)
    agg = c.line(df, "x", "y", agg=dsa.summary(s=dsa.sum("v"), m=dsa.mean("v"), mn=dsa.min("v"),
                                              mx=dsa.max("v"), var=dsa.var("v"),
                                              std=dsa.std("v"), n=dsa.count("v")))
    cv = r["countv"]
    np.testing.assert_array_equal(agg["n"].values, cv)
    has = cv > 0
    for name, exp, tol in (("s", r["sumz"], 1e-12), ("m", r["sumz"] / np.maximum(cv, 1), 1e-12),
                           ("var", r["m2"] / np.maximum(cv, 1), 1e-9),
                           ("std", np.sqrt(r["m2"] / np.maximum(cv, 1)), 1e-9)):
        got = agg[name].values
        assert np.all(np.isnan(got[~has]))
        np.testing.assert_allclose(got[has], exp[has], rtol=tol, atol=1e-9)
    G.assert_same(agg["mn"].values, r["min"])
    G.assert_same(agg["mx"].values, r["max"])
    assert has.sum() > 100


@pytest.mark.parametrize("dt", ["f4", "f8"])
def test_ragged_lines_vs_oracle(dt):
    """Ragged rows of random lengths (x and y lengths differ per row, NaN gaps inside):
    count and max against the oracle drawing each row's first min(len x, len y)
    vertices, with autorange over every stored value."""
    rng = np.random.default_rng(21)
    rows = 1500
    xs, ys, vs = [], [], rng.uniform(0, 1, rows)
    for _ in range(rows):
        lx = int(rng.integers(1, 40))
        ly = lx + int(rng.integers(-3, 4))
        ly = max(1, ly)
        x = np.cumsum(rng.normal(0, 0.3, lx)).astype(dt)
        y = np.cumsum(rng.normal(0, 0.3, ly)).astype(dt)
        if lx > 5 and rng.random() < 0.2:
            x[int(rng.integers(1, lx - 1))] = np.nan
        xs.append(x)
        ys.append(y)
    df = pd.DataFrame({"x": pd.Series(xs, dtype=object), "y": pd.Series(ys, dtype=object),
                       "v": vs})
    W, H = 211, 157
    for agg, kind in ((dsa.count(), "count"), (dsa.max("v"), "max")):
        got = dsa.Canvas(W, H).line(df, "x", "y", agg=agg, axis=1)
        case = {"frame": None, "kwargs": {"x": "x", "y": "y", "axis": 1,
                                          "agg": {"kind": kind,
                                                  "column": None if kind == "count" else "v"}},
                "canvas": {"W": W, "H": H, "x_range": None, "y_range": None}}
        want, (xr_, yr_) = _oracle_frame_case(case, df)
        assert np.allclose(got.attrs["x_range"], xr_) and np.allclose(got.attrs["y_range"], yr_)
        G.assert_same(got.values, want)


def _oracle_frame_case(case, df):
    """G.line_case_oracle for an in-memory frame."""
    orig = G.line_frame
    G.line_frame = lambda spec: df
    try:
        return G.line_case_oracle(case)
    finally:
        G.line_frame = orig


# ---- float32 vertices exactly on a view bound -----------------------------------------
# The reference computes a segment's delta in the column's float32 and its distance to
# each bound in float64 (line.py:752-759); a vertex on a bound whose float32 delta rounds
# above that distance makes _clipt clip the segment (line.py:792-798), which changes the
# pixels of a zero-length segment (line.py:1003-1008).

@pytest.mark.parametrize("case", G.f32_bound_cases(), ids=lambda c: c["name"])
def test_f32_vertex_on_bound_axis0(case):
    W, H = case["W"], case["H"]
    view = G.view_for(W, H, case["x_range"], case["y_range"], xf32=True, yf32=True)
    exp = oracle.line_axis0(case["xs"].astype(float), case["ys"].astype(float), view, W, H)["count"]
    assert exp.sum() == case["total_f32"]
    got = abi_lines(case["xs"], case["ys"], view, W, H, xdtype=np.float32, ydtype=np.float32)
    np.testing.assert_array_equal(got, exp)
    df = pd.DataFrame({"x": case["xs"], "y": case["ys"]})
    agg = dsa.Canvas(W, H, x_range=case["x_range"], y_range=case["y_range"]).line(
        df, "x", "y", agg=dsa.count())
    np.testing.assert_array_equal(agg.values, exp)


@pytest.mark.parametrize("xdtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", G.f32_bound_cases(), ids=lambda c: c["name"])
def test_f32_vertex_on_bound_xconst_strips(case, xdtype):
    """The same polylines as x-constant axis-1 lines (the LDS-strip kernel): many lines
    sharing x, every one with the case's float32 y."""
    W, H = case["W"], case["H"]
    nlines = 3000
    xs = case["xs"].astype(xdtype)
    ys = np.tile(case["ys"], (nlines, 1))
    view = G.view_for(W, H, case["x_range"], case["y_range"], xf32=xdtype == np.float32,
                      yf32=True)
    exp = oracle.lines_xconst(xs.astype(float), ys.astype(float), view, W, H)["count"]
    one = oracle.line_axis0(xs.astype(float), case["ys"].astype(float), view, W, H)["count"]
    np.testing.assert_array_equal(exp, one * nlines)
    got = abi_lines(xs, ys.T, view, W, H, axis=1, x_ls=0, x_vs=1, y_ls=1, y_vs=nlines,
                    nlines=nlines, nverts=3, xdtype=xdtype, ydtype=np.float32)
    np.testing.assert_array_equal(got, exp)
    names = ["y0", "y1", "y2"]
    agg = dsa.Canvas(W, H, x_range=case["x_range"], y_range=case["y_range"]).line(
        dsa.DeviceFrame.from_pandas(pd.DataFrame(ys, columns=names)), x=xs, y=names,
        agg=dsa.count(), axis=1)
    np.testing.assert_array_equal(agg.values, exp)


@pytest.mark.parametrize("side", ["ymin", "ymax"])
def test_f32_random_vertices_on_bounds_vs_oracle(side):
    """Random float32 lines whose vertices sit on a bound next to vertices of the other
    sign inside the bound's pixel (rounded float32 deltas): x-constant strips, per-line x
    (the per-segment kernel) and one axis-0 polyline, all against the oracle's float32
    typing."""
    rng = np.random.default_rng(31 if side == "ymin" else 32)
    f = np.float32
    W, H = 9, 7
    b = float(f(0.043172628))
    yr = (-b, 0.5) if side == "ymin" else (-0.5, b)
    bound = f(yr[0] if side == "ymin" else yr[1])
    xr = (0.0, 1.0)
    nlines, nverts = 4000, 6
    xs = np.sort(rng.uniform(0.23, 0.32, nverts)).astype(f)  # one pixel column
    ys = rng.uniform(yr[0], yr[1], (nlines, nverts)).astype(f)
    pick = rng.random((nlines, nverts)) < 0.4
    ys[pick] = bound
    near = (rng.random((nlines, nverts)) < 0.4) & ~pick
    ys[near] = (-bound * rng.uniform(0.05, 0.75, near.sum())).astype(f)
    view = G.view_for(W, H, xr, yr, xf32=True, yf32=True)
    exp = oracle.lines_xconst(xs.astype(float), ys.astype(float), view, W, H)["count"]
    strip = abi_lines(xs, ys.T, view, W, H, axis=1, x_ls=0, x_vs=1, y_ls=1, y_vs=nlines,
                      nlines=nlines, nverts=nverts, xdtype=np.float32, ydtype=np.float32)
    np.testing.assert_array_equal(strip, exp)
    general = abi_lines(np.tile(xs, (nlines, 1)), ys, view, W, H, axis=1, x_ls=nverts,
                        x_vs=1, y_ls=nverts, y_vs=1, nlines=nlines, nverts=nverts,
                        xdtype=np.float32, ydtype=np.float32)
    np.testing.assert_array_equal(general, exp)
    # the same vertices as one long axis-0 polyline
    px = np.tile(xs, nlines)
    py = ys.reshape(-1)
    exp0 = oracle.line_axis0(px.astype(float), py.astype(float), view, W, H)["count"]
    got0 = abi_lines(px, py, view, W, H, xdtype=np.float32, ydtype=np.float32)
    np.testing.assert_array_equal(got0, exp0)
    # the float64 typing of the same vertices differs (the case is not vacuous)
    v64 = G.view_for(W, H, xr, yr)
    assert not np.array_equal(exp, oracle.lines_xconst(xs.astype(float), ys.astype(float),
                                                        v64, W, H)["count"])

"""Pin the CPU oracle before trusting it: against the reference's own test expectations
(golden fixtures), the touched-pixel masks of the reference's line images, numpy closed
forms and brute force on tiny inputs.  CPU only."""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

import oracle
from oracle import closed_forms as cf

import golden_utils as G

GC = G.cases()


# ---- points: the reference's expectations ------------------------------------------

@pytest.mark.parametrize("case", GC["points"], ids=[c["name"] for c in GC["points"]])
def test_points_golden(case):
    got, (xr_, yr_) = G.point_case_oracle(case)
    G.assert_same(got, G.expected_array(case))
    if "x_range" in case:
        assert np.allclose(xr_, case["x_range"]) and np.allclose(yr_, case["y_range"])


@pytest.mark.parametrize("high", [9, 10, 99, 100])
def test_uniform_diagonal_points(high):
    """test_pandas.py:1147-1168: a square grid of integer points, 2x2 canvas."""
    low = 0
    width = height = high - low
    n = width * height
    x = np.array([np.arange(low, high, dtype="f8")] * width).flatten()
    y = np.array([np.arange(low, high, dtype="f8")] * height).flatten()
    view = G.view_for(2, 2, (low, high), (low, high))
    agg = oracle.points(x, y, view, 2, 2, ("count",))["count"]
    diagonal = agg.diagonal(0)
    assert sum(diagonal) == n
    assert abs(high - low) % 2 == abs(diagonal[1] / high - diagonal[0] / high)


# ---- points: closed forms on random inputs ----------------------------------------

def _random_points(rng, n, nan_frac=0.01, f32=False):
    x = rng.normal(0.0, 3.0, n)
    y = rng.normal(0.5, 2.0, n)
    v = rng.lognormal(0.0, 2.0, n)
    if f32:
        x, y, v = x.astype(np.float32), y.astype(np.float32), v.astype(np.float32)
    k = int(n * nan_frac)
    x[rng.choice(n, k, replace=False)] = np.nan
    y[rng.choice(n, k, replace=False)] = np.nan
    v[rng.choice(n, k, replace=False)] = np.nan
    return x, y, v


@pytest.mark.parametrize("f32", [False, True])
def test_points_closed_forms(rng, f32):
    W, H = 97, 61
    x, y, v = _random_points(rng, 200_000, f32=f32)
    view = G.view_for(W, H, (-6.0, 7.5), (-3.0, 4.0))
    bins_c = oracle.point_bins(x, y, view, W, H)
    bins_np = cf.point_bins(x, y, view, W, H)
    np.testing.assert_array_equal(bins_c, bins_np)
    r = oracle.points(x, y, view, W, H, BASES_ALL, val=v)
    np.testing.assert_array_equal(r["count"], cf.count(bins_np, W, H))
    np.testing.assert_array_equal(r["countv"], cf.count(bins_np, W, H, v))
    # bincount adds sequentially in row order: bit-identical to the row-ordered f64 sum
    np.testing.assert_array_equal(r["sumz"], cf.sum_zero(bins_np, v, W, H))
    G.assert_same(r["sum"], cf.sum_nan(bins_np, v, W, H))
    mn, mx = cf.minmax(bins_np, v, W, H)
    G.assert_same(r["min"], mn)
    G.assert_same(r["max"], mx)
    np.testing.assert_array_equal(r["any"], cf.count(bins_np, W, H) > 0)
    np.testing.assert_array_equal(r["anyv"], cf.count(bins_np, W, H, v) > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        var = np.where(r["countv"] > 0, r["m2"] / r["countv"], np.nan)
    ref = cf.var2pass(bins_np, v, W, H)
    ok = ~np.isnan(ref)
    assert np.array_equal(np.isnan(var), np.isnan(ref))
    # Welford-on-sums vs two-pass: same quantity, different rounding
    assert np.allclose(var[ok], ref[ok], rtol=1e-9, atol=0)


BASES_ALL = ("count", "countv", "any", "anyv", "sum", "sumz", "min", "max", "m2")


def test_points_edges_and_wraparound():
    """Upper bound maps into the last bin; values below xmin are dropped; a tiny
    negative product truncates to bin 0."""
    view = G.view_for(4, 4, (0.0, 1.0), (0.0, 1.0))
    x = np.array([0.0, 1.0, -1e-300, 0.999999, 1.0000001, np.nan, 0.5])
    y = np.array([0.0, 1.0, 0.0, 0.999999, 0.5, 0.5, np.inf])
    bins = oracle.point_bins(x, y, view, 4, 4)
    assert bins.tolist() == [0, 15, -1, 15, -1, -1, -1]


def test_points_brute_force_tiny(rng):
    """Exhaustive small inputs: every row's bin computed by hand in Python floats."""
    W, H = 3, 2
    xr_, yr_ = (-1.0, 2.0), (0.0, 1.0)
    view = G.view_for(W, H, xr_, yr_)
    vals = [-1.0, -0.5, 0.0, 0.3333333333333333, 1.0, 1.9999999, 2.0, 2.5, math.nan]
    ys = [-0.1, 0.0, 0.49999, 0.5, 1.0, math.nan]
    pts = list(itertools.product(vals, ys))
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    exp = np.zeros((H, W), np.uint32)
    for a, b in pts:
        if xr_[0] <= a <= xr_[1] and yr_[0] <= b <= yr_[1]:
            xx = int(a * view.sx + view.tx)
            yy = int(b * view.sy + view.ty)
            exp[min(yy, H - 1), min(xx, W - 1)] += 1
    got = oracle.points(x, y, view, W, H, ("count",))["count"]
    np.testing.assert_array_equal(got, exp)


def test_by_category_wraparound():
    """A code of -1 (pandas' NaN category) lands in the last category."""
    view = G.view_for(2, 1, (0.0, 1.0), (0.0, 1.0))
    x = np.array([0.1, 0.1, 0.9, 0.9])
    y = np.zeros(4)
    cat = np.array([0, -1, 1, 2])
    r = oracle.points(x, y, view, 2, 1, ("count",), cat=cat, ncat=3, squeeze=False)["count"]
    assert r[0, 0].tolist() == [1, 0, 1] and r[0, 1].tolist() == [0, 1, 1]
    # sum over categories equals the plain count
    plain = oracle.points(x, y, view, 2, 1, ("count",))["count"]
    np.testing.assert_array_equal(r.sum(axis=2), plain)


def test_bounds():
    a = np.array([np.nan, 3.0, -2.5, np.inf, np.nan])
    assert oracle.bounds(a) == (-2.5, np.inf)
    assert oracle.bounds(np.array([np.nan])) == (np.inf, -np.inf)
    r = np.random.default_rng(0).normal(size=1000)
    assert oracle.bounds(r) == (np.nanmin(r), np.nanmax(r))


# ---- lines: the reference's expectations --------------------------------------------

@pytest.mark.parametrize("case", GC["glyph_lines"], ids=[c["name"] for c in GC["glyph_lines"]])
def test_glyph_lines_golden(case):
    v = case["view"]
    view = oracle.View(v["sx"], v["tx"], v["sy"], v["ty"], v["xmin"], v["xmax"], v["ymin"],
                       v["ymax"])
    W, H = case["W"], case["H"]
    if case["kind"] == "segment":
        (x0, y0), (x1, y1) = case["p0"], case["p1"]
        agg = oracle.draw_segment(view, W, H, case["segment_start"], x0, y0, x1, y1)["count"]
    else:
        agg = oracle.line_axis0(G.dec(case["xs"]), G.dec(case["ys"]), view, W, H,
                                plot_start=case["plot_start"])["count"]
    if case["expected"] is not None:
        np.testing.assert_array_equal(agg, np.array(case["expected"]))
    if case.get("total") is not None:
        assert agg.sum() == case["total"]
    if case.get("at"):
        r, c_, val = case["at"]
        assert agg[r, c_] == val


@pytest.mark.parametrize("case", GC["lines"], ids=[c["name"] for c in GC["lines"]])
def test_lines_golden(case):
    got, (xr_, yr_) = G.line_case_oracle(case)
    if case.get("ones_at") is not None:
        ones = np.argwhere(got == 1).tolist()
        assert ones == case["ones_at"]
        assert got.sum() == len(ones)
    else:
        G.assert_same(got, G.expected_array(case))
    if "x_range" in case:
        assert np.allclose(xr_, case["x_range"]) and np.allclose(yr_, case["y_range"])


@pytest.mark.parametrize("case", GC["noaa_masks"], ids=[c["name"] for c in GC["noaa_masks"]])
def test_noaa_masks(case):
    """Union of the pixels the oracle's Bresenham touches == the non-white pixels of
    the reference's rendered image."""
    cv = case["canvas"]
    W, H = cv["W"], cv["H"]
    view = G.view_for(W, H, cv["x_range"], cv["y_range"])
    touched = np.zeros((H, W), bool)
    for pl in case["polylines"]:
        pl = np.asarray(pl, float)
        touched |= oracle.line_axis0(pl[:, 0], pl[:, 1], view, W, H, ("count",))["count"] > 0
    np.testing.assert_array_equal(touched, np.array(case["expected"], bool))


# ---- lines: brute-force properties ---------------------------------------------------

def test_lines_xconst_equals_axis1_and_axis0(rng):
    """The three layouts draw the same segments: x-constant == axis1 with broadcast x ==
    one axis0 call per row."""
    nrows, ncols, W, H = 40, 17, 31, 23
    xs = np.sort(rng.uniform(-1, 1, ncols))
    ys = rng.normal(0, 1, (nrows, ncols)).cumsum(axis=1) * 0.3
    ys[3, 5] = np.nan
    view = G.view_for(W, H, (-0.8, 0.9), (-1.5, 1.2))
    a = oracle.lines_xconst(xs, ys, view, W, H)["count"]
    b = oracle.lines_axis1(np.broadcast_to(xs, ys.shape), ys, view, W, H)["count"]
    np.testing.assert_array_equal(a, b)
    c = sum(oracle.line_axis0(xs, ys[i], view, W, H, plot_start=True)["count"].astype(np.int64)
            for i in range(nrows))
    np.testing.assert_array_equal(a, c)


def test_line_unclipped_pixel_count(rng):
    """An unclipped segment drawn with segment_start touches max(|dx|,|dy|)+1 pixels
    (Bresenham steps once per major-axis pixel), its two snapped ends included."""
    W = H = 64
    view = G.view_for(W, H, (0.0, 1.0), (0.0, 1.0))
    for _ in range(200):
        p = rng.uniform(0.0, 0.999, 4)
        agg = oracle.draw_segment(view, W, H, True, p[0], p[1], p[2], p[3])["count"]
        x0, y0 = int(p[0] * W), int(p[1] * H)
        x1, y1 = int(p[2] * W), int(p[3] * H)
        n = max(abs(x1 - x0), abs(y1 - y0)) + 1
        assert agg.sum() == n
        assert agg[y0, x0] >= 1 and agg[y1, x1] >= 1


def test_line_f32_delta_flag():
    """The float32 segment delta only changes results when a segment is clipped."""
    W = H = 16
    v64 = G.view_for(W, H, (0.0, 1.0), (0.0, 1.0))
    v32 = G.view_for(W, H, (0.0, 1.0), (0.0, 1.0), xf32=True, yf32=True)
    xs = np.array([0.1, 0.7, 0.3], dtype=np.float32).astype(float)
    ys = np.array([0.2, 0.9, 0.4], dtype=np.float32).astype(float)
    a = oracle.line_axis0(xs, ys, v64, W, H)["count"]
    b = oracle.line_axis0(xs, ys, v32, W, H)["count"]
    np.testing.assert_array_equal(a, b)


def test_line_f32_delta_known_answer():
    """Known answer, derived by hand from line.py's arithmetic for the float32 polyline
    (0.30, -0.5) -> (0.3397695, -0.027073061) -> (0.33822325, 0.043172628) on a 34 x 7
    canvas, x_range (0, 1), y_range (-1, 0.043172628f):
    sx = 34, sy = 7 / 1.043172628 = 6.7103, so the vertices snap to (10, 3), (11, 6) and
    (11, 7 -> 6).  Segment 1 is inside the view: Bresenham from (10, 3) to (11, 6) (y major,
    error -2, +2 -> step x, ...) gives (10, 3), (10, 4), (11, 5), (11, 6).  Segment 2 ends
    on ymax; its float32 dy = fl32(0.043172628 + 0.027073061) = 0.0702456906 exceeds
    q = ymax - y0 = 0.0702456888 (float64), so _clipt sets t1 = q / p < 1: the segment is
    clipped, zero-length in pixels, and draws nothing.  With float64 deltas it is not
    clipped and appends (11, 6) once more."""
    f = np.float32
    assert float(f(0.043172628) - f(-0.027073061)) > float(f(0.043172628)) - float(f(-0.027073061))
    xs = np.array([0.30, 0.3397695, 0.33822325], f).astype(float)
    ys = np.array([-0.5, -0.027073061, 0.043172628], f).astype(float)
    yr = (-1.0, float(f(0.043172628)))
    want = np.zeros((7, 34), np.uint32)
    for r, c in ((3, 10), (4, 10), (5, 11), (6, 11)):
        want[r, c] = 1
    v32 = G.view_for(34, 7, (0.0, 1.0), yr, xf32=True, yf32=True)
    np.testing.assert_array_equal(oracle.line_axis0(xs, ys, v32, 34, 7)["count"], want)
    v64 = G.view_for(34, 7, (0.0, 1.0), yr)
    want[6, 11] = 2
    np.testing.assert_array_equal(oracle.line_axis0(xs, ys, v64, 34, 7)["count"], want)


@pytest.mark.parametrize("case", G.f32_bound_cases(), ids=lambda c: c["name"])
def test_line_f32_delta_on_each_bound(case):
    """A float32 vertex exactly on each of the four bounds with a rounded-up float32
    delta: float32 mode clips the second segment (4 pixels), float64 mode draws its
    unclipped zero-length pixel once more (5 pixels, one cell differing by one)."""
    xs, ys = case["xs"].astype(float), case["ys"].astype(float)
    W, H = case["W"], case["H"]
    v32 = G.view_for(W, H, case["x_range"], case["y_range"], xf32=True, yf32=True)
    v64 = G.view_for(W, H, case["x_range"], case["y_range"])
    a = oracle.line_axis0(xs, ys, v32, W, H)["count"].astype(np.int64)
    b = oracle.line_axis0(xs, ys, v64, W, H)["count"].astype(np.int64)
    assert a.sum() == case["total_f32"] and b.sum() == case["total_f64"]
    d = b - a
    assert d.min() == 0 and d.sum() == 1 and a[d == 1].item() == 1


def test_line_f32_delta_matches_f64_when_deltas_agree():
    """When both ends are inside the view and each float32 delta equals the float64
    difference of its ends, every _clipt call leaves (t0, t1) = (0, 1) (q >= p by the
    monotonicity of rounding), so float32 and float64 modes draw the same pixels."""
    rng = np.random.default_rng(23)
    W, H = 29, 11
    f = np.float32
    xr = (float(f(-0.61)), float(f(0.47)))
    yr = (float(f(-0.043)), float(f(0.39)))
    v32 = G.view_for(W, H, xr, yr, xf32=True, yf32=True)
    v64 = G.view_for(W, H, xr, yr)
    same = 0
    for _ in range(3000):
        xs = rng.uniform(xr[0], xr[1], 2).astype(f)
        ys = rng.uniform(yr[0], yr[1], 2).astype(f)
        k = rng.integers(0, 4)  # one end on a bound
        (xs if k < 2 else ys)[1] = f((xr if k < 2 else yr)[k % 2])
        exact = (float(xs[1] - xs[0]) == float(xs[1]) - float(xs[0]) and
                 float(ys[1] - ys[0]) == float(ys[1]) - float(ys[0]))
        if not exact:
            continue
        a = oracle.line_axis0(xs.astype(float), ys.astype(float), v32, W, H)["count"]
        b = oracle.line_axis0(xs.astype(float), ys.astype(float), v64, W, H)["count"]
        np.testing.assert_array_equal(a, b)
        same += 1
    assert same > 500


@pytest.mark.parametrize("side", ["xmin", "xmax", "ymin", "ymax"])
def test_line_f32_zero_length_clip_rule(side):
    """A non-start segment inside one pixel that ends exactly on a bound, from a vertex of
    the opposite sign: with float64 deltas it is unclipped and zero-length and appends
    its pixel once (line.py:1003-1005); with float32 deltas it is clipped exactly when
    p = fl32(delta toward the bound) exceeds q = the float64 distance to the bound
    (_clipt's q < p branch, line.py:792-798), and then draws nothing.  The rule is
    evaluated with numpy's float32 / float64 arithmetic on random vertices."""
    rng = np.random.default_rng({"xmin": 1, "xmax": 2, "ymin": 3, "ymax": 4}[side])
    f = np.float32
    W = H = 7
    b = float(f(0.043172628))
    lo_hi = (-b, 0.5) if side.endswith("min") else (-0.5, b)
    rest = (0.0, 1.0)
    xr, yr = (lo_hi, rest) if side[0] == "x" else (rest, lo_hi)
    v32 = G.view_for(W, H, xr, yr, xf32=True, yf32=True)
    v64 = G.view_for(W, H, xr, yr)
    clipped = kept = 0
    for _ in range(2000):
        bound = f(lo_hi[0] if side.endswith("min") else lo_hi[1])
        # the other vertex: opposite sign, within the bound's pixel (pixel size 0.0776)
        other = f(-bound * rng.uniform(0.05, 0.75))
        o2 = f(rng.uniform(0.31, 0.32))   # the other coordinate, same pixel for both ends
        p = abs(float(f(bound - other)))
        q = abs(float(bound) - float(other))
        if side[0] == "x":
            x0, x1, y0, y1 = float(other), float(bound), float(o2), float(o2)
        else:
            x0, x1, y0, y1 = float(o2), float(o2), float(other), float(bound)
        a = oracle.draw_segment(v32, W, H, False, x0, y0, x1, y1)["count"].sum()
        c = oracle.draw_segment(v64, W, H, False, x0, y0, x1, y1)["count"].sum()
        assert c == 1
        assert a == (0 if p > q else 1), (x0, y0, x1, y1)
        clipped += p > q
        kept += p <= q
    assert clipped > 100 and kept > 100


@pytest.mark.parametrize("mode", ["first", "last", "max", "min"])
def test_row_index_loop_matches_closed_form(mode):
    """The sequential row-index restatement against its vectorised statement, on random
    cells with ties (integer values) and NaNs."""
    from oracle import rowindex as ri
    rng = np.random.default_rng(17)
    n, ncells = 20_000, 97
    cell = rng.integers(-1, ncells, n)
    val = rng.integers(0, 6, n).astype(float)
    val[rng.random(n) < 0.1] = np.nan
    for v in (val, None) if mode in ("first", "last") else (val,):
        np.testing.assert_array_equal(ri.select_rows(cell, ncells, mode, v),
                                      ri.selected_rows_closed_form(cell, ncells, mode, v))


# ---- the timed CPU baseline (oracle/cpu_baseline.c) equals the sequential oracle -------

@pytest.mark.parametrize("threads", [1, 3, 8])
def test_cpu_baseline_count_mean_equals_oracle(threads):
    from oracle import baseline
    rng = np.random.default_rng(31)
    n, W, H = 200_003, 97, 61
    x = rng.normal(0.5, 0.3, n).astype(np.float32)
    y = rng.normal(0.5, 0.3, n).astype(np.float32)
    v = rng.lognormal(0, 2, n).astype(np.float32)
    x[::101] = np.nan
    v[::7] = np.nan
    x[5], y[5] = 1.0, 1.0  # on the upper bounds: the W-1 / H-1 clamp
    sx, sy = W / 1.0, H / 1.0
    view = oracle.View(sx, 0.0, sy, 0.0, 0.0, 1.0, 0.0, 1.0)
    got = baseline.points_count_mean_f32(x, y, v, view, W, H, threads)
    ref = oracle.points(x, y, view, W, H, ("count", "countv", "sumz"), val=v)
    np.testing.assert_array_equal(got["count"], ref["count"])
    np.testing.assert_array_equal(got["countv"], ref["countv"])
    if threads == 1:  # one partition: the oracle's row order exactly
        np.testing.assert_array_equal(got["sumz"], ref["sumz"])
    else:             # partitions merged by the combine tree: another summation order
        np.testing.assert_allclose(got["sumz"], ref["sumz"], rtol=1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = np.where(ref["countv"] > 0, ref["sumz"] / ref["countv"], np.nan)
    np.testing.assert_allclose(got["mean"], mean, rtol=1e-12)
    assert np.array_equal(np.isnan(got["mean"]), np.isnan(mean))


@pytest.mark.parametrize("threads", [1, 5])
def test_cpu_baseline_count_f64_and_categories_equal_oracle(threads):
    from oracle import baseline
    rng = np.random.default_rng(32)
    n, W, H = 150_000, 80, 60
    x = rng.normal(0, 5, n)
    y = rng.normal(0, 5, n)
    x[::997] = np.nan
    view = oracle.View(4.0, 40.0, 3.0, 30.0, -10, 10, -10, 10)
    got = baseline.points_count_f64(x, y, view, W, H, threads)
    np.testing.assert_array_equal(got, oracle.points(x, y, view, W, H, ("count",))["count"])
    cat = rng.integers(0, 16, n).astype(np.uint8)
    got = baseline.points_count_cat_f32(x.astype(np.float32), y.astype(np.float32), cat, view,
                                        W, H, 16, threads)
    ref = oracle.points(x.astype(np.float32), y.astype(np.float32), view, W, H, ("count",),
                        cat=cat.astype(np.int64), ncat=16, squeeze=False)["count"]
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_baseline_lines_equal_oracle(threads):
    from oracle import baseline
    rng = np.random.default_rng(33)
    nl, nv, W, H = 600, 64, 120, 90
    xs = np.linspace(0, 1, nv)
    ys = np.cumsum(rng.normal(0, 0.05, (nl, nv)), axis=1).astype(np.float32)
    ys[3, 10] = np.nan
    ys[7, :] = 0.25            # a constant line: horizontal segments only
    lo, hi = float(np.nanmin(ys)), float(np.nanmax(ys))
    sx, sy = W / 1.0, H / (hi - lo)
    view = oracle.View(sx, 0.0, sy, -lo * sy, 0.0, 1.0, lo, hi, yf32=True)
    got = baseline.lines_xconst_f32(xs, ys, view, W, H, threads)
    ref = oracle.lines_xconst(xs, ys.astype(np.float64), view, W, H)["count"]
    np.testing.assert_array_equal(got, ref)
    # a narrower view clips many segments (float32 deltas on the clipped ones)
    view = oracle.View(W / 0.5, -0.25 * W / 0.5, H / 0.4, 0.2 * H / 0.4, 0.25, 0.75, -0.2, 0.2,
                       yf32=True)
    got = baseline.lines_xconst_f32(xs, ys, view, W, H, threads)
    ref = oracle.lines_xconst(xs, ys.astype(np.float64), view, W, H)["count"]
    np.testing.assert_array_equal(got, ref)

"""GPU parity of the viewport server: the grid index, the row gather, and requests served
through them (alone and overlapping on several streams) against the plain
Canvas.points call over the whole resident set."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import datashader_amd as dsa  # noqa: E402
from datashader_amd import serving, synthetic  # noqa: E402


@pytest.fixture(scope="module")
def resident():
    n = 2_000_000
    x, y, v = synthetic.points_mixture_dev(n, seed=7, nan_frac=0.001, dtype=torch.float32,
                                           with_value=True)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
    index = serving.PointIndex(x, y, v, grid=64)
    return frame, index


def test_index_is_a_sort_by_cell(resident):
    frame, index = resident
    x = frame["x"].cpu().numpy().astype(np.float64)
    y = frame["y"].cpu().numpy().astype(np.float64)
    ox = index.frame["x"].cpu().numpy().astype(np.float64)
    oy = index.frame["y"].cpu().numpy().astype(np.float64)
    ov = index.frame["v"].cpu().numpy()
    G = index.grid

    def cells(a, b):
        nan = np.isnan(a) | np.isnan(b)
        fx = np.clip((np.nan_to_num(a) - index.x0) * index.sx, 0, G - 1).astype(np.int64)
        fy = np.clip((np.nan_to_num(b) - index.y0) * index.sy, 0, G - 1).astype(np.int64)
        return np.where(nan, G * G, fy * G + fx)

    k = cells(ox, oy)
    assert np.all(np.diff(k) >= 0)                       # sorted by cell, NaN rows last
    np.testing.assert_array_equal(np.sort(cells(x, y)), k)
    cs = index.cell_start
    assert cs[0] == 0 and np.all(np.diff(cs) >= 0)
    np.testing.assert_array_equal(cs[:-1], np.searchsorted(k, np.arange(G * G)))
    assert index.indexed == int((k < G * G).sum())
    # the rows moved together: the (x, y, v) triples are a permutation of the input
    v = frame["v"].cpu().numpy()
    a = np.lexsort((v, np.nan_to_num(y, nan=9e9), np.nan_to_num(x, nan=9e9)))
    b = np.lexsort((ov, np.nan_to_num(oy, nan=9e9), np.nan_to_num(ox, nan=9e9)))
    np.testing.assert_array_equal(np.nan_to_num(x[a], nan=9e9), np.nan_to_num(ox[b], nan=9e9))
    np.testing.assert_array_equal(v[a], ov[b])


def _requests(index, k, seed):
    frame_rows = (index.frame["x"][:index.indexed].cpu().numpy(),
                  index.frame["y"][:index.indexed].cpu().numpy())
    reqs = serving.make_requests(k, frame_rows, ((index.x0, index.x1), (index.y0, index.y1)),
                                 seed=seed, sizes=((64, 48), (200, 150), (512, 384)))
    # a view of the whole set (full scan), one outside it, one on a cell boundary
    agg = reqs[0].agg
    reqs.append(serving.Request((index.x0, index.x1), (index.y0, index.y1), 300, 200, agg))
    reqs.append(serving.Request((index.x1 + 1, index.x1 + 2), (0, 1), 32, 32, agg))
    edge = index.x0 + 17 / index.sx
    reqs.append(serving.Request((edge - 0.5, edge), (-1.0, 1.0), 128, 128, agg))
    return reqs


def _check_same(a, b):
    np.testing.assert_array_equal(a["count"].values, b["count"].values)
    ma, mb = a["mean"].values, b["mean"].values
    np.testing.assert_array_equal(np.isnan(ma), np.isnan(mb))
    ok = ~np.isnan(ma)
    # float32 values summed in float64 in another order: a few ulps of float64
    np.testing.assert_allclose(ma[ok], mb[ok], rtol=1e-12, atol=0)


def test_indexed_requests_match_plain(resident):
    frame, index = resident
    for req in _requests(index, 24, seed=3):
        plain = serving.run_plain(frame, req)
        got = serving.run_indexed(index, req)
        _check_same(got, plain)
        # the rows behind the view, gathered, aggregate to the same result
        sub, rows = index.candidates(req.x_range, req.y_range)
        _check_same(req.canvas().points(sub, "x", "y", req.agg), plain)
        assert int(got["count"].values.sum()) == int(plain["count"].values.sum())


def test_server_alone_and_in_the_mix(resident):
    frame, index = resident
    reqs = _requests(index, 40, seed=4)
    alone = serving.ViewportServer(index, slots=1).serve(reqs, realtime=False, keep_results=True)
    mixed = serving.ViewportServer(index, slots=4, mem_budget=1 << 28).serve(
        reqs, realtime=False, keep_results=True)
    torch.cuda.synchronize()
    for r, a, m in zip(reqs, alone["results"], mixed["results"]):
        _check_same(a, m)
        _check_same(m, serving.run_plain(frame, r))
    assert np.all(mixed["latency"] >= 0)


def test_indexed_all_reductions(resident):
    """Every points reduction over the indexed frame equals the plain call."""
    frame, index = resident
    agg = dsa.summary(cnt=dsa.count(), cntv=dsa.count("v"), any=dsa.any(), s=dsa.sum("v"),
                      mean=dsa.mean("v"), mn=dsa.min("v"), mx=dsa.max("v"),
                      var=dsa.var("v"), std=dsa.std("v"))
    for xr, yr, W, H in (((index.x0, index.x1), (index.y0, index.y1), 640, 480),
                         ((-1.0, 0.5), (-2.0, 1.0), 300, 500),
                         ((0.25, 0.26), (0.1, 0.12), 97, 61)):
        c = dsa.Canvas(W, H, x_range=xr, y_range=yr)
        a, b = c.points(index.frame, "x", "y", agg), c.points(frame, "x", "y", agg)
        for k in ("cnt", "cntv", "any", "mn", "mx"):
            np.testing.assert_array_equal(a[k].values, b[k].values)
        for k, tol in (("s", 1e-12), ("mean", 1e-12), ("var", 1e-9), ("std", 1e-9)):
            va, vb = a[k].values, b[k].values
            np.testing.assert_array_equal(np.isnan(va), np.isnan(vb))
            ok = ~np.isnan(va)
            np.testing.assert_allclose(va[ok], vb[ok], rtol=tol, atol=1e-12)


def test_indexed_vs_oracle_at_serving_canvases(resident):
    """dsx_points_agg_indexed (Canvas.points over the indexed frame) against the oracle on
    the unsorted rows, at the serving workload's canvases {256^2, 512^2, 1024x768,
    1920x1080} and views drawn by make_requests: counts bit-exact, means within 1e-12
    (float32 values summed in float64, another order)."""
    import oracle
    import golden_utils as G
    from datashader_amd import backend
    frame, index = resident
    x = frame["x"].cpu().numpy()
    y = frame["y"].cpu().numpy()
    v = frame["v"].cpu().numpy()
    rows = (index.frame["x"][:index.indexed].cpu().numpy(),
            index.frame["y"][:index.indexed].cpu().numpy())
    sizes = ((256, 256), (512, 512), (1024, 768), (1920, 1080))
    reqs = serving.make_requests(12, rows, ((index.x0, index.x1), (index.y0, index.y1)),
                                 seed=11, sizes=sizes, zoom_max_log2=6)
    reqs += [serving.Request((index.x0, index.x1), (index.y0, index.y1), w, h, reqs[0].agg)
             for w, h in sizes]
    seen = set()
    for req in reqs:
        W, H = req.width, req.height
        seen.add((W, H))
        view = backend.make_view(req.canvas(), req.x_range, req.y_range)
        p = backend.Pass("v", None)
        assert backend._grid_ref(index.frame, dsa.glyphs.Point("x", "y"), p, view) is not None
        got = serving.run_indexed(index, req)
        ov = G.view_for(W, H, req.x_range, req.y_range)
        r = oracle.points(x, y, ov, W, H, ("count", "countv", "sumz"), val=v)
        np.testing.assert_array_equal(got["count"].values, r["count"])
        with np.errstate(divide="ignore", invalid="ignore"):
            mean = np.where(r["countv"] > 0, r["sumz"] / r["countv"], np.nan)
        gm = got["mean"].values
        np.testing.assert_array_equal(np.isnan(gm), np.isnan(mean))
        ok = ~np.isnan(mean)
        np.testing.assert_allclose(gm[ok], mean[ok], rtol=1e-12, atol=0)
    assert seen == set(sizes)

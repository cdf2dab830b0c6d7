"""GPU parity of the batched served form (dsx_batch through datashader_amd.batching):
every request of a batch against the plain Canvas.points call over the whole resident set
and against the oracle's closed forms; each result alone and in a mix of overlapping
batches; the captured graph against kernel-by-kernel launches; tuned against untuned
knobs.  Tolerances: counts bit-exact; sum / mean within 1e-12 relative (float64 sums of
float32 values in another order)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import closed_forms as cf
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import datashader_amd as dsa  # noqa: E402
from datashader_amd import batching, serving, synthetic  # noqa: E402

SIZES = ((256, 256), (512, 512), (1024, 768), (1920, 1080), (7, 3))


@pytest.fixture(scope="module", params=[1024, 4096], ids=["grid1024", "grid4096"])
def resident(request):
    n = 3_000_000
    x, y, v = synthetic.points_mixture_dev(n, seed=9, nan_frac=0.001, dtype=torch.float32,
                                           with_value=True)
    v[::97] = float("nan")  # rows with a NaN value: counted by count(), not by mean / count(v)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
    index = serving.PointIndex(x, y, v, grid=request.param)
    return frame, index


def _requests(index, k, seed, sizes=SIZES):
    rows = (index.frame["x"][:index.indexed].cpu().numpy(),
            index.frame["y"][:index.indexed].cpu().numpy())
    return serving.make_requests(k, rows, ((index.x0, index.x1), (index.y0, index.y1)),
                                 seed=seed, sizes=sizes,
                                 agg=dsa.summary(c=dsa.count(), m=dsa.mean("v"),
                                                 s=dsa.sum("v"), cv=dsa.count("v")))


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b) & (b != 0)
    return float(np.max(np.abs(a[ok] - b[ok]) / np.abs(b[ok]))) if ok.any() else 0.0


def _same(got, want, what):
    for k in ("c", "cv"):
        np.testing.assert_array_equal(got[k].values, want[k].values, err_msg=f"{what} {k}")
    for k in ("m", "s"):
        assert _rel(got[k].values, want[k].values) <= 1e-12, (what, k)
    np.testing.assert_array_equal(got["c"].coords["x"], want["c"].coords["x"])
    np.testing.assert_array_equal(got["c"].coords["y"], want["c"].coords["y"])


def test_batched_requests_equal_plain_calls(resident):
    frame, index = resident
    reqs = _requests(index, 40, seed=21)
    server = batching.BatchServer(index, agg=reqs[0].agg, engines=2, max_batch=16)
    out = server.serve(reqs, realtime=False, keep_results=True)
    print(f"batching: {len(reqs)} requests in {out['batches']} batches "
          f"(mean {out['mean_batch']:.1f})", flush=True)
    assert out["batches"] < len(reqs)  # requests were batched
    for i, r in enumerate(reqs):
        want = serving.run_plain(frame, r)
        _same(out["results"][i], want, f"request {i} {r.width}x{r.height}")
    server.close()


def test_batched_counts_and_means_vs_oracle(resident):
    frame, index = resident
    reqs = _requests(index, 6, seed=22, sizes=((256, 256), (1920, 1080)))
    server = batching.BatchServer(index, agg=reqs[0].agg, engines=1, max_batch=8)
    out = server.serve(reqs, realtime=False, keep_results=True)
    x = frame["x"].cpu().numpy().astype(np.float64)
    y = frame["y"].cpu().numpy().astype(np.float64)
    v = frame["v"].cpu().numpy()
    for i, r in enumerate(reqs):
        W, H = r.width, r.height
        sx, sy = W / (r.x_range[1] - r.x_range[0]), H / (r.y_range[1] - r.y_range[0])
        view = oracle.View(sx, -r.x_range[0] * sx, sy, -r.y_range[0] * sy, r.x_range[0],
                           r.x_range[1], r.y_range[0], r.y_range[1])
        bins = cf.point_bins(x, y, view, W, H)
        got = out["results"][i]
        np.testing.assert_array_equal(got["c"].values, cf.count(bins, W, H))
        np.testing.assert_array_equal(got["cv"].values, cf.count(bins, W, H, v))
        assert _rel(got["m"].values, cf.mean(bins, v, W, H)) <= 1e-12
        assert _rel(got["s"].values, cf.sum_nan(bins, v, W, H)) <= 1e-12
    server.close()


def test_each_request_alone_equals_in_the_mix(resident):
    """A request served alone (a batch of one) and inside mixed, overlapping batches
    (Poisson arrivals, two engines) gives the same result."""
    frame, index = resident
    reqs = _requests(index, 24, seed=23)
    for k, r in enumerate(reqs):  # arrivals 0.5 ms apart: batches of varying sizes overlap
        r.arrival = 0.0005 * k
    server = batching.BatchServer(index, agg=reqs[0].agg, engines=2, max_batch=8)
    mix = server.serve(reqs, realtime=True, keep_results=True)
    for i, r in enumerate(reqs):
        r.arrival = 0.0
        alone = server.serve([r], realtime=False, keep_results=True)["results"][0]
        _same(mix["results"][i], alone, f"request {i}")
    server.close()


def test_graph_replay_equals_direct_launches(resident):
    frame, index = resident
    reqs = _requests(index, 12, seed=24)
    a = batching.BatchServer(index, agg=reqs[0].agg, engines=1, max_batch=12, use_graph=True)
    b = batching.BatchServer(index, agg=reqs[0].agg, engines=1, max_batch=12, use_graph=False)
    for rep in range(3):  # the graph is replayed: every replay against the eager launches
        ra = a.serve(reqs, realtime=False, keep_results=True)["results"]
        rb = b.serve(reqs, realtime=False, keep_results=True)["results"]
        for i in range(len(reqs)):
            _same(ra[i], rb[i], f"replay {rep} request {i}")
    a.close()
    b.close()


@pytest.mark.parametrize("tile_h,tile_cells,part_rows", [(1, 0, 1 << 10), (16, 0, 1 << 12),
                                                         (128, 0, 1 << 20), (8, 300, 1 << 14)])
def test_results_do_not_depend_on_the_knobs(resident, tile_h, tile_cells, part_rows):
    """Narrow or small tiles, many small parts (heavy tiles split and merged) or none."""
    frame, index = resident
    reqs = _requests(index, 8, seed=25, sizes=((512, 512), (1024, 768)))
    tuning = batching.Tuning(threads=512, per_class={})
    tuning.DEFAULT = {"tile_h": tile_h, "tile_cells": tile_cells, "part_rows": part_rows}
    server = batching.BatchServer(index, agg=reqs[0].agg, engines=1, max_batch=8,
                                  tuning=tuning)
    got = server.serve(reqs, realtime=False, keep_results=True)["results"]
    for i, r in enumerate(reqs):
        _same(got[i], serving.run_plain(frame, r), f"request {i}")
    server.close()


def test_views_outside_the_data_and_tiny_canvases(resident):
    frame, index = resident
    agg = dsa.summary(c=dsa.count(), m=dsa.mean("v"), s=dsa.sum("v"), cv=dsa.count("v"))
    far = (index.x1 + 10.0, index.x1 + 20.0)
    reqs = [serving.Request(far, (index.y0, index.y1), 64, 32, agg),
            serving.Request((index.x0, index.x1), (index.y0, index.y1), 1, 1, agg),
            serving.Request((index.x0 - 1, index.x1 + 1), (index.y0 - 1, index.y1 + 1), 3, 700,
                            agg)]
    server = batching.BatchServer(index, agg=agg, engines=1, max_batch=4)
    got = server.serve(reqs, realtime=False, keep_results=True)["results"]
    assert int(got[0]["c"].values.sum()) == 0
    assert np.isnan(got[0]["m"].values).all()
    for i, r in enumerate(reqs):
        _same(got[i], serving.run_plain(frame, r), f"request {i}")
    server.close()


def test_tuner_picks_knobs_that_keep_results(resident):
    frame, index = resident
    sample = _requests(index, 12, seed=26)
    tuner = batching.BatchTuner(index, planes=("count", "countv", "sum", "mean"))
    tuner.TILE_H, tuner.TILE_DIV, tuner.PART_ROWS = (32, 64), (1, 8), (1 << 15, 1 << 17)
    tuner.THREADS, tuner.WG_PER_CU = (512, 1024), (1,)
    tuning = tuner.tune(sample, max_batch=8)
    assert tuning.per_class and tuning.threads in (512, 1024)
    reqs = _requests(index, 10, seed=27)
    server = batching.BatchServer(index, agg=reqs[0].agg, engines=2, max_batch=8,
                                  tuning=tuning)
    got = server.serve(reqs, realtime=False, keep_results=True)["results"]
    for i, r in enumerate(reqs):
        _same(got[i], serving.run_plain(frame, r), f"request {i}")
    server.close()

"""GPU parity of the two ways a points call splits its rows (``dsx_strategy``): the
window strategy (a hot window aggregated in LDS during the scan, the other rows spilled
and scattered by tile) and the tile strategy (every row bucketed by tile in the scan,
full buckets written as blocks of 16-bit tile offsets + values), each forced, and AUTO.
Each is held to the oracle with the tolerances of test_gpu_points (counts, min, max
bit-exact; float32 sums / means 1e-6, float64 1e-12; var / std 1e-5 / 1e-9), plus the
cases the tile strategy's bucket flush must get right: one hot cell overflowing its
bucket in every chunk (rows written straight to the block stream), bucket-size
boundaries, all rows in one tile."""
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
from datashader_amd import _abi, backend, synthetic  # noqa: E402

import test_gpu_points as TP  # noqa: E402

STRATS = {"window": _abi.STRATEGY_WINDOW, "tiles": _abi.STRATEGY_TILES}


@pytest.fixture(params=list(STRATS))
def strategy(request):
    backend.FORCE_STRATEGY = STRATS[request.param]
    print(f"strategy {request.param}", flush=True)
    yield request.param
    backend.FORCE_STRATEGY = None


@pytest.mark.parametrize("case", TP.GC["points"], ids=[c["name"] for c in TP.GC["points"]])
def test_golden_points(strategy, case):
    TP.test_points_golden(case, "device")


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("shape", [(97, 61), (800, 600)])
def test_random_points_vs_oracle(strategy, dtype, shape):
    TP.test_random_points_vs_oracle(dtype, shape)


def test_by_category_vs_oracle(strategy):
    TP.test_by_category_vs_oracle()
    TP.test_by_category_1024x1024x16_vs_oracle()


def test_skewed_vs_oracle(strategy):
    TP.test_skewed_hot_bins_compensated_sum()
    TP.test_skewed_min_max_std_vs_oracle()


def test_log_axes_empty_and_tiny(strategy):
    TP.test_log_axis_points()
    TP.test_empty_and_tiny_inputs()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4096, 4097, 12_288, 1_000_003])
def test_one_hot_cell(strategy, n):
    """All rows but every 97th in one cell: the tile strategy's bucket for that cell's
    tile overflows in every chunk, so most of its rows go straight to full blocks and
    the rest to the front of the bucket; the counts around the bucket and chunk sizes
    (64, 4096) check the partial block at the end.  Values are small integers, so every
    sum is exact in float64 and the checks are bit-exact."""
    rng = np.random.default_rng(n)
    x = np.full(n, 0.3, np.float32)
    y = np.full(n, 0.7, np.float32)
    spread = np.arange(n) % 97 == 5
    x[spread] = rng.random(int(spread.sum())).astype(np.float32)
    y[spread] = rng.random(int(spread.sum())).astype(np.float32)
    v = (np.arange(n) % 7).astype(np.float32)
    v[np.arange(n) % 11 == 3] = np.nan
    W, H = 640, 480
    f = dsa.DeviceFrame({k: torch.from_numpy(a).cuda() for k, a in (("x", x), ("y", y), ("v", v))})
    ds = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(
        f, "x", "y", dsa.summary(c=dsa.count(), cv=dsa.count("v"), s=dsa.sum("v"),
                                 mn=dsa.min("v"), mx=dsa.max("v")))
    bins = cf.point_bins(x, y, G.view_for(W, H, (0, 1), (0, 1)), W, H)
    np.testing.assert_array_equal(ds["c"].values, cf.count(bins, W, H))
    np.testing.assert_array_equal(ds["cv"].values, cf.count(bins, W, H, v))
    G.assert_same(ds["s"].values, cf.sum_nan(bins, v, W, H))
    r = oracle.points(x, y, G.view_for(W, H, (0, 1), (0, 1)), W, H, ("countv", "min", "max"),
                      val=v)
    G.assert_same(ds["mn"].values, np.where(r["countv"] > 0, r["min"], np.nan))
    G.assert_same(ds["mx"].values, np.where(r["countv"] > 0, r["max"], np.nan))


def test_rows_in_one_tile_by_category(strategy):
    """1024 x 1024 x 16 cells: a tile is one canvas row of 16-category cells; 3M rows on
    canvas row 500 all fall in it (one tile split over many parts, the others empty)."""
    rng = np.random.default_rng(5)
    n = 3_000_000
    W = H = 1024
    x = rng.random(n).astype(np.float32)
    y = np.full(n, (500.5) / H, np.float32)
    cat = rng.integers(0, 16, n).astype(np.int8)
    cats = [f"c{k}" for k in range(16)]
    f = dsa.DeviceFrame({"x": torch.from_numpy(x).cuda(), "y": torch.from_numpy(y).cuda(),
                         "cat": torch.from_numpy(cat).cuda()}, categories={"cat": cats})
    agg = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(
        f, "x", "y", dsa.by("cat", dsa.count()))
    view = G.view_for(W, H, (0, 1), (0, 1))
    exp = oracle.points(x, y, view, W, H, ("count",), cat=cat.astype(np.int64), ncat=16,
                        squeeze=False)["count"]
    np.testing.assert_array_equal(agg.values, exp)
    assert agg.values[500].sum() == n


@pytest.mark.parametrize("n", [1_500_003, 1_500_004])
@pytest.mark.parametrize("dt", [np.int8, np.uint8])
def test_byte_codes_out_of_range_vs_oracle(strategy, dt, n):
    """One-byte category codes outside [0, ncat): an int8 code k < 0 counts as k + ncat
    (once; still negative is dropped), codes >= ncat are dropped, for int8 and uint8
    columns on a 1024 x 256 x 16 canvas, bit-exact vs the oracle (n a multiple of 4: the
    tile scan's 16-byte row loads, four codes in one 4-byte load)."""
    rng = np.random.default_rng(11)
    W, H, ncat = 1024, 256, 16
    x = rng.random(n).astype(np.float32)
    y = rng.random(n).astype(np.float32)
    lo_, hi_ = (-40, 40) if dt is np.int8 else (0, 60)
    cat = rng.integers(lo_, hi_, n).astype(dt)
    cats = [f"c{k}" for k in range(ncat)]
    f = dsa.DeviceFrame({"x": torch.from_numpy(x).cuda(), "y": torch.from_numpy(y).cuda(),
                         "cat": torch.from_numpy(cat).cuda()}, categories={"cat": cats})
    agg = dsa.Canvas(W, H, x_range=(0, 1), y_range=(0, 1)).points(
        f, "x", "y", dsa.by("cat", dsa.count()))
    view = G.view_for(W, H, (0, 1), (0, 1))
    exp = oracle.points(x, y, view, W, H, ("count",), cat=cat.astype(np.int64), ncat=ncat,
                        squeeze=False)["count"]
    np.testing.assert_array_equal(agg.values, exp)
    kept = cat.astype(np.int64)
    kept = np.where(kept < 0, kept + ncat, kept)
    assert agg.values.sum() == int(((kept >= 0) & (kept < ncat)).sum())


def test_auto_picks_by_the_sample_and_agrees():
    """AUTO on a resident frame: spread ridge rows by 16 categories pick the tile
    strategy, a 0.01-sigma blob picks the window; the record keeps the pick, later calls
    pass it explicitly once read back, and every result equals the forced strategies'."""
    backend.FORCE_STRATEGY = None
    n = 2_000_000
    x, y, cat = synthetic.points_ridges_np(n, seed=3)
    cats = [f"c{k}" for k in range(16)]
    fr = dsa.DeviceFrame({"x": torch.from_numpy(x).cuda(), "y": torch.from_numpy(y).cuda(),
                          "cat": torch.from_numpy(cat).cuda()}, categories={"cat": cats})
    xs, ys, vs = synthetic.points_skewed_np(n, seed=2)
    fs = dsa.DeviceFrame({k: torch.from_numpy(a).cuda() for k, a in (("x", xs), ("y", ys), ("v", vs))})
    cases = [(fr, dsa.Canvas(1024, 1024, x_range=(0, 1), y_range=(0, 1)),
              dsa.by("cat", dsa.count()), _abi.STRATEGY_TILES),
             (fs, dsa.Canvas(1920, 1080, x_range=(0, 1), y_range=(0, 1)),
              dsa.summary(c=dsa.count(), m=dsa.mean("v")), _abi.STRATEGY_WINDOW)]
    try:
        for frame, canvas, agg, picked in cases:
            first = canvas.points(frame, "x", "y", agg)
            (slot,) = frame._window_cache.values()
            assert int(slot.rec[4]) == picked
            torch.cuda.synchronize()
            second = canvas.points(frame, "x", "y", agg)
            assert slot.strategy == picked  # read back, passed explicitly
            forced = {}
            for name, s in STRATS.items():
                backend.FORCE_STRATEGY = s
                forced[name] = canvas.points(frame, "x", "y", agg)
            backend.FORCE_STRATEGY = None
            for other in (second, forced["window"], forced["tiles"]):
                if hasattr(first, "data_vars"):  # the summary
                    np.testing.assert_array_equal(other["c"].values, first["c"].values)
                    assert TP._rel_err(other["m"].values, first["m"].values) <= 1e-6
                else:
                    np.testing.assert_array_equal(other.values, first.values)
    finally:
        backend.FORCE_STRATEGY = None

"""Accuracy of the float reductions on the full float-reduction workload: 100M float32
points (90% in one 0.01-sigma Gaussian, 10% uniform), value lognormal(0, 2), Canvas
1920 x 1080 over the data extent, summary(count, sum, mean, min, max, std, var) -- the
same seeded rows bench.py aggregates.  The maximum per-bin relative error of each
reduction against the float64 oracle is printed (run with -s to see it) and held to the
stated tolerances:
  count, min, max                          bit-exact
  sum, mean                                <= 1e-6 relative (float64 closed form)
  var, std                                 <= 1e-5 relative (two-pass float64 closed
                                           form; the reference's sequential m2 is checked
                                           against the same bound at 4M rows in
                                           test_gpu_points.py)
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import closed_forms as cf

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import datashader_amd as dsa  # noqa: E402


def _max_rel(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want) & (want != 0)
    if not ok.any():
        return 0.0
    return float(np.max(np.abs(got[ok] - want[ok]) / np.abs(want[ok])))


def test_float_reductions_error_on_the_100M_skewed_workload():
    from datashader_amd import _device as dev
    from datashader_amd import synthetic
    torch.cuda.set_device(0)
    n, W, H = 100_000_000, 1920, 1080
    print(f"accuracy: {n} skewed rows, {W}x{H}", flush=True)
    x, y, v = synthetic.points_skewed_dev(n, seed=2)
    xr = dev.maybe_expand_bounds(dev.bounds(x))
    yr = dev.maybe_expand_bounds(dev.bounds(y))
    cvs = dsa.Canvas(W, H, x_range=xr, y_range=yr)
    ds = cvs.points(dsa.DeviceFrame({"x": x, "y": y, "v": v}), "x", "y",
                    dsa.summary(count=dsa.count(), sum=dsa.sum("v"), mean=dsa.mean("v"),
                                min=dsa.min("v"), max=dsa.max("v"), std=dsa.std("v"),
                                var=dsa.var("v")))
    got = {k: ds[k].values for k in ("count", "sum", "mean", "min", "max", "std", "var")}
    xs, ys, vs = (t.cpu().numpy() for t in (x, y, v))
    del x, y, v
    torch.cuda.empty_cache()
    sx, tx = cvs.x_axis.compute_scale_and_translate(xr, W)
    sy, ty = cvs.y_axis.compute_scale_and_translate(yr, H)
    view = oracle.View(sx, tx, sy, ty, xr[0], xr[1], yr[0], yr[1])
    bins = oracle.point_bins(xs.astype(np.float64), ys.astype(np.float64), view, W, H)
    del xs, ys
    count = cf.count(bins, W, H)
    total = cf.sum_nan(bins, vs, W, H)
    mean = cf.mean(bins, vs, W, H)
    mn, mx = cf.minmax(bins, vs, W, H)
    var = cf.var2pass(bins, vs, W, H)
    err = {
        "count_mismatched_cells": int(np.count_nonzero(got["count"] != count)),
        "sum": _max_rel(got["sum"], total),
        "mean": _max_rel(got["mean"], mean),
        "min_mismatched_cells": int(np.count_nonzero(
            ~((got["min"] == mn) | (np.isnan(got["min"]) & np.isnan(mn))))),
        "max_mismatched_cells": int(np.count_nonzero(
            ~((got["max"] == mx) | (np.isnan(got["max"]) & np.isnan(mx))))),
        "var": _max_rel(got["var"], var),
        "std": _max_rel(got["std"], np.sqrt(var)),
    }
    print("accuracy (max per-bin relative error vs float64): " +
          ", ".join(f"{k} {val:.3e}" if isinstance(val, float) else f"{k} {val}"
                    for k, val in err.items()), flush=True)
    print(f"accuracy: busiest cell {int(count.max())} rows, non-empty cells "
          f"{int(np.count_nonzero(count))}", flush=True)
    assert err["count_mismatched_cells"] == 0
    assert err["min_mismatched_cells"] == 0 and err["max_mismatched_cells"] == 0
    assert err["sum"] <= 1e-6 and err["mean"] <= 1e-6
    assert err["var"] <= 1e-5 and err["std"] <= 1e-5

"""GPU parity of tf.shade (the dsx_shade kernels) against the reference's expected images
and the numpy oracle (oracle/shade.py).

Tolerances: every golden case exactly, except the ones the reference itself compares
within 1 per channel.  Random planes: linear and eq_hist exactly (pure IEEE arithmetic
in both); log / cbrt within 1 per channel (log1p / pow of the GPU's math library vs the
host's, a last-bit difference can move a truncated channel by one); float categorical
mixes within 1 per channel (float32 sums in another order than numpy's BLAS matmul).
"""
from __future__ import annotations

import zlib

import numpy as np
import pytest

import datashader_amd as dsa
import datashader_amd.transfer_functions as tf
from golden_utils import channels_close, shade_case, shade_cases, shade_oracle

pytestmark = pytest.mark.gpu

CASES = shade_cases()


def _dataarray(data, device):
    import torch
    dims = ["y", "x"] + (["cat"] if data.ndim == 3 else [])
    coords = {"x": np.arange(data.shape[1]), "y": np.arange(data.shape[0])}
    if data.ndim == 3:
        coords["cat"] = [f"c{i}" for i in range(data.shape[2])]
    if device:
        if data.dtype == np.uint32:
            t = torch.from_numpy(data.view(np.int32)).cuda().view(torch.uint32)
        else:
            t = torch.from_numpy(np.ascontiguousarray(data)).cuda()
        return dsa.DataArray(t, coords=coords, dims=dims)
    return dsa.DataArray(data, coords=coords, dims=dims)


def _img(img):
    return np.asarray(img.values, dtype=np.uint32)


@pytest.mark.parametrize("device", [False, True], ids=["host-plane", "device-plane"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden(case, device):
    data, kw, expected = shade_case(case)
    img = tf.shade(_dataarray(data, device), **kw)
    got = _img(img)
    if case["tol"]:
        assert channels_close(got, expected, case["tol"]), (got, expected)
    else:
        np.testing.assert_array_equal(got, expected)
    if device:
        assert img.data.is_cuda  # the image stays on the device


def _plane(rng, dtype, shape, kind):
    """Random plane: counts with zeros, floats with NaNs, signed ints, booleans."""
    if dtype == "bool":
        return rng.random(shape) < 0.6
    if dtype == "u4":
        hi = 200 if kind == "small" else 3_000_000
        a = rng.integers(0, hi, size=shape).astype("u4")
        a[rng.random(shape) < 0.3] = 0
        return a
    if dtype in ("i4", "i8"):
        hi = 150 if kind == "small" else 2_000_000
        return rng.integers(-hi, hi, size=shape).astype(dtype)
    a = rng.lognormal(0, 2, size=shape).astype(dtype)
    a[rng.random(shape) < 0.25] = np.nan
    return a


DTYPES = ["u4", "i4", "i8", "f4", "f8", "bool"]
HOWS = ["linear", "log", "cbrt", "eq_hist"]


def _tol(how, data):
    return 1 if how in ("log", "cbrt") else 0


@pytest.mark.parametrize("shape", [(61, 97), (384, 512)], ids=["61x97", "384x512"])
@pytest.mark.parametrize("kind", ["small", "wide"])
@pytest.mark.parametrize("how", HOWS)
@pytest.mark.parametrize("dtype", DTYPES)
def test_random_cmap_vs_oracle(dtype, how, kind, shape):
    rng = np.random.default_rng(zlib.crc32(repr((dtype, how, kind, shape)).encode()))
    data = _plane(rng, dtype, shape, kind)
    for kw in ({"cmap": ["lightblue", "darkblue"]},
               {"cmap": ["#000000", (0, 255, 0), "orange", "#FF00FF", "white"]},
               {"cmap": "red", "min_alpha": 30},
               {"cmap": ["pink", "red"], "rescale_discrete_levels": True}):
        kw = dict(kw, how=how)
        want = shade_oracle(data, kw)
        got = _img(tf.shade(_dataarray(data, True), **kw))
        tol = _tol(how, data)
        if tol:
            assert channels_close(got, want, tol), (dtype, how, kw)
        else:
            bad = np.argwhere(got != want)
            assert bad.size == 0, (dtype, how, kw, bad[:5], got.ravel()[:4], want.ravel()[:4])


@pytest.mark.parametrize("how", ["linear", "log", "cbrt"])
@pytest.mark.parametrize("dtype", ["u4", "i4", "f4", "f8", "bool"])
def test_random_span_vs_oracle(dtype, how):
    rng = np.random.default_rng(7)
    data = _plane(rng, dtype, (128, 96), "small")
    finite = data[~np.isnan(data.astype(float))].astype(float) if dtype != "bool" else None
    if dtype == "bool":
        spans = ([0, 1], [0.5, 2.5])
    else:
        q = np.quantile(finite, [0.1, 0.2, 0.8, 0.9])
        spans = ([float(q[0]), float(q[3])], [int(q[1]), int(q[2]) + 1],
                 [float(q[1]), float(finite.max()) * 4])
    for span in spans:
        for cmap in (["pink", "red"], "black"):
            kw = {"cmap": cmap, "how": how, "span": span}
            want = shade_oracle(data, kw)
            got = _img(tf.shade(_dataarray(data, True), **kw))
            assert channels_close(got, want, _tol(how, data)), (dtype, how, span, cmap)


@pytest.mark.parametrize("how", HOWS)
@pytest.mark.parametrize("dtype", ["u4", "i4", "i8", "f4", "f8"])
def test_random_categorical_vs_oracle(dtype, how):
    rng = np.random.default_rng(11)
    data = _plane(rng, dtype, (70, 90, 6), "small")
    if dtype == "u4":
        data[rng.random(data.shape[:2]) < 0.3] = 0  # empty pixels
    if dtype[0] == "f":
        data[rng.random(data.shape[:2]) < 0.1] = np.nan  # all-NaN pixels
    colours = ["red", "#00FF00", (0, 0, 255), "orange", "purple", "teal"]
    tol = 0 if dtype[0] in "ui" and how in ("linear", "eq_hist") else 1
    for kw in ({}, {"min_alpha": 0, "alpha": 200}, {"rescale_discrete_levels": True},
               {"color_baseline": 0 if dtype[0] in "ui" else 0.5}):
        kw = dict(kw, how=how, color_key=colours)
        want = shade_oracle(data, kw)
        got = _img(tf.shade(_dataarray(data, True), **kw))
        assert channels_close(got, want, tol), (dtype, how, kw)
    if how != "eq_hist":
        lo, hi = (2, 40) if dtype[0] in "ui" else (0.5, 40.0)
        kw = {"how": how, "color_key": colours, "span": [lo, hi]}
        want = shade_oracle(data, kw)
        got = _img(tf.shade(_dataarray(data, True), **kw))
        assert channels_close(got, want, tol), (dtype, how, kw)


def test_histogram_path_of_wide_counts_is_exact():
    """Counts spanning more than 65536 values take np.histogram's float path (65536 bins,
    numpy's edge corrections): bit-exact against the oracle."""
    rng = np.random.default_rng(3)
    counts = (rng.pareto(1.2, size=(300, 400)) * 1000).astype("u4")
    counts[rng.random(counts.shape) < 0.4] = 0
    assert np.ptp(counts[counts > 0]) >= 65536
    kw = {"cmap": ["lightblue", "darkblue"], "how": "eq_hist"}
    np.testing.assert_array_equal(_img(tf.shade(_dataarray(counts, True), **kw)),
                                  shade_oracle(counts, kw))


@pytest.mark.parametrize("data", [np.zeros((3, 4), "u4"), np.full((3, 4), np.nan),
                                  np.zeros((2, 2, 2), "u4"), np.full((2, 2, 2), np.nan)],
                         ids=["zeros2d", "nan2d", "zeros3d", "nan3d"])
def test_all_masked(data):
    """test_shade_all_masked (test_transfer_functions.py:569-590) and
    test_shade_should_handle_zeros_array (1170-1178)."""
    img = tf.shade(_dataarray(data, True), how="eq_hist", cmap=["white", "white"])
    assert img.shape == data.shape[:2]
    np.testing.assert_array_equal(_img(img), shade_oracle(data, {"how": "eq_hist",
                                                                 "cmap": ["white", "white"]}))


def test_canvas_then_shade_stays_on_device():
    """Aggregate on the device, shade on the device; same image as the oracle's shade of
    the downloaded aggregate."""
    from datashader_amd import synthetic
    n = 2_000_000
    x, y, v = synthetic.points_skewed_dev(n, seed=5)
    frame = dsa.DeviceFrame({"x": x, "y": y, "v": v})
    cvs = dsa.Canvas(300, 200, x_range=(0, 1), y_range=(0, 1))
    agg = cvs.points(frame, "x", "y", dsa.count())
    img = tf.shade(agg)
    assert img.data.is_cuda and img.dims == agg.dims
    np.testing.assert_array_equal(_img(img), shade_oracle(agg.values, {}))
    mean = cvs.points(frame, "x", "y", dsa.mean("v"))
    for how in ("linear", "eq_hist"):
        np.testing.assert_array_equal(_img(tf.shade(mean, how=how)),
                                      shade_oracle(mean.values, {"how": how}))


def test_shade_category_mixed_nan():
    """test_shade_category_mixed_nan (test_transfer_functions.py:499-517)."""
    import pandas as pd
    df = pd.DataFrame({"x": [0.25, 0.25, 0.75, 0.75, 0.75], "y": [0.5] * 5,
                       "cat": pd.Categorical(["a", "b", "a", "b", "c"]),
                       "val": [1.0, 2.0, 1.0, 2.0, 3.0]})
    cvs = dsa.Canvas(plot_width=2, plot_height=1, x_range=(0, 1), y_range=(0, 1))
    agg = cvs.points(df, "x", "y", agg=dsa.by("cat", dsa.mean("val")))
    assert np.isnan(agg.values[0, 0, 2])
    assert not np.any(np.isnan(agg.values[0, 1, :]))
    img = tf.shade(agg, color_key={"a": "red", "b": "green", "c": "blue"})
    alpha = _img(img).view(np.uint8).reshape(img.shape + (4,))[..., 3]
    assert alpha[0, 0] > 0 and alpha[0, 1] > 0
    want = shade_oracle(agg.values, {"color_key": {"a": "red", "b": "green", "c": "blue"}})
    np.testing.assert_array_equal(_img(img), want)

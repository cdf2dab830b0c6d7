"""ChunkedFrame host logic (no GPU): validation and chunk iteration."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from datashader_amd import ChunkedFrame


def _df(n, seed=0):
    rng = np.random.default_rng(seed)
    return pd.DataFrame({"x": rng.normal(size=n).astype("f4"), "y": rng.normal(size=n),
                         "c": pd.Categorical(rng.choice(list("abc"), n), categories=list("abc"))})


def test_chunks_cover_every_row_in_order():
    parts = [_df(1000, 1), _df(2500, 2), _df(7, 3)]
    cf = ChunkedFrame(parts, chunk_rows=600)
    got = [c for c, head, ps in cf.chunks(["x", "c"])]
    assert all(head is None and ps for _, head, ps in cf.chunks(["x"]))
    assert [len(c["x"]) for c in got] == [600, 400, 600, 600, 600, 600, 100, 7]
    np.testing.assert_array_equal(np.concatenate([c["x"] for c in got]),
                                  np.concatenate([p["x"].to_numpy() for p in parts]))
    codes = np.concatenate([c["c"] for c in got])
    np.testing.assert_array_equal(codes, np.concatenate([p["c"].cat.codes for p in parts]))
    assert len(cf) == 3507 and cf.schema["c"] == "category"


def test_validation():
    with pytest.raises(ValueError, match="same columns"):
        ChunkedFrame([_df(10), _df(10).rename(columns={"y": "z"})])
    other = _df(10)
    other["c"] = pd.Categorical(other["c"].astype(str), categories=list("abcd"))
    with pytest.raises(ValueError, match="categories"):
        ChunkedFrame([_df(10), other])
    with pytest.raises(ValueError, match="chunk_rows"):
        ChunkedFrame(_df(10), chunk_rows=0)
    with pytest.raises(ValueError, match="DataFrame"):
        ChunkedFrame([])


def _drawn(cf, stitch):
    """The (rows, plot_start) each uploaded chunk of an axis-0 line holds."""
    out = []
    for cols, head, ps in cf.chunks(["x"], stitch):
        rows = cols["x"] if head is None else np.concatenate([head["x"], cols["x"]])
        out.append((rows.tolist(), ps))
    return out


def test_axis0_chunks_carry_the_rows_before_them():
    """Within a partition a chunk is uploaded behind the two rows before it, the first as
    context (plot_start DSX_PLOT_CONTEXT), or behind the one row before it when that row
    opens the line; LineAxis0 partitions after the first start with the previous
    partition's last row and plot_start False (dask.py:242-248); LinesAxis0Multi
    partitions are independent."""
    from datashader_amd import _abi
    C = _abi.PLOT_CONTEXT
    p0 = pd.DataFrame({"x": np.arange(0.0, 5.0)})
    p1 = pd.DataFrame({"x": np.arange(10.0, 13.0)})
    cf = ChunkedFrame([p0, p1], chunk_rows=2)
    assert _drawn(cf, "axis0") == [
        ([0.0, 1.0], True), ([0.0, 1.0, 2.0, 3.0], C), ([2.0, 3.0, 4.0], C),
        ([4.0, 10.0, 11.0], False), ([10.0, 11.0, 12.0], C)]
    assert _drawn(cf, "multi") == [
        ([0.0, 1.0], True), ([0.0, 1.0, 2.0, 3.0], C), ([2.0, 3.0, 4.0], C),
        ([10.0, 11.0], True), ([10.0, 11.0, 12.0], C)]
    # one-row chunks: the second chunk's only row before it opens the line
    cf1 = ChunkedFrame(pd.DataFrame({"x": [1.0, 2.0, 3.0]}), chunk_rows=1)
    assert _drawn(cf1, "axis0") == [([1.0], True), ([1.0, 2.0], True), ([1.0, 2.0, 3.0], C)]
    # an empty partition hands on no row (dask concatenates an empty frame's last row),
    # and the next partition still starts with plot_start False
    cf2 = ChunkedFrame([p0.iloc[:2], p0.iloc[:0], p1], chunk_rows=5)
    assert _drawn(cf2, "axis0") == [([0.0, 1.0], True), ([10.0, 11.0, 12.0], False)]


def test_row_index_reductions_are_not_streamed():
    import datashader_amd as dsa
    df = pd.DataFrame({"x": np.arange(10.0), "y": np.arange(10.0), "v": np.arange(10.0)})
    for red in (dsa.first("v"), dsa.where(dsa.max("v")), dsa.max_n("v", n=2)):
        with pytest.raises(NotImplementedError, match="row-index"):
            dsa.Canvas(10, 10).points(ChunkedFrame(df, chunk_rows=3), "x", "y", red)

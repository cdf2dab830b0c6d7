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
    got = list(cf.chunks(["x", "c"]))
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

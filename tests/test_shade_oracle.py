"""Pin the shade oracle (oracle/shade.py) on the reference's own expected images and on
properties the mathematics fixes.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

from golden_utils import channels_close, shade_case, shade_cases, shade_oracle
from oracle import shade as osh

CASES = shade_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_expectation(case):
    data, kw, expected = shade_case(case)
    got = shade_oracle(data, kw)
    if case["tol"]:
        assert channels_close(got, expected, case["tol"])
    else:
        np.testing.assert_array_equal(got, expected)


def _cdf_slope(eq):
    """test_transfer_functions.py:1128-1137: the sorted equalised values rise linearly."""
    hist, _ = np.histogram(eq[~np.isnan(eq)], bins=256)
    cdf = np.cumsum(hist)
    cdf = cdf / cdf[-1]
    return np.polyfit(np.linspace(0, 1, cdf.size), cdf, 1)[0]


def test_eq_hist_flattens_float_and_integer_data():
    """test_eq_hist (1140-1153): float data with NaNs and integer data."""
    rng = np.random.default_rng(0)
    data = rng.normal(size=300 ** 2)
    data[rng.integers(300 ** 2, size=100)] = np.nan
    data = (data - np.nanmin(data)).reshape((300, 300))
    mask = np.isnan(data)
    eq, levels = osh.eq_hist(data, mask)
    assert 0.9 < _cdf_slope(eq) < 1.1
    assert (np.isnan(eq) == mask).all()
    assert levels > 1000
    ints = rng.normal(scale=100, size=(300, 300)).astype("i8")
    ints = ints - ints.min()
    eq, levels = osh.eq_hist(ints)
    assert 0.9 < _cdf_slope(eq) < 1.1
    assert levels == len(np.unique(ints))


def test_eq_hist_integer_path_is_the_empirical_cdf():
    """For integers the equalised value of v is #(values <= v) / n, exactly."""
    rng = np.random.default_rng(1)
    ints = rng.integers(0, 50, size=(40, 50)).astype("u4")
    eq, _ = osh.eq_hist(ints)
    flat = np.sort(ints.ravel())
    want = np.searchsorted(flat, ints, side="right") / flat.size
    np.testing.assert_array_equal(eq, want)


def test_linear_two_colours_is_a_blend():
    """With two colours and how='linear' every channel is the blend
    c0 + (c1 - c0) * (v - min) / (max - min), truncated (closed form)."""
    rng = np.random.default_rng(2)
    data = rng.uniform(1, 100, size=(30, 40))
    data[rng.random(data.shape) < 0.2] = np.nan
    c0, c1 = (200, 40, 10), (10, 220, 90)
    img = osh.interpolate(data, [c0, c1], how="linear")
    lo, hi = np.nanmin(data), np.nanmax(data)
    f = (data - lo) / (hi - lo)
    px = img.view(np.uint8).reshape(data.shape + (4,)).astype(int)
    ok = ~np.isnan(data)
    for ch in range(3):
        exact = c0[ch] + (c1[ch] - c0[ch]) * f
        assert np.all(np.abs(px[..., ch][ok] - np.trunc(exact[ok])) <= 1)
    assert np.all(px[..., 3][ok] == 255) and np.all(px[~ok] == 0)


def test_oracle_catches_a_planted_mask_mistake(monkeypatch):
    """Plant a plausible slip in the oracle (zeros of an unsigned plane left unmasked):
    the golden cases no longer hold, so they pin the masking rule."""
    monkeypatch.setattr(osh, "_mask_of", lambda d: np.isnan(d) if d.dtype.kind == "f"
                        else np.zeros(d.shape, dtype=bool))
    broken = 0
    for case in CASES:
        data, kw, expected = shade_case(case)
        if data.dtype.kind != "u" or data.ndim != 2:
            continue
        if not np.array_equal(shade_oracle(data, kw), expected):
            broken += 1
    assert broken > 10

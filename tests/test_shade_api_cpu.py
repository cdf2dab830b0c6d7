"""tf.shade argument validation (raised on the host, before any device work)."""
from __future__ import annotations

import numpy as np
import pytest

import datashader_amd as dsa
import datashader_amd.transfer_functions as tf
from datashader_amd.colors import rgb


def _agg2d():
    a = np.arange(10, 19, dtype="u4").reshape((3, 3))
    return dsa.DataArray(a, coords={"x_axis": [3, 4, 5], "y_axis": [0, 1, 2]},
                         dims=["y_axis", "x_axis"])


def test_not_a_dataarray():
    with pytest.raises(TypeError, match="agg must be instance of DataArray"):
        tf.shade(np.zeros((2, 2)))


@pytest.mark.parametrize("alpha,min_alpha", [(256, 40), (255, -1), (-3, 0)])
def test_alpha_range(alpha, min_alpha):
    with pytest.raises(ValueError, match="must be between 0 and 255"):
        tf.shade(_agg2d(), alpha=alpha, min_alpha=min_alpha)


def test_cmap_errors():
    """test_shade_cmap_errors (test_transfer_functions.py:295-300)."""
    with pytest.raises(ValueError):
        tf.shade(_agg2d(), cmap="foo")
    with pytest.raises(ValueError):
        tf.shade(_agg2d(), cmap=[])
    with pytest.raises(TypeError):
        tf.shade(_agg2d(), cmap=3)


def test_unknown_how_and_eq_hist_span():
    with pytest.raises(ValueError, match="Unknown interpolation method"):
        tf.shade(_agg2d(), how="sqrt")
    with pytest.raises(ValueError, match="span is not"):
        tf.shade(_agg2d(), how="eq_hist", span=[0, 10])


def test_callables_are_not_taken():
    with pytest.raises(NotImplementedError):
        tf.shade(_agg2d(), how=lambda x, m: x)


def test_dimension_errors():
    a = dsa.DataArray(np.zeros(4, dtype="u4"), dims=["x"])
    with pytest.raises(ValueError, match="2D or 3D"):
        tf.shade(a)


def test_categorical_colour_key_errors():
    data = np.zeros((2, 2, 3), dtype="u4")
    agg = dsa.DataArray(data, coords={"y": [0, 1], "x": [0, 1], "cat": ["a", "b", "c"]},
                        dims=["y", "x", "cat"])
    with pytest.raises(ValueError, match="Color key must be provided"):
        tf.shade(agg, color_key=None)
    with pytest.raises(ValueError, match="Insufficient colors"):
        tf.shade(agg, color_key=["red", "blue"])


def test_categorical_without_categories_is_blank():
    agg = dsa.DataArray(np.zeros((2, 3, 0), dtype="u4"),
                        coords={"y": [0, 1], "x": [0, 1, 2], "cat": []}, dims=["y", "x", "cat"])
    img = tf.shade(agg)
    assert img.shape == (2, 3) and not img.values.any()
    assert list(img.dims) == ["y", "x"] and list(img.coords) == ["x", "y"]


def test_rgb():
    """colors.py:99-127 behaviour."""
    assert rgb("plum") == (221, 160, 221)
    assert rgb("#FFFFFF") == (255, 255, 255)
    assert rgb((1, 2, 3)) == (1, 2, 3)
    assert rgb("sandybrown") == (250, 164, 96)
    for bad in ("#FFF", "#GGGGGG", "nocolour"):
        with pytest.raises(ValueError):
            rgb(bad)
    with pytest.raises(ValueError):
        rgb((0, 0, 300))
    with pytest.raises(TypeError):
        rgb([0, 0, 0])

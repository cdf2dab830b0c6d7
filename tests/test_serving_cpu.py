"""Host logic of the viewport server: the row intervals a viewport reads from the grid
index cover every row that can land in it (brute force over a synthetic index)."""
from __future__ import annotations

import numpy as np

from datashader_amd import serving


def _fake_index(x, y, grid, extent):
    idx = object.__new__(serving.PointIndex)
    (x0, x1), (y0, y1) = extent
    idx.grid, idx.n = grid, len(x)
    idx.x0, idx.x1, idx.y0, idx.y1 = x0, x1, y0, y1
    idx.sx, idx.sy = grid / (x1 - x0), grid / (y1 - y0)
    fx = np.clip((x - x0) * idx.sx, 0, grid - 1).astype(np.int64)
    fy = np.clip((y - y0) * idx.sy, 0, grid - 1).astype(np.int64)
    key = fy * grid + fx
    order = np.argsort(key, kind="stable")
    idx.cell_start = np.searchsorted(key[order], np.arange(grid * grid + 1))
    return idx, x[order], y[order]


def test_ranges_cover_every_row_in_the_view():
    rng = np.random.default_rng(0)
    x = rng.normal(0, 1, 200_000)
    y = rng.normal(0, 2, 200_000)
    extent = ((x.min(), x.max()), (y.min(), y.max()))
    idx, sx, sy = _fake_index(x, y, 32, extent)
    for _ in range(200):
        cx, cy = rng.normal(0, 1), rng.normal(0, 2)
        w, h = 10 ** rng.uniform(-3, 1), 10 ** rng.uniform(-3, 1)
        xr, yr = (cx - w, cx + w), (cy - h, cy + h)
        inside = (sx >= xr[0]) & (sx <= xr[1]) & (sy >= yr[0]) & (sy <= yr[1])
        covered = np.zeros(len(sx), bool)
        iv = idx.ranges(xr, yr)
        for s, e in iv:
            covered[s:e] = True
        assert not np.any(inside & ~covered)
        assert all(e > s for s, e in iv)
        assert all(iv[k][1] < iv[k + 1][0] for k in range(len(iv) - 1))  # merged, ordered
    # views outside the extent read nothing; the whole extent reads everything
    assert idx.ranges((x.max() + 1, x.max() + 2), (0, 1)) == []
    assert idx.ranges(extent[0], extent[1]) == [(0, len(x))]


def test_request_stream_shape():
    rng = np.random.default_rng(1)
    pts = (rng.normal(0, 1, 1000), rng.normal(0, 1, 1000))
    reqs = serving.make_requests(500, pts, ((-4, 4), (-4, 4)), seed=2, rate=200.0)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    assert abs(len(reqs) / arr[-1] - 200.0) < 30.0           # Poisson at 200 per second
    zooms = np.array([8.0 / (r.x_range[1] - r.x_range[0]) for r in reqs])
    assert zooms.min() >= 1.0 - 1e-9 and zooms.max() <= 2 ** 12 * (1 + 1e-9)
    assert {(r.width, r.height) for r in reqs} == {(256, 256), (512, 512), (1024, 768),
                                                    (1920, 1080)}

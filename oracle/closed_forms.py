"""Vectorised numpy closed forms of the points path (TEST INFRASTRUCTURE ONLY).

Independent of ``dsx_oracle.c``: they pin it on large random inputs and serve as the
float64 accuracy oracle for the GPU's float32 compensated sums.

* ``point_bins``   -- ``glyphs/points.py:199-203`` vectorised: numpy evaluates
  ``x*sx + tx`` as two separately rounded float64 operations (numpy never contracts to
  FMA), then truncates toward zero.
* ``count`` / ``sum`` -- ``np.bincount``.  numpy's bincount adds the weights of each bin
  sequentially in row order, so its float64 sum equals the reference's row-ordered
  ``agg[y, x] += field`` bit for bit (``reductions.py:958-962``).
* ``minmax``       -- ``np.fmin.at`` / ``np.fmax.at`` (order-free, exact).
* ``var2pass``     -- two-pass float64 variance per bin (the accuracy oracle for m2).
"""
from __future__ import annotations

import numpy as np


def point_bins(xs, ys, view, W, H):
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    with np.errstate(invalid="ignore"):
        keep = (view.xmin <= xs) & (xs <= view.xmax) & (view.ymin <= ys) & (ys <= view.ymax)
    xx = np.zeros(len(xs), np.int64)
    yy = np.zeros(len(xs), np.int64)
    xx[keep] = np.trunc(xs[keep] * view.sx + view.tx).astype(np.int64)
    yy[keep] = np.trunc(ys[keep] * view.sy + view.ty).astype(np.int64)
    xx = np.where(xx >= W, W - 1, xx)
    yy = np.where(yy >= H, H - 1, yy)
    xx = np.where(xx < 0, xx + W, xx)
    yy = np.where(yy < 0, yy + H, yy)
    return np.where(keep, yy * W + xx, -1)


def _valid(bins, val=None):
    m = bins >= 0
    if val is not None:
        m &= ~np.isnan(np.asarray(val, np.float64))
    return m


def count(bins, W, H, val=None):
    m = _valid(bins, val)
    return np.bincount(bins[m], minlength=W * H).astype(np.uint32).reshape(H, W)


def sum_zero(bins, val, W, H):
    m = _valid(bins, val)
    w = np.asarray(val, np.float64)[m]
    return np.bincount(bins[m], weights=w, minlength=W * H).reshape(H, W)


def sum_nan(bins, val, W, H):
    s = sum_zero(bins, val, W, H)
    c = count(bins, W, H, val)
    return np.where(c > 0, s, np.nan)


def minmax(bins, val, W, H):
    m = _valid(bins, val)
    v = np.asarray(val, np.float64)[m]
    mn = np.full(W * H, np.inf)
    mx = np.full(W * H, -np.inf)
    np.minimum.at(mn, bins[m], v)
    np.maximum.at(mx, bins[m], v)
    c = count(bins, W, H, val).ravel()
    mn = np.where(c > 0, mn, np.nan).reshape(H, W)
    mx = np.where(c > 0, mx, np.nan).reshape(H, W)
    return mn, mx


def mean(bins, val, W, H):
    s = sum_zero(bins, val, W, H)
    c = count(bins, W, H, val)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(c > 0, s / c, np.nan)


def var2pass(bins, val, W, H):
    """Population variance per bin by the two-pass float64 formula."""
    m = _valid(bins, val)
    b = bins[m]
    v = np.asarray(val, np.float64)[m]
    c = np.bincount(b, minlength=W * H).astype(np.float64)
    s = np.bincount(b, weights=v, minlength=W * H)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = s / c
    d = v - mu[b]
    m2 = np.bincount(b, weights=d * d, minlength=W * H)
    with np.errstate(divide="ignore", invalid="ignore"):
        var = np.where(c > 0, m2 / c, np.nan)
    return var.reshape(H, W)

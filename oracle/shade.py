"""CPU oracle for ``tf.shade`` (TEST INFRASTRUCTURE ONLY).

A numpy restatement of datashader 0.19.1's colouring of aggregates
(``transfer_functions/__init__.py``), used by ``tests/`` as the checker of the
``dsx_shade`` kernels.  The product package never imports it.

* :func:`eq_hist`       -- ``eq_hist`` (148-223): exact value counts for integer data
  whose range is below ``nbins``, otherwise ``np.histogram`` with the empty bins
  removed; cdf = cumsum / total; ``np.interp`` of the data onto it.
* :func:`interpolate`   -- ``_interpolate`` (251-357) for a list colormap or one colour.
* :func:`colorize`      -- ``_colorize`` (360-463) + ``_interpolate_alpha`` (466-538).
* :func:`discrete`      -- ``_apply_discrete_colorkey`` (541-613).

Colours come in as (r, g, b) triples (the names are resolved by the caller).  The
arithmetic is numpy's, in the dtypes the reference's expressions produce: the offset is
subtracted in the plane's dtype, float32 planes transform in float32, the spans of
float32 values give float32 knots.

Pinning: every expected image of the reference's ``tests/test_transfer_functions.py``
for these functions (transcribed in ``tests/golden/shade_cases.json``); ``eq_hist``'s
cdf-slope property (``test_eq_hist``); closed forms for the linear case (a pixel's
channels are the linear blend of its two neighbouring colours).
"""
from __future__ import annotations

import numpy as np

_HOWS = ("linear", "log", "cbrt", "eq_hist")


def _mask_of(data):
    """The missing-value mask of a plane (transfer_functions:268-275)."""
    if data.dtype == np.bool_:
        return ~data
    if data.dtype.kind == "u":
        return data == 0
    if data.dtype.kind == "f":
        return np.isnan(data)
    return np.zeros(data.shape, dtype=bool)


def eq_hist(data, mask=None, nbins=256 * 256):
    """(equalised values, number of discrete levels); NaN where ``mask``."""
    if mask is not None and np.all(mask):
        return np.full_like(data, np.nan, dtype=np.float64), 0
    valid = data if mask is None else data[~mask]
    if valid.dtype == bool or (np.issubdtype(valid.dtype, np.integer) and np.ptp(valid) < nbins):
        values, counts = np.unique(valid, return_counts=True)
        lo, hi = values[0].item(), values[-1].item()
        centers = np.arange(lo, hi + 1)
        hist = np.zeros(hi - lo + 1, dtype="uint64")
        hist[values - lo] = counts
        levels = len(values)
    else:
        hist, edges = np.histogram(valid, bins=nbins)
        centers = (edges[:-1] + edges[1:]) / 2
        nonempty = hist > 0
        levels = int(np.count_nonzero(nonempty))
        if levels != len(hist):
            hist = hist[nonempty]
            centers = centers[nonempty]
    cdf = hist.cumsum()
    cdf = cdf / float(cdf[-1])
    out = np.interp(data, centers, cdf).reshape(data.shape)
    if mask is not None:
        out = np.where(mask, np.nan, out)
    return out, levels


def _transform(how, d, m):
    """The interpolaters of transfer_functions:225-228; returns (values, levels)."""
    if how == "eq_hist":
        return eq_hist(d, m)
    blanked = np.where(m, np.nan, d)
    if how == "log":
        return np.log1p(blanked), None
    if how == "cbrt":
        return blanked ** (1 / 3.), None
    return blanked, None


def _rescale(levels, span):
    """_rescale_discrete_levels (transfer_functions:236-248)."""
    m = -0.5 / 98.0
    c = 1.5 - 2 * m
    multiple = m * levels + c
    if multiple > 1:
        return (max(span[1] - multiple * (span[1] - span[0]), 0), 1)
    return span


def _clip(data, mask, lo, hi):
    """masked_clip_2d (_cpu_utils.py): clip the unmasked values in place."""
    keep = ~mask
    vals = data[keep]
    wide = vals.astype(np.float64)  # numba compares in float64
    lo_v = np.array(lo).astype(data.dtype)
    hi_v = np.array(hi).astype(data.dtype)
    vals = np.where(wide < float(lo), lo_v, np.where(wide > float(hi), hi_v, vals))
    data[keep] = vals


def _pack(r, g, b, a):
    rgba = np.dstack([r.astype(np.uint8), g.astype(np.uint8), b.astype(np.uint8),
                      a.astype(np.uint8)])
    return rgba.view(np.uint32).reshape(r.shape)


def interpolate(data, colors, how="eq_hist", alpha=255, span=None, min_alpha=40,
                rescale_discrete_levels=False, single=False):
    """RGBA image of a 2D plane: ``colors`` is a list of (r, g, b) (``single=False``) or
    one (r, g, b) whose alpha follows the value (``single=True``)."""
    data = np.array(data, copy=True)
    mask = _mask_of(data)
    if data.dtype == np.bool_:
        data = data.astype(np.int8)
    if mask.all():
        return np.zeros(data.shape, dtype=np.uint32)
    if span is None:
        offset = np.nanmin(data[~mask])
    else:
        offset = np.array(span, dtype=data.dtype)[0]
        _clip(data, mask, span[0], span[1])
    data -= offset
    with np.errstate(invalid="ignore", divide="ignore"):
        t, levels = _transform(how, data, mask)
        if span is None:
            blanked = np.where(~mask, t, np.nan)
            rng = np.nanmin(blanked), np.nanmax(blanked)
            if rescale_discrete_levels and levels is not None:
                rng = _rescale(levels, rng)
        else:
            rng, _ = _transform(how, np.array([0, span[1] - span[0]]), np.zeros(2, bool))
    if not single:
        chans = np.array(colors).T  # 3 x n
        knots = np.linspace(rng[0], rng[1], len(colors))
        rgb = [np.nan_to_num(np.interp(t, knots, ch, left=255)).astype(np.uint8)
               for ch in chans]
        a = np.where(np.isnan(t), 0, alpha).astype(np.uint8)
        return _pack(rgb[0], rgb[1], rgb[2], a)
    steps = np.arange(min_alpha, alpha + 1)
    knots = np.linspace(rng[0], rng[1], len(steps))
    a = np.nan_to_num(np.interp(t, knots, steps, left=0, right=255)).astype(np.uint8)
    full = [np.full(data.shape, c, dtype=np.uint8) for c in colors]
    return _pack(full[0], full[1], full[2], a)


def _alpha(data, total, mask, how, alpha, span, min_alpha, rescale_discrete_levels):
    """_interpolate_alpha (transfer_functions:466-538)."""
    if span is None:
        offset = np.nanmin(total)
        if total.dtype.kind == "u" and offset == 0:
            mask = mask | (total == 0)
            if not np.all(mask):
                offset = total[total > 0].min()
            total = np.where(~mask, total, np.nan)
        scaled, levels = _transform(how, total - offset, mask)
        rng = [np.nanmin(scaled).item(), np.nanmax(scaled).item()]
        if rescale_discrete_levels and levels is not None:
            rng = _rescale(levels, rng)
    else:
        offset = np.array(span, dtype=data.dtype)[0]
        if total.dtype.kind == "u" and np.nanmin(total) == 0:
            mask = mask | (total <= 0)
            total = np.where(~mask, total, np.nan)
        _clip(total, mask, span[0], span[1])
        scaled, _ = _transform(how, total - offset, mask)
        rng, _ = _transform(how, np.array([0, span[1] - span[0]]), np.zeros(2, bool))
    a = np.interp(scaled, np.hstack(rng), np.array([min_alpha, alpha]), left=0, right=255)
    return np.nan_to_num(a).astype(np.uint8)


def colorize(data, colors, how="eq_hist", alpha=255, span=None, min_alpha=40,
             color_baseline=None, rescale_discrete_levels=False):
    """RGBA image of an [H, W, C] plane with one (r, g, b) per category."""
    data = np.asarray(data)
    cd = np.array(data, copy=True)
    present = ~np.isnan(data)
    with np.errstate(invalid="ignore"):
        base = np.nanmin(cd) if color_baseline is None else color_baseline
        if base > 0:
            np.subtract(cd, base, out=cd, where=present)
        else:
            np.add(cd, -base, out=cd, where=present)
    if color_baseline is not None and cd.dtype.kind != "u":
        np.maximum(cd, 0, out=cd)
    cd = np.nan_to_num(cd.astype(np.float32))
    weight = cd.sum(axis=2)
    palette = np.array(colors, dtype=np.float32)  # C x 3
    mixed = cd @ palette
    present_f = present.astype(np.float32)
    plain = present_f @ palette
    with np.errstate(divide="ignore", invalid="ignore"):
        rgb = (mixed / weight[..., None]).astype(np.uint8)
        rgb_plain = (plain / present_f.sum(axis=2)[..., None]).astype(np.uint8)
    rgb = np.where((weight == 0)[..., None], rgb_plain, rgb)
    missing = np.isnan(data)
    every_missing = missing.all(axis=2)
    total = np.where(missing & ~every_missing[..., None], 0, data).sum(axis=2)
    with np.errstate(invalid="ignore", divide="ignore"):
        a = _alpha(data, total, np.isnan(total), how, alpha, span, min_alpha,
                   rescale_discrete_levels)
    return _pack(rgb[..., 0], rgb[..., 1], rgb[..., 2], a)


def discrete(data, keys, colors, alpha=255, color_baseline=None):
    """RGBA image of a 2D plane of category values, one (r, g, b) per key."""
    data = np.asarray(data)
    hit = np.full(data.shape, -1)
    for i, k in enumerate(keys):
        hit[data == k] = i
    matched = hit >= 0
    value = 1.0
    if color_baseline is not None:
        if color_baseline > 0:
            value = 1.0 - color_baseline
        elif color_baseline < 0:
            value = 1.0 + (-color_baseline)
        value = max(value, 0.0)
    black = color_baseline is not None and value != 0 and matched.any()
    pal = np.array(colors, dtype=np.uint8).reshape(-1, 3)
    rgb = np.zeros(data.shape + (3,), dtype=np.uint8)
    if not black:
        rgb[matched] = pal[hit[matched]]
    a = np.where(matched, alpha, 0).astype(np.uint8)
    return _pack(rgb[..., 0], rgb[..., 1], rgb[..., 2], a)

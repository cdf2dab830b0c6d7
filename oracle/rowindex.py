"""Row-index reductions restated for the points path (TEST INFRASTRUCTURE ONLY).

``select_rows`` walks the rows in order and applies the reference's CPU appends one row
at a time (reductions.py): ``first`` keeps the first non-NaN value it sees
(``first._append`` 1398-1403), ``last`` the last (1436-1440), ``_min_row_index`` /
``_max_row_index`` the smallest / largest row (2319-2325, 2261-2266), ``max`` / ``min``
replace the kept row only on a strictly larger / smaller value (``max._append``
1223-1227, ``min._append`` 1178-1182), so ties keep the earlier row; ``where`` looks up
another column at the kept row (1914-1925).  ``selected_rows_closed_form`` is an
independent vectorised statement of the same result (np.minimum.at / np.maximum.at /
lexsort) that the tests use to pin the loop on large random inputs.
"""
from __future__ import annotations

import numpy as np


def cells(bins, codes=None, ncat=1):
    """Cell of each row (-1: dropped): bin * ncat + category, with a negative code
    wrapping once (numba's negative indexing) and out-of-range categories dropped."""
    bins = np.asarray(bins, np.int64)
    if codes is None:
        return bins
    c = np.asarray(codes, np.int64)
    c = np.where(c < 0, c + ncat, c)
    ok = (bins >= 0) & (c >= 0) & (c < ncat)
    return np.where(ok, bins * ncat + c, -1)


def select_rows(cell, ncells, mode, val=None):
    """int64 row index per cell (-1 where none) for mode first / last / max / min;
    ``val`` given: rows whose value is NaN are skipped."""
    idx = np.full(ncells, -1, np.int64)
    best = np.full(ncells, np.nan)
    v = None if val is None else np.asarray(val, np.float64)
    for i, b in enumerate(np.asarray(cell, np.int64)):
        if b < 0:
            continue
        if v is not None and v[i] != v[i]:
            continue
        if mode == "first":
            if idx[b] == -1:
                idx[b] = i
        elif mode == "last":
            idx[b] = i
        elif mode == "max":
            if best[b] != best[b] or best[b] < v[i]:
                best[b] = v[i]
                idx[b] = i
        elif mode == "min":
            if best[b] != best[b] or best[b] > v[i]:
                best[b] = v[i]
                idx[b] = i
        else:
            raise ValueError(mode)
    return idx


def selected_rows_closed_form(cell, ncells, mode, val=None):
    cell = np.asarray(cell, np.int64)
    rows = np.arange(len(cell), dtype=np.int64)
    keep = cell >= 0
    if val is not None:
        keep &= ~np.isnan(np.asarray(val, np.float64))
    c, r = cell[keep], rows[keep]
    idx = np.full(ncells, -1, np.int64)
    if mode in ("first", "last"):
        if mode == "first":
            big = np.full(ncells, np.iinfo(np.int64).max)
            np.minimum.at(big, c, r)
            idx = np.where(big == np.iinfo(np.int64).max, -1, big)
        else:
            np.maximum.at(idx, c, r)
        return idx
    v = np.asarray(val, np.float64)[keep]
    # order rows by (cell, value best first, row); the first of each cell is selected
    key = -v if mode == "max" else v
    order = np.lexsort((r, key, c))
    cs = c[order]
    head = np.ones(len(cs), bool)
    head[1:] = cs[1:] != cs[:-1]
    idx[cs[head]] = r[order][head]
    return idx


def take(index, col):
    """float64 lookup of ``col`` at ``index``; NaN where index is -1."""
    index = np.asarray(index, np.int64)
    col = np.asarray(col)
    out = np.full(index.shape, np.nan)
    ok = index >= 0
    out[ok] = col[index[ok]].astype(np.float64)
    return out


def _shift_insert(arr, value, k):
    """shift_and_insert (utils.py:672): arr[k+1:] = arr[k:-1]; arr[k] = value."""
    arr[k + 1:] = arr[k:-1].copy()
    arr[k] = value


def select_n(cell, ncells, n, mode, val=None):
    """[ncells, n] rows of each cell's n-slot list, -1 padded, applying the reference's
    appends row by row: first_n fills the first empty slot while one is left
    (reductions.py:1552-1566), last_n inserts at the front (1608-1613), max_n / min_n
    insert before the first slot holding a strictly smaller / larger value (1649-1659,
    1739-1749); the row-index forms are first / last (2410-2421, 2493-2503)."""
    rows = np.full((ncells, n), -1, np.int64)
    vals = np.full((ncells, n), np.nan)
    v = None if val is None else np.asarray(val, np.float64)
    for i, b in enumerate(np.asarray(cell, np.int64)):
        if b < 0 or (v is not None and v[i] != v[i]):
            continue
        r = rows[b]
        if mode == "first":
            if r[n - 1] != -1:
                continue
            r[int(np.argmax(r == -1))] = i
        elif mode == "last":
            _shift_insert(r, i, 0)
        else:
            for k in range(n):
                if r[k] == -1 or (v[i] > vals[b, k] if mode == "max" else v[i] < vals[b, k]):
                    _shift_insert(r, i, k)
                    _shift_insert(vals[b], v[i], k)
                    break
    return rows


def select_n_closed_form(cell, ncells, n, mode, val=None):
    """The same lists by sorting each cell's rows (lexsort) and keeping the first n."""
    cell = np.asarray(cell, np.int64)
    rows = np.arange(len(cell), dtype=np.int64)
    keep = cell >= 0
    if val is not None:
        keep &= ~np.isnan(np.asarray(val, np.float64))
    c, r = cell[keep], rows[keep]
    if mode == "first":
        key = r
    elif mode == "last":
        key = -r
    else:
        v = np.asarray(val, np.float64)[keep]
        key = -v if mode == "max" else v
    order = np.lexsort((r, key, c))
    cs, rs = c[order], r[order]
    start = np.ones(len(cs), bool)
    start[1:] = cs[1:] != cs[:-1]
    first_pos = np.maximum.accumulate(np.where(start, np.arange(len(cs)), 0))
    rank = np.arange(len(cs)) - first_pos
    out = np.full((ncells, n), -1, np.int64)
    sel = rank < n
    out[cs[sel], rank[sel]] = rs[sel]
    return out

"""CPU oracle for the points / line aggregation path (TEST INFRASTRUCTURE ONLY).

This package is the checker the HIP path is measured against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it;
the product package ``datashader_amd`` never imports, links or calls anything here.

Two independent restatements of the reference live here:

* ``libdsx_oracle.so`` (``dsx_oracle.c``): a sequential, row-ordered C restatement of
  datashader 0.19.1's numba kernels (``glyphs/points.py:188-212``,
  ``glyphs/line.py:688-1097,1099-1136,1244-1386``, ``reductions.py`` appends,
  ``compiler.py:321-475`` append order).  Wrapped by :func:`points`,
  :func:`line_axis0`, :func:`lines_xconst`, :func:`lines_axis1`, :func:`draw_segment`
  and :func:`bounds`.
* :mod:`oracle.closed_forms`: vectorised numpy closed forms (``np.bincount``,
  ``np.fmin.at`` ...) used to pin the C restatement on large random inputs.

Pinning (what each function is checked against, in ``tests/test_oracle.py``):
  points + count/any/sum/min/max/mean/var/std/by   reference test expectations
      (``tests/test_pandas.py`` fixtures transcribed in ``tests/golden``), numpy closed
      forms, brute force on tiny inputs.
  line drawing                                      reference expectations
      (``tests/test_glyphs.py``, ``tests/test_pandas.py`` line tests) and the touched-
      pixel masks of the reference's ``tests/data/test_00*_noaa_*.png`` images.
  bounds                                            ``np.nanmin`` / ``np.nanmax``.
Parity unpinned (properties of numba's generated code that no reference fixture fixes):
  FMA contraction of ``x*sx+tx`` (assumed absent), float32 segment deltas in
  ``_liang_barsky`` for float32 coordinates that are clipped, ``round`` ties in
  ``map_onto_pixel_snap`` (assumed half-to-even), negative category codes (assumed to
  wrap to the last category).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdsx_oracle.so")

BASES = ("count", "countv", "any", "anyv", "sum", "sumz", "min", "max", "m2")


class _Aggs(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int32), ("H", ctypes.c_int32), ("ncat", ctypes.c_int32),
        ("val", ctypes.c_void_p), ("cat", ctypes.c_void_p),
        ("count", ctypes.c_void_p), ("countv", ctypes.c_void_p),
        ("any", ctypes.c_void_p), ("anyv", ctypes.c_void_p),
        ("sum", ctypes.c_void_p), ("sumz", ctypes.c_void_p),
        ("min", ctypes.c_void_p), ("max", ctypes.c_void_p),
        ("m2", ctypes.c_void_p), ("m2_sum", ctypes.c_void_p), ("m2_cnt", ctypes.c_void_p),
    ]


class _View(ctypes.Structure):
    _fields_ = [
        ("sx", ctypes.c_double), ("tx", ctypes.c_double),
        ("sy", ctypes.c_double), ("ty", ctypes.c_double),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("xxmax", ctypes.c_int64), ("yymax", ctypes.c_int64),
        ("xf32", ctypes.c_int32), ("yf32", ctypes.c_int32),
    ]


class View:
    """Scale/translate and bounds of one aggregation call (``extend``'s ``vt`` and
    ``bounds`` arguments, ``glyphs/points.py:223``), plus the line snapping limits."""

    def __init__(self, sx, tx, sy, ty, xmin, xmax, ymin, ymax, xxmax=None, yymax=None,
                 xf32=False, yf32=False):
        self.sx, self.tx, self.sy, self.ty = float(sx), float(tx), float(sy), float(ty)
        self.xmin, self.xmax = float(xmin), float(xmax)
        self.ymin, self.ymax = float(ymin), float(ymax)
        # line.py:716-717: xxmax = round(xmax*sx + tx) (Python round: half to even)
        self.xxmax = int(round(self.xmax * self.sx + self.tx)) if xxmax is None else int(xxmax)
        self.yymax = int(round(self.ymax * self.sy + self.ty)) if yymax is None else int(yymax)
        self.xf32, self.yf32 = bool(xf32), bool(yf32)

    def _c(self):
        return _View(self.sx, self.tx, self.sy, self.ty, self.xmin, self.xmax, self.ymin,
                     self.ymax, self.xxmax, self.yymax, int(self.xf32), int(self.yf32))


_lib = None


def build():
    """Compile ``libdsx_oracle.so`` (and the timed CPU baseline ``libdsx_cpubase.so``)
    with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        lib.dsx_oracle_points.argtypes = [vp, vp, i64, P(_View), P(_Aggs)]
        lib.dsx_oracle_point_bins.argtypes = [vp, vp, i64, P(_View), i32, i32, vp]
        lib.dsx_oracle_line_axis0.argtypes = [vp, vp, i64, i32, P(_View), P(_Aggs)]
        lib.dsx_oracle_lines_xconst.argtypes = [vp, i64, vp, i64, P(_View), P(_Aggs)]
        lib.dsx_oracle_lines_axis1.argtypes = [vp, vp, i64, i64, P(_View), P(_Aggs)]
        lib.dsx_oracle_draw_segment.argtypes = [P(_View), P(_Aggs), i64, i32, d, d, d, d]
        lib.dsx_oracle_bounds.argtypes = [vp, i64, vp]
        for f in ("dsx_oracle_points", "dsx_oracle_point_bins", "dsx_oracle_line_axis0",
                  "dsx_oracle_lines_xconst", "dsx_oracle_lines_axis1",
                  "dsx_oracle_draw_segment", "dsx_oracle_bounds"):
            getattr(lib, f).restype = None
        _lib = lib
    return _lib


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    return None if a is None else a.ctypes.data


class Aggs:
    """Output arrays of one oracle call, initialised as ``reductions.py`` creates them
    (``_create_uint32`` zeros, ``_create_float64_nan`` NaN, ``_create_float64_zero`` 0,
    ``_create_bool`` False; ``reductions.py:446-472``)."""

    def __init__(self, W, H, bases, ncat=1, val=None, cat=None):
        bases = tuple(bases)
        for b in bases:
            if b not in BASES:
                raise ValueError(f"unknown base {b!r}")
        self.W, self.H, self.ncat = int(W), int(H), int(ncat)
        shape = (self.H, self.W, self.ncat)
        self.val = None if val is None else _f64(val)
        self.cat = None if cat is None else np.ascontiguousarray(np.asarray(cat, np.int64))
        self.out = {}
        for b in bases:
            if b in ("count", "countv"):
                self.out[b] = np.zeros(shape, np.uint32)
            elif b in ("any", "anyv"):
                self.out[b] = np.zeros(shape, np.uint8)
            elif b in ("sum", "min", "max"):
                self.out[b] = np.full(shape, np.nan, np.float64)
            else:
                self.out[b] = np.zeros(shape, np.float64)
        self._m2_sum = np.zeros(shape, np.float64) if "m2" in bases else None
        self._m2_cnt = np.zeros(shape, np.uint32) if "m2" in bases else None

    def _c(self):
        o = self.out
        return _Aggs(self.W, self.H, self.ncat, _ptr(self.val), _ptr(self.cat),
                     _ptr(o.get("count")), _ptr(o.get("countv")), _ptr(o.get("any")),
                     _ptr(o.get("anyv")), _ptr(o.get("sum")), _ptr(o.get("sumz")),
                     _ptr(o.get("min")), _ptr(o.get("max")), _ptr(o.get("m2")),
                     _ptr(self._m2_sum), _ptr(self._m2_cnt))

    def result(self, squeeze=True):
        res = {}
        for k, a in self.out.items():
            if k in ("any", "anyv"):
                a = a.astype(bool)
            res[k] = a[:, :, 0] if (squeeze and self.ncat == 1) else a
        return res


def _check_rows(n, val, cat):
    if val is not None and len(val) != n:
        raise ValueError("value column length differs from x/y")
    if cat is not None and len(cat) != n:
        raise ValueError("category column length differs from x/y")


def points(xs, ys, view, W, H, bases, val=None, cat=None, ncat=1, squeeze=True):
    """Sequential ``Point`` extend (``glyphs/points.py:188-212``) with the append of
    each base in ``bases``; returns ``{base: array[H, W(, ncat)]}``."""
    xs, ys = _f64(xs), _f64(ys)
    if xs.shape != ys.shape or xs.ndim != 1:
        raise ValueError("xs and ys must be 1-D of equal length")
    _check_rows(len(xs), val, cat)
    a = Aggs(W, H, bases, ncat, val, cat)
    _load().dsx_oracle_points(_ptr(xs), _ptr(ys), len(xs), ctypes.byref(view._c()),
                              ctypes.byref(a._c()))
    return a.result(squeeze)


def point_bins(xs, ys, view, W, H):
    """Flat bin index ``y*W + x`` of each row, -1 for dropped rows."""
    xs, ys = _f64(xs), _f64(ys)
    out = np.empty(len(xs), np.int64)
    _load().dsx_oracle_point_bins(_ptr(xs), _ptr(ys), len(xs), ctypes.byref(view._c()),
                                  int(W), int(H), _ptr(out))
    return out


def line_axis0(xs, ys, view, W, H, bases=("count",), plot_start=True, val=None, cat=None,
               ncat=1, squeeze=True):
    """``LineAxis0`` extend (``glyphs/line.py:1099-1136``): one line through all rows."""
    xs, ys = _f64(xs), _f64(ys)
    _check_rows(len(xs), val, cat)
    a = Aggs(W, H, bases, ncat, val, cat)
    _load().dsx_oracle_line_axis0(_ptr(xs), _ptr(ys), len(xs), int(bool(plot_start)),
                                  ctypes.byref(view._c()), ctypes.byref(a._c()))
    return a.result(squeeze)


def lines_xconst(xs, ys, view, W, H, bases=("count",), val=None, cat=None, ncat=1,
                 squeeze=True):
    """``LinesAxis1XConstant`` extend (``glyphs/line.py:1340-1386``): ``ys[i, :]`` is
    line ``i``, all sharing ``xs``."""
    xs = _f64(xs)
    ys = _f64(ys)
    if ys.ndim != 2 or ys.shape[1] != len(xs):
        raise ValueError("ys must be [nrows, len(xs)]")
    _check_rows(ys.shape[0], val, cat)
    a = Aggs(W, H, bases, ncat, val, cat)
    _load().dsx_oracle_lines_xconst(_ptr(xs), len(xs), _ptr(ys), ys.shape[0],
                                    ctypes.byref(view._c()), ctypes.byref(a._c()))
    return a.result(squeeze)


def lines_axis1(xs, ys, view, W, H, bases=("count",), val=None, cat=None, ncat=1,
                squeeze=True):
    """``LinesAxis1`` extend (``glyphs/line.py:1244-1336``): row ``i`` of ``xs``/``ys``
    is line ``i``."""
    xs, ys = _f64(xs), _f64(ys)
    if xs.shape != ys.shape or xs.ndim != 2:
        raise ValueError("xs and ys must be 2-D of equal shape")
    _check_rows(xs.shape[0], val, cat)
    a = Aggs(W, H, bases, ncat, val, cat)
    _load().dsx_oracle_lines_axis1(_ptr(xs), _ptr(ys), xs.shape[0], xs.shape[1],
                                   ctypes.byref(view._c()), ctypes.byref(a._c()))
    return a.result(squeeze)


def draw_segment(view, W, H, segment_start, x0, y0, x1, y1, i=0, bases=("count",)):
    """One ``draw_segment`` call (``glyphs/line.py:1045-1095``) into a fresh canvas."""
    a = Aggs(W, H, bases)
    _load().dsx_oracle_draw_segment(ctypes.byref(view._c()), ctypes.byref(a._c()), int(i),
                                    int(bool(segment_start)), float(x0), float(y0),
                                    float(x1), float(y1))
    return a.result()


def bounds(arr):
    """NaN-skipping (min, max) of ``_compute_bounds_numba`` (``glyphs/glyph.py:66-78``)."""
    arr = _f64(np.ravel(arr))
    out = np.empty(2, np.float64)
    _load().dsx_oracle_bounds(_ptr(arr), len(arr), _ptr(out))
    return float(out[0]), float(out[1])

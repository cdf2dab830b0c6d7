"""ctypes wrappers of ``libdsx_cpubase.so`` (``cpu_baseline.c``): the CPU baseline that
``bench.py`` times beside the GPU numbers (TEST INFRASTRUCTURE ONLY, like the rest of
this package).

Each function runs the reference's CPU algorithm for one workload on ``threads`` host
threads: 1 is the pandas path (numba ``ngjit``, one core, ``utils.py:52``), P > 1 the
dask path with ``npartitions = P`` -- private canvases per partition and a binary combine
tree (``data_libraries/dask.py:168-217``, ``split_every=2``).  ``tests/test_oracle.py``
checks every function, at 1 and several threads, against the sequential oracle.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdsx_cpubase.so")
_lib = None


class _View(ctypes.Structure):
    _fields_ = [("sx", ctypes.c_double), ("tx", ctypes.c_double),
                ("sy", ctypes.c_double), ("ty", ctypes.c_double),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("xxmax", ctypes.c_int64), ("yymax", ctypes.c_int64)]


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        lib.cpub_points_count_mean_f32.argtypes = [vp, vp, vp, i64, P(_View), i32, i32, i32,
                                                   vp, vp, vp, vp]
        lib.cpub_points_count_f64.argtypes = [vp, vp, i64, P(_View), i32, i32, i32, vp]
        lib.cpub_points_count_cat_f32.argtypes = [vp, vp, vp, i64, P(_View), i32, i32, i32,
                                                  i32, vp]
        lib.cpub_lines_xconst_f32.argtypes = [vp, i64, vp, i64, P(_View), i32, i32, i32, vp]
        for f in ("cpub_points_count_mean_f32", "cpub_points_count_f64",
                  "cpub_points_count_cat_f32", "cpub_lines_xconst_f32"):
            getattr(lib, f).restype = ctypes.c_int
        _lib = lib
    return _lib


def _view(v):
    return _View(v.sx, v.tx, v.sy, v.ty, v.xmin, v.xmax, v.ymin, v.ymax, v.xxmax, v.yymax)


def _c(a, dtype):
    a = np.ascontiguousarray(a)
    if a.dtype != dtype:
        raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
    return a


def _ok(rc, what):
    if rc != 0:
        raise MemoryError(f"{what}: the CPU baseline could not allocate its canvases")


def points_count_mean_f32(x, y, v, view, W, H, threads=1):
    """count() + mean(v) over float32 columns: {"count", "countv", "sumz", "mean"}."""
    x, y, v = (_c(a, np.float32) for a in (x, y, v))
    count = np.empty((H, W), np.uint32)
    countv = np.empty((H, W), np.uint32)
    total = np.empty((H, W), np.float64)
    mean = np.empty((H, W), np.float64)
    _ok(_load().cpub_points_count_mean_f32(
        x.ctypes.data, y.ctypes.data, v.ctypes.data, len(x), ctypes.byref(_view(view)), W, H,
        threads, count.ctypes.data, countv.ctypes.data, total.ctypes.data, mean.ctypes.data),
        "points_count_mean_f32")
    return {"count": count, "countv": countv, "sumz": total, "mean": mean}


def points_count_f64(x, y, view, W, H, threads=1):
    """count() over float64 columns."""
    x, y = (_c(a, np.float64) for a in (x, y))
    count = np.empty((H, W), np.uint32)
    _ok(_load().cpub_points_count_f64(x.ctypes.data, y.ctypes.data, len(x),
                                      ctypes.byref(_view(view)), W, H, threads,
                                      count.ctypes.data), "points_count_f64")
    return count


def points_count_cat_f32(x, y, cat, view, W, H, ncat, threads=1):
    """by(cat, count()) over float32 columns and uint8 codes: uint32[H, W, ncat]."""
    x, y = (_c(a, np.float32) for a in (x, y))
    cat = _c(cat, np.uint8)
    count = np.empty((H, W, ncat), np.uint32)
    _ok(_load().cpub_points_count_cat_f32(x.ctypes.data, y.ctypes.data, cat.ctypes.data,
                                          len(x), ctypes.byref(_view(view)), W, H, ncat,
                                          threads, count.ctypes.data), "points_count_cat_f32")
    return count


def lines_xconst_f32(xs, ys, view, W, H, threads=1):
    """LinesAxis1XConstant count: float64 shared x[ncols], float32 ys[nrows, ncols]."""
    xs = _c(xs, np.float64)
    ys = _c(ys, np.float32)
    count = np.empty((H, W), np.uint32)
    _ok(_load().cpub_lines_xconst_f32(xs.ctypes.data, ys.shape[1], ys.ctypes.data,
                                      ys.shape[0], ctypes.byref(_view(view)), W, H, threads,
                                      count.ctypes.data), "lines_xconst_f32")
    return count


def host_info():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}

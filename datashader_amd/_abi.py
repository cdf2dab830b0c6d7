"""ctypes binding of libdsx.so (``include/dsx.h``).

The structures below mirror the header field for field.  Every call checks the returned
status and raises; there is no fallback: without the built library (or without a GPU)
the aggregation path fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdsx.so")

ABI_VERSION = 6

# dtype codes
F32, F64, I8, U8, I16, I32, I64 = 1, 2, 3, 4, 5, 6, 7
U32, BOOL = 8, 9
# row-index reduction modes
RI_FIRST, RI_LAST, RI_ARGMAX, RI_ARGMIN = 1, 2, 3, 4
NS_FIRST, NS_LAST, NS_MAX, NS_MIN = 1, 2, 3, 4
# shade kinds and transforms
SHADE_CMAP, SHADE_COLOR, SHADE_CATEGORICAL, SHADE_COLOR_KEY = 0, 1, 2, 3
HOW_CODES = {"linear": 0, "log": 1, "cbrt": 2, "eq_hist": 3}
# lines: plot_start of a chunk whose first vertex is context only (axis 0)
PLOT_CONTEXT = -1
# fill modes
FILL_ZERO32, FILL_ONES32, FILL_ZERO64, FILL_ONES64 = 1, 2, 3, 4
# finalize modes
FIN_ANY, FIN_SUM, FIN_MEAN, FIN_VAR, FIN_STD, FIN_MIN, FIN_MAX, FIN_SUMZERO = range(1, 9)

EXPORTS = ("dsx_points_workspace_size", "dsx_points_agg", "dsx_lines_agg",
           "dsx_bounds_workspace_size", "dsx_bounds", "dsx_bounds_noalloc", "dsx_fill", "dsx_finalize",
           "dsx_grid_index_workspace_size", "dsx_grid_index_build", "dsx_gather_rows",
           "dsx_points_indexed_workspace_size", "dsx_points_agg_indexed",
           "dsx_shade_workspace_size", "dsx_shade", "dsx_combine", "dsx_points_rowindex",
           "dsx_points_rowindex_workspace_size",
           "dsx_take", "dsx_points_nselect_workspace_size", "dsx_points_nselect",
           "dsx_batch_workspace_size", "dsx_batch_table_bytes", "dsx_batch_tile_cells",
           "dsx_batch_create", "dsx_batch_table", "dsx_batch_submit", "dsx_batch_destroy",
           "dsx_abi_version", "dsx_strerror", "dsx_last_hip_error")

vp = ctypes.c_void_p


class Planes(ctypes.Structure):
    _fields_ = [("count", vp), ("countv", vp), ("sum", vp), ("min", vp), ("max", vp),
                ("mean", vp)]


class View(ctypes.Structure):
    _fields_ = [("sx", ctypes.c_double), ("tx", ctypes.c_double),
                ("sy", ctypes.c_double), ("ty", ctypes.c_double),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("xxmax", ctypes.c_int64), ("yymax", ctypes.c_int64),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("xlog", ctypes.c_int32), ("ylog", ctypes.c_int32)]


CAT_NONE, CAT_CODES, CAT_MODULO, CAT_BINNING = 0, 1, 2, 3
# points: how a call splits the rows between its passes (dsx_strategy)
STRATEGY_AUTO, STRATEGY_WINDOW, STRATEGY_TILES = 0, 1, 2
# words of a window record {x0, y0, w, h, strategy, 0, 0, 0}
WINDOW_RECORD_WORDS = 8


class Categorizer(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("dtype", ctypes.c_int32), ("col", vp),
                ("ncat", ctypes.c_int32), ("nbins", ctypes.c_int32),
                ("bin_under", ctypes.c_int32), ("bin_over", ctypes.c_int32),
                ("offset", ctypes.c_int64), ("modulo", ctypes.c_int64),
                ("bin0", ctypes.c_double), ("binsize", ctypes.c_double)]


class PointsDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", vp), ("y", vp), ("xy_dtype", ctypes.c_int32),
                ("val_dtype", ctypes.c_int32), ("val", vp), ("cat", Categorizer),
                ("window", vp), ("reuse_window", ctypes.c_int32),
                ("strategy", ctypes.c_int32)]


class LinesDesc(ctypes.Structure):
    _fields_ = [("nlines", ctypes.c_int64), ("nverts", ctypes.c_int64), ("x", vp), ("y", vp),
                ("x_dtype", ctypes.c_int32), ("y_dtype", ctypes.c_int32),
                ("x_line_stride", ctypes.c_int64), ("x_vert_stride", ctypes.c_int64),
                ("y_line_stride", ctypes.c_int64), ("y_vert_stride", ctypes.c_int64),
                ("axis", ctypes.c_int32), ("plot_start", ctypes.c_int32),
                ("val", vp), ("val_dtype", ctypes.c_int32), ("cat", Categorizer),
                ("mean", vp)]


class FinalizeDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("key_bytes", ctypes.c_int32),
                ("ncells", ctypes.c_int64), ("count", vp), ("countv", vp), ("sum", vp),
                ("key", vp), ("out", vp)]


class GridIndexDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("grid", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("x", vp), ("y", vp), ("val", vp),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double),
                ("out_x", vp), ("out_y", vp), ("out_val", vp), ("cell_start", vp)]


class GridRef(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double), ("cell_start", vp)]


class GatherDesc(ctypes.Structure):
    _fields_ = [("ncols", ctypes.c_int32), ("elem_bytes", ctypes.c_int32 * 4),
                ("src", vp * 4), ("dst", vp * 4), ("items", vp), ("nitems", ctypes.c_int64)]


class RowIndexDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("nan_check", ctypes.c_int32), ("key", vp),
                ("index", vp)]


class NSelectDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("n", ctypes.c_int32), ("nan_check", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("index", vp)]


class TakeDesc(ctypes.Structure):
    _fields_ = [("ncells", ctypes.c_int64), ("index", vp), ("col", vp),
                ("dtype", ctypes.c_int32), ("reserved", ctypes.c_int32), ("out", vp)]


class CombineDesc(ctypes.Structure):
    _fields_ = [("ncells", ctypes.c_int64), ("key_bytes", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("into", Planes), ("part", Planes)]


BATCH_COUNT, BATCH_COUNTV, BATCH_SUM, BATCH_MEAN = 1, 2, 4, 8


class BatchReq(ctypes.Structure):
    _fields_ = [("sx", ctypes.c_double), ("tx", ctypes.c_double),
                ("sy", ctypes.c_double), ("ty", ctypes.c_double),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32),
                ("part_rows", ctypes.c_int64), ("tile_base", ctypes.c_int32),
                ("reserved", ctypes.c_int32),
                ("count", vp), ("countv", vp), ("sum", vp), ("mean", vp)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", vp), ("y", vp), ("val", vp),
                ("grid", ctypes.c_int32), ("planes", ctypes.c_uint32),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double), ("cell_start", vp),
                ("max_requests", ctypes.c_int32), ("max_tiles", ctypes.c_int32),
                ("max_regions", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("workgroups", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ShadeDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("H", ctypes.c_int64), ("W", ctypes.c_int64), ("ncat", ctypes.c_int64),
                ("data", vp), ("out", vp), ("how", ctypes.c_int32), ("ncolors", ctypes.c_int32),
                ("colors", vp), ("keys", vp),
                ("alpha", ctypes.c_double), ("min_alpha", ctypes.c_double),
                ("nalpha", ctypes.c_int32), ("has_span", ctypes.c_int32),
                ("span_lo", ctypes.c_double), ("span_hi", ctypes.c_double),
                ("span_delta", ctypes.c_double),
                ("rescale_discrete_levels", ctypes.c_int32), ("has_baseline", ctypes.c_int32),
                ("baseline", ctypes.c_double)]


class DsxError(RuntimeError):
    """A libdsx call returned a non-zero status."""


_lib = None


def load():
    """Load libdsx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m datashader_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        i32, i64, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
        lib.dsx_points_workspace_size.argtypes = [P(PointsDesc), P(View), P(Planes)]
        lib.dsx_points_workspace_size.restype = sz
        lib.dsx_points_agg.argtypes = [P(PointsDesc), P(View), P(Planes), vp, vp, sz, vp]
        lib.dsx_lines_agg.argtypes = [P(LinesDesc), P(View), P(Planes), vp]
        lib.dsx_bounds_workspace_size.argtypes = [i64]
        lib.dsx_bounds_workspace_size.restype = sz
        lib.dsx_bounds.argtypes = [vp, i32, i64, vp, vp, sz, vp]
        lib.dsx_bounds_noalloc.argtypes = [vp, i32, i64, vp, vp]
        lib.dsx_fill.argtypes = [vp, i32, i64, vp]
        lib.dsx_finalize.argtypes = [P(FinalizeDesc), vp]
        lib.dsx_grid_index_workspace_size.argtypes = [P(GridIndexDesc)]
        lib.dsx_grid_index_workspace_size.restype = sz
        lib.dsx_grid_index_build.argtypes = [P(GridIndexDesc), vp, sz, vp]
        lib.dsx_gather_rows.argtypes = [P(GatherDesc), vp]
        lib.dsx_points_indexed_workspace_size.argtypes = [P(PointsDesc), P(View), P(Planes),
                                                          P(GridRef)]
        lib.dsx_points_indexed_workspace_size.restype = sz
        lib.dsx_points_agg_indexed.argtypes = [P(PointsDesc), P(View), P(Planes), vp,
                                               P(GridRef), vp, sz, vp]
        lib.dsx_combine.argtypes = [P(CombineDesc), vp]
        lib.dsx_points_rowindex.argtypes = [P(PointsDesc), P(View), P(RowIndexDesc), vp, sz, vp]
        lib.dsx_points_rowindex_workspace_size.argtypes = [P(PointsDesc), P(View), P(RowIndexDesc)]
        lib.dsx_points_rowindex_workspace_size.restype = sz
        lib.dsx_take.argtypes = [P(TakeDesc), vp]
        lib.dsx_points_nselect_workspace_size.argtypes = [P(PointsDesc), P(View), P(NSelectDesc)]
        lib.dsx_points_nselect_workspace_size.restype = sz
        lib.dsx_points_nselect.argtypes = [P(PointsDesc), P(View), P(NSelectDesc), vp, sz, vp]
        lib.dsx_shade_workspace_size.argtypes = [P(ShadeDesc)]
        lib.dsx_shade_workspace_size.restype = sz
        lib.dsx_shade.argtypes = [P(ShadeDesc), vp, sz, vp]
        lib.dsx_batch_workspace_size.argtypes = [P(BatchConfig)]
        lib.dsx_batch_workspace_size.restype = sz
        lib.dsx_batch_table_bytes.argtypes = [i32]
        lib.dsx_batch_table_bytes.restype = sz
        lib.dsx_batch_tile_cells.argtypes = [P(BatchConfig)]
        lib.dsx_batch_tile_cells.restype = i32
        lib.dsx_batch_create.argtypes = [P(BatchConfig), vp, sz, vp, i32, P(vp)]
        lib.dsx_batch_table.argtypes = [vp]
        lib.dsx_batch_table.restype = P(BatchReq)
        lib.dsx_batch_submit.argtypes = [vp, i32, vp]
        lib.dsx_batch_destroy.argtypes = [vp]
        lib.dsx_abi_version.argtypes = []
        lib.dsx_strerror.argtypes = [ctypes.c_int]
        lib.dsx_strerror.restype = ctypes.c_char_p
        lib.dsx_last_hip_error.argtypes = []
        lib.dsx_last_hip_error.restype = ctypes.c_char_p
        for f in ("dsx_points_agg", "dsx_lines_agg", "dsx_bounds", "dsx_bounds_noalloc", "dsx_fill", "dsx_finalize",
                  "dsx_grid_index_build", "dsx_gather_rows", "dsx_points_agg_indexed",
                  "dsx_shade", "dsx_combine", "dsx_points_rowindex", "dsx_take",
                  "dsx_points_nselect", "dsx_batch_create", "dsx_batch_submit",
                  "dsx_batch_destroy", "dsx_abi_version"):
            getattr(lib, f).restype = ctypes.c_int
        if lib.dsx_abi_version() != ABI_VERSION:
            raise RuntimeError("libdsx.so ABI version mismatch; rebuild the library")
        _lib = lib
    return _lib


def check(status: int, what: str):
    if status != 0:
        lib = load()
        msg = lib.dsx_strerror(status).decode()
        if status == -5:
            msg += ": " + lib.dsx_last_hip_error().decode()
        raise DsxError(f"{what} failed ({status}): {msg}")

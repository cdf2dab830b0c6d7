"""datashader_amd -- an MI355X-native (gfx950) implementation of datashader's points / line
projection + aggregation path.

Drop-in for ``datashader.Canvas.points`` / ``Canvas.line`` with the reductions
``count, any, sum, mean, min, max, var, std, by, count_cat, summary``: the same call
signatures and the same DataArray / Dataset results (dims, coords, attrs, dtypes).  The
projection and aggregation run as hand-written HIP kernels in ``libdsx.so`` (C ABI in
``include/dsx.h``), called through ctypes; columns live in HBM as torch-ROCm tensors.
"""
from __future__ import annotations

__version__ = "0.1.0"

from .core import Canvas, LinearAxis, LogAxis, bypixel  # noqa: F401
from .frame import DeviceFrame, Schema  # noqa: F401
from .reductions import (_max_n_row_index, _max_row_index,  # noqa: F401,A004
                         _min_n_row_index, _min_row_index, any, by, category_binning,
                         category_codes, category_modulo, count, count_cat, first, first_n,
                         last, last_n, max, max_n, mean, min, min_n, std, sum, summary, var,
                         where)
from . import backend as _backend  # noqa: F401  (registers the pipelines)
from .streaming import ChunkedFrame  # noqa: F401  (registers the chunked pipeline)
from .xrlite import DataArray, Dataset  # noqa: F401

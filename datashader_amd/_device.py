"""Device plumbing shared by the glyphs and the backend: stream handle, dtype codes,
column conversion and the autorange bounds call.  torch is used for device memory only;
every computation on the columns is a libdsx kernel."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi


def torch():
    import torch as _t
    return _t


def stream_handle():
    """The current torch stream as the ``void*`` the C ABI takes."""
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream().cuda_stream)


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("datashader_amd needs a ROCm GPU (MI355X / gfx950); none is visible")
    _abi.load()


def dtype_code(t):
    tt = torch()
    m = {tt.float32: _abi.F32, tt.float64: _abi.F64, tt.int8: _abi.I8, tt.uint8: _abi.U8,
         tt.int16: _abi.I16, tt.int32: _abi.I32, tt.int64: _abi.I64}
    if t.dtype not in m:
        raise TypeError(f"unsupported column dtype {t.dtype}")
    return m[t.dtype]


def as_float_column(t):
    """A float32 / float64 view of a numeric column.  Integer and bool columns are
    promoted to float64 -- the reference's numba code promotes them the same way in
    every comparison and accumulation, exactly for |v| < 2**53."""
    tt = torch()
    if t.dtype in (tt.float32, tt.float64):
        return t if t.is_contiguous() else t.contiguous()
    return t.to(tt.float64)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def bounds(col):
    """NaN-skipping (min, max) of a device column on the GPU (glyph.py:66-78).  Returns
    Python floats (the reference's numba function returns float64)."""
    tt = torch()
    lib = _abi.load()
    col = as_float_column(col)
    n = int(col.numel())
    ws = int(lib.dsx_bounds_workspace_size(n))
    work = tt.empty(max(ws, 16), dtype=tt.uint8, device=col.device)
    out = tt.empty(2, dtype=tt.float64, device=col.device)
    _abi.check(lib.dsx_bounds(ptr(col), dtype_code(col), n, ptr(out), ptr(work), ws,
                              stream_handle()), "dsx_bounds")
    lo, hi = out.cpu().tolist()
    return lo, hi


def maybe_expand_bounds(b):
    """``Glyph.maybe_expand_bounds`` (glyph.py:42-49)."""
    minval, maxval = b
    if not (np.isfinite(minval) and np.isfinite(maxval)):
        minval, maxval = -1.0, 1.0
    elif minval == maxval:
        minval, maxval = minval - 1, minval + 1
    return minval, maxval


def column_matrix(frame, names):
    """``names`` as a strided matrix: returns ``(base, col_stride, row_stride)`` such that
    element (column j, row r) is ``base.data[j*col_stride + r*row_stride]`` (element
    strides).  Columns that are equally spaced views of one device allocation -- rows of
    a column-major block or columns of a row-major one -- are read in place; otherwise
    they are stacked once (as the reference's ``df.loc[:, names].to_numpy()`` copies)."""
    tt = torch()
    cols = [frame[n] for n in names]
    dtypes = {c.dtype for c in cols}
    if len(dtypes) == 1 and cols[0].dtype in (tt.float32, tt.float64):
        rs = cols[0].stride(0)
        if len(cols) == 1:
            return cols[0], 0, rs
        item = cols[0].element_size()
        d = cols[1].data_ptr() - cols[0].data_ptr()
        base_storage = cols[0].untyped_storage().data_ptr()
        if (d % item == 0 and d != 0 and all(c.stride(0) == rs for c in cols)
                and all(c.untyped_storage().data_ptr() == base_storage for c in cols)
                and all(cols[j].data_ptr() - cols[0].data_ptr() == j * d
                        for j in range(len(cols)))):
            return cols[0], d // item, rs
    # mixed dtypes are promoted to float64 (numpy's common type of the columns)
    target = cols[0].dtype if len(dtypes) == 1 and cols[0].dtype in (tt.float32, tt.float64) \
        else tt.float64
    m = tt.stack([c.to(target) for c in cols]).contiguous()
    return m[0], int(m.shape[1]), 1

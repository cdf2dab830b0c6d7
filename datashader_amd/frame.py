"""Device-resident column frames and the schema the glyphs / reductions validate against.

``DeviceFrame`` is this package's device-frame type, registered with
``bypixel.pipeline`` the way the reference registers ``cudf.DataFrame``
(``data_libraries/cudf.py:7-14``).  Its columns live in HBM as torch-ROCm tensors
(torch is used only for device memory).  Columns of one float dtype are stored as rows
of a single ``[ncols, nrows]`` block, so a group of them -- e.g. the 256 y columns of an
axis-1 line frame -- is one strided matrix the line kernel reads in place, with no
``to_numpy()`` materialisation (the reference's ``line.py:370``).
"""
from __future__ import annotations

import numpy as np
import pandas as pd

from . import _abi

_KIND = {np.dtype("float32"): "f4", np.dtype("float64"): "f8", np.dtype("int8"): "i1",
         np.dtype("int16"): "i2", np.dtype("int32"): "i4", np.dtype("int64"): "i8",
         np.dtype("uint8"): "u1", np.dtype("uint16"): "u2", np.dtype("uint32"): "u4",
         np.dtype("uint64"): "u8", np.dtype("bool"): "b1"}

DSX_DTYPE = {"f4": _abi.F32, "f8": _abi.F64, "i1": _abi.I8, "u1": _abi.U8, "i2": _abi.I16,
             "i4": _abi.I32, "i8": _abi.I64}


def _torch():
    import torch
    return torch


def is_ragged(series):
    """A pandas column whose rows are variable-length sequences: the reference's
    ``RaggedArray`` (anything exposing ``start_indices`` / ``flat_array``) or an object
    column of arrays / lists."""
    arr = getattr(series, "array", None)
    if hasattr(arr, "start_indices") and hasattr(arr, "flat_array"):
        return True
    if series.dtype != object or len(series) == 0:
        return False
    first = series.iloc[0]
    return isinstance(first, (list, tuple, np.ndarray))


def ragged_parts(series):
    """(start indices with the total appended, flat values) of a ragged column."""
    arr = getattr(series, "array", None)
    if hasattr(arr, "start_indices") and hasattr(arr, "flat_array"):
        flat = np.asarray(arr.flat_array)
        starts = np.append(np.asarray(arr.start_indices, np.int64), len(flat))
        return starts, flat
    rows = [np.asarray(r) for r in series]
    lens = np.array([len(r) for r in rows], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    dt = np.float32 if rows and all(r.dtype == np.float32 for r in rows) else np.float64
    flat = np.concatenate([r.astype(dt) for r in rows]) if rows else np.zeros(0, dt)
    return starts, flat


class RaggedColumn:
    """A ragged column on the device: row r holds ``flat[starts[r]:starts[r+1]]``
    (``starts`` has one entry more than rows).  Counts as a 1-D column of ``len(starts)-1``
    rows in a DeviceFrame."""

    def __init__(self, starts, flat):
        self.starts = starts
        self.flat = flat

    @property
    def shape(self):
        return (int(self.starts.numel()) - 1,)

    def dim(self):
        return 1

    @property
    def dtype(self):
        return self.flat.dtype

    @property
    def device(self):
        return self.flat.device

    def __len__(self):
        return self.shape[0]

    @classmethod
    def from_series(cls, series, device="cuda"):
        torch = _torch()
        starts, flat = ragged_parts(series)
        return cls(torch.from_numpy(starts).to(device),
                   torch.from_numpy(np.ascontiguousarray(flat)).to(device))


class Schema:
    """Column name -> kind ('f4', 'f8', 'i4', ..., 'category'), plus categories."""

    def __init__(self, kinds, categories=None):
        self._kinds = dict(kinds)
        self._cats = dict(categories or {})

    def __contains__(self, col):
        return col in self._kinds

    def __getitem__(self, col):
        return self._kinds[col]

    def categories(self, col):
        return self._cats[col]

    @property
    def dict(self):
        return self._kinds

    @classmethod
    def from_pandas(cls, df):
        kinds, cats = {}, {}
        for name, dt in df.dtypes.items():
            if isinstance(dt, pd.CategoricalDtype):
                kinds[name] = "category"
                cats[name] = list(dt.categories)
            elif is_ragged(df[name]):
                kinds[name] = "ragged"
            else:
                kinds[name] = _KIND.get(np.dtype(dt), "O") if dt != object else "O"
        return cls(kinds, cats)


def _torch_kind(t):
    if isinstance(t, RaggedColumn):
        return "ragged"
    torch = _torch()
    m = {torch.float32: "f4", torch.float64: "f8", torch.int8: "i1", torch.int16: "i2",
         torch.int32: "i4", torch.int64: "i8", torch.uint8: "u1", torch.bool: "b1"}
    return m.get(t.dtype, "O")


class DeviceFrame:
    """Columns resident in device memory.

    ``columns`` maps name -> 1-D device tensor; a categorical column holds its integer
    codes and ``categories[name]`` lists its categories (like pandas' Categorical)."""

    def __init__(self, columns, categories=None):
        self._cols = dict(columns)
        self._cats = dict(categories or {})
        lens = {int(t.shape[0]) for t in self._cols.values()}
        if len(lens) > 1:
            raise ValueError("all columns of a DeviceFrame must have the same length")
        self._len = lens.pop() if lens else 0
        # resident frames keep per-canvas tuning state (the hot-window cache)
        self.persistent = True
        for name, t in self._cols.items():
            if t.dim() != 1:
                raise ValueError(f"column {name!r} is not 1-D")

    @classmethod
    def from_pandas(cls, df, device="cuda"):
        """Upload a pandas DataFrame.  Float columns of one dtype share a block."""
        torch = _torch()
        cols, cats = {}, {}
        groups = {}
        for name in df.columns:
            dt = df.dtypes[name]
            if isinstance(dt, pd.CategoricalDtype):
                codes = np.ascontiguousarray(df[name].cat.codes.to_numpy())
                cols[name] = torch.from_numpy(codes).to(device)
                cats[name] = list(dt.categories)
            elif is_ragged(df[name]):
                cols[name] = RaggedColumn.from_series(df[name], device)
            elif not isinstance(dt, np.dtype):
                continue  # other extension types are never aggregated
            elif np.dtype(dt).kind == "f":
                groups.setdefault(np.dtype(dt), []).append(name)
            elif np.dtype(dt).kind in "iub":
                cols[name] = torch.from_numpy(np.ascontiguousarray(df[name].to_numpy())).to(device)
            else:
                continue  # non-numeric columns stay on the host (never aggregated)
        for dt, names in groups.items():
            host = np.empty((len(names), len(df)), dtype=dt)
            for j, name in enumerate(names):
                host[j] = df[name].to_numpy()
            block = torch.from_numpy(host).to(device)
            for j, name in enumerate(names):
                cols[name] = block[j]
        ordered = {name: cols[name] for name in df.columns if name in cols}
        return cls(ordered, cats)

    @classmethod
    def from_block(cls, names, block, categories=None, **other):
        """Columns ``names`` as the rows of a ``[len(names), nrows]`` device tensor."""
        if block.dim() != 2 or block.shape[0] != len(names):
            raise ValueError("block must be [len(names), nrows]")
        cols = {n: block[j] for j, n in enumerate(names)}
        cols.update(other)
        return cls(cols, categories)

    def __len__(self):
        return self._len

    @property
    def columns(self):
        return list(self._cols)

    def __getitem__(self, name):
        return self._cols[name]

    def __contains__(self, name):
        return name in self._cols

    @property
    def schema(self):
        kinds = {}
        for n, t in self._cols.items():
            kinds[n] = "category" if n in self._cats else _torch_kind(t)
        return Schema(kinds, self._cats)

    def categories(self, name):
        return self._cats[name]

    def select(self, names):
        return DeviceFrame({n: self._cols[n] for n in names},
                           {n: c for n, c in self._cats.items() if n in names})

    def to_pandas(self):
        out = {}
        for n, t in self._cols.items():
            a = t.detach().cpu().numpy()
            if n in self._cats:
                out[n] = pd.Categorical.from_codes(a, self._cats[n])
            else:
                out[n] = a
        return pd.DataFrame(out)

    @property
    def device(self):
        for t in self._cols.values():
            return t.device
        return None

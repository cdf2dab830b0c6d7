"""Reductions of the points / line path, mirroring datashader's ``reductions.py``.

Same class names, constructor arguments, validation errors and output dtypes as the
reference (datashader 0.19.1):

  count(column=None)      reductions.py:532-664   uint32
  any(column=None)        reductions.py:825-890   bool
  sum(column)             reductions.py:1030-1098 float64, NaN where no value
  mean(column)            reductions.py:1280-1297 float64
  min(column), max(column) reductions.py:1160-1260 float64
  var(column), std(column) reductions.py:1300-1337 float64
  by(cat, reduction)      reductions.py:696-823   reduction's dtype, extra category dim
  count_cat(column)       reductions.py:1263-1278
  summary(**named)        reductions.py:2169-2233 Dataset, keys sorted
  category_codes / category_modulo / category_binning   reductions.py:143-259

Each reduction says which accumulator planes of the HIP kernels it needs
(:meth:`Reduction.planes`) and how its result is finalized (:attr:`finalize_mode`);
the backend groups reductions that share a value column into one kernel pass.
"""
from __future__ import annotations

import numpy as np

from . import _abi

__all__ = ["Reduction", "count", "any", "sum", "mean", "min", "max", "var", "std", "by",
           "count_cat", "summary", "category_codes", "category_modulo", "category_binning"]


# ---- schema helpers (datashape predicates isreal / isnumeric, utils.py:125-158) ------

def _is_numeric(kind):
    return kind in ("f4", "f8", "i1", "i2", "i4", "i8", "u1", "u2", "u4", "u8")


def _is_integer(kind):
    return kind in ("i1", "i2", "i4", "i8", "u1", "u2", "u4", "u8", "b1")


# ---- categorizers --------------------------------------------------------------------

class CategoryPreprocess:
    """Base class of categorizers (reductions.py:123-141)."""

    def __init__(self, column):
        self.column = column

    @property
    def cat_column(self):
        return self.column

    def _key(self):
        return (type(self).__name__, self.column)

    def __eq__(self, other):
        return type(self) is type(other) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())


class category_codes(CategoryPreprocess):
    """Category codes of a pandas categorical column (reductions.py:143-169)."""

    def categories(self, schema):
        return list(schema.categories(self.column))

    def validate(self, schema):
        if self.column not in schema:
            raise ValueError("specified column not found")
        if schema[self.column] != "category":
            raise ValueError("input must be categorical")

    def kernel_args(self, schema):
        return dict(kind=_abi.CAT_CODES, ncat=len(self.categories(schema)))


class category_modulo(category_codes):
    """``(column - offset) % modulo`` of an integer column (reductions.py:171-206)."""

    def __init__(self, column, modulo, offset=0):
        super().__init__(column)
        self.offset = offset
        self.modulo = modulo

    def _key(self):
        return super()._key() + (self.offset, self.modulo)

    def categories(self, schema):
        return list(range(self.modulo))

    def validate(self, schema):
        if self.column not in schema:
            raise ValueError("specified column not found")
        if not _is_integer(schema[self.column]):
            raise ValueError("input must be an integer column")

    def kernel_args(self, schema):
        return dict(kind=_abi.CAT_MODULO, ncat=self.modulo, offset=int(self.offset),
                    modulo=int(self.modulo))


class category_binning(category_modulo):
    """Bins of a continuous column; category ``nbins`` holds NaN and, unless included,
    values outside [lower, upper) (reductions.py:208-259)."""

    def __init__(self, column, lower, upper, nbins, include_under=True, include_over=True):
        super().__init__(column, nbins + 1)
        self.bin0 = lower
        self.binsize = (upper - lower) / float(nbins)
        self.nbins = nbins
        self.bin_under = 0 if include_under else nbins
        self.bin_over = nbins - 1 if include_over else nbins

    def _key(self):
        return super()._key() + (self.bin0, self.binsize, self.bin_under, self.bin_over)

    def validate(self, schema):
        if self.column not in schema:
            raise ValueError("specified column not found")

    def kernel_args(self, schema):
        return dict(kind=_abi.CAT_BINNING, ncat=self.nbins + 1, nbins=self.nbins,
                    bin_under=self.bin_under, bin_over=self.bin_over,
                    bin0=float(self.bin0), binsize=float(self.binsize))


# ---- reductions ------------------------------------------------------------------------

class Reduction:
    """Base class for per-bin reductions (reductions.py:309-475)."""

    #: accumulator planes the kernel pass must produce for this reduction
    _planes: tuple = ()
    #: dsx finalize mode (None: the plane itself is the result)
    finalize_mode = None
    #: numpy dtype of the finalized result
    out_dtype = np.float64
    needs_m2 = False

    def __init__(self, column=None):
        self.column = column

    def validate(self, schema):
        if self.column not in schema:
            raise ValueError("specified column not found")
        if not _is_numeric(schema[self.column]):
            raise ValueError("input must be numeric")

    def planes(self):
        return self._planes

    def is_categorical(self):
        return False

    def _key(self):
        return (type(self).__name__, self.column)

    def __eq__(self, other):
        return type(self) is type(other) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())

    def __repr__(self):
        return f"{type(self).__name__}({self.column!r})"


class OptionalFieldReduction(Reduction):
    """count / any: the column is optional (reductions.py:478-493)."""

    def validate(self, schema):
        if self.column is not None:
            super().validate(schema)


class count(OptionalFieldReduction):
    """Count elements in each bin (uint32).  With a column, NaN values are not counted."""

    out_dtype = np.uint32

    def __init__(self, column=None, self_intersect=True):
        super().__init__(column)
        self.self_intersect = self_intersect

    def planes(self):
        return ("count",) if self.column is None else ("countv",)

    def __repr__(self):
        return "count()" if self.column is None else f"count({self.column!r})"


class any(OptionalFieldReduction):
    """Whether any elements map to each bin (bool); NaN values of the column skipped."""

    out_dtype = np.bool_
    finalize_mode = _abi.FIN_ANY

    def planes(self):
        return ("count",) if self.column is None else ("countv",)


class FloatingReduction(Reduction):
    """Reductions with a float64 result (reductions.py:928-935)."""


class sum(FloatingReduction):
    """Sum of the column in each bin; NaN where no non-NaN value lands."""

    _planes = ("countv", "sum")
    finalize_mode = _abi.FIN_SUM

    def __init__(self, column, self_intersect=True):
        super().__init__(column)
        self.self_intersect = self_intersect


class mean(Reduction):
    """Mean of the column in each bin (sum / count of non-NaN values)."""

    _planes = ("countv", "sum")
    finalize_mode = _abi.FIN_MEAN


class min(FloatingReduction):
    _planes = ("min",)
    finalize_mode = _abi.FIN_MIN


class max(FloatingReduction):
    _planes = ("max",)
    finalize_mode = _abi.FIN_MAX


class var(Reduction):
    """Population variance of the column in each bin."""

    _planes = ("countv", "sum")
    finalize_mode = _abi.FIN_VAR
    needs_m2 = True


class std(Reduction):
    """Population standard deviation of the column in each bin."""

    _planes = ("countv", "sum")
    finalize_mode = _abi.FIN_STD
    needs_m2 = True


class _first_or_last(FloatingReduction):
    """first / last (reductions.py:1340-1456): the value of the first / last row (in row
    order) whose value is not NaN; computed as the reference's GPU path does, through
    the smallest / largest row index of the bin and a lookup of that row."""

    row_index = True
    ri_mode = None

    def planes(self):
        return ()


class first(_first_or_last):
    ri_mode = _abi.RI_FIRST


class last(_first_or_last):
    ri_mode = _abi.RI_LAST


class _max_or_min_row_index(OptionalFieldReduction):
    """The largest / smallest row index in each bin, int64, -1 where empty
    (reductions.py:2236-2372)."""

    row_index = True
    ri_mode = None
    out_dtype = np.int64

    def __init__(self):
        super().__init__(None)

    def planes(self):
        return ()

    def __repr__(self):
        return f"{type(self).__name__}()"


class _min_row_index(_max_or_min_row_index):
    ri_mode = _abi.RI_FIRST


class _max_row_index(_max_or_min_row_index):
    ri_mode = _abi.RI_LAST


class FloatingNReduction(OptionalFieldReduction):
    """Reductions that keep ``n`` values per bin, along a trailing ``n`` dimension
    (reductions.py:1457-1494); float64 with NaN padding (int64 with -1 for the row-index
    forms).  Computed from the rows of each bin sorted in the reduction's order."""

    row_index = True
    ns_mode = None

    def __init__(self, column=None, n=1):
        super().__init__(column)
        self.n = n if n >= 1 else 1

    def planes(self):
        return ()

    def _key(self):
        return (type(self).__name__, self.column, self.n)

    def __repr__(self):
        return f"{type(self).__name__}(column={self.column!r}, n={self.n!r})"


class max_n(FloatingNReduction):
    """The ``n`` largest values of the bin, descending (reductions.py:1629-1713)."""
    ns_mode = _abi.NS_MAX


class min_n(FloatingNReduction):
    """The ``n`` smallest values of the bin, ascending (reductions.py:1716-1803)."""
    ns_mode = _abi.NS_MIN


class first_n(FloatingNReduction):
    """The first ``n`` non-NaN values of the bin in row order (reductions.py:1535-1588)."""
    ns_mode = _abi.NS_FIRST


class last_n(FloatingNReduction):
    """The last ``n`` non-NaN values of the bin, latest first (reductions.py:1591-1626)."""
    ns_mode = _abi.NS_LAST


class _max_n_or_min_n_row_index(FloatingNReduction):
    out_dtype = np.int64

    def __init__(self, n=1):
        super().__init__(None, n)

    def __repr__(self):
        return f"{type(self).__name__}(n={self.n!r})"


class _max_n_row_index(_max_n_or_min_n_row_index):
    """The ``n`` largest row indices of the bin, descending (reductions.py:2398-2475)."""
    ns_mode = _abi.NS_LAST


class _min_n_row_index(_max_n_or_min_n_row_index):
    """The ``n`` smallest row indices of the bin, ascending (reductions.py:2478-2562)."""
    ns_mode = _abi.NS_FIRST


class where(FloatingReduction):
    """Values of ``lookup_column`` at the row a ``selector`` (first, last, max, min or a
    row-index selector) picks in each bin; with no lookup column the row index itself
    (int64, -1 where empty) (reductions.py:1842-2166).  Among rows tied for the max / min
    the first in row order is picked, as the reference's row-order appends do."""

    row_index = True

    def __init__(self, selector, lookup_column=None):
        if not isinstance(selector, (first, last, max, min, _max_or_min_row_index,
                                     FloatingNReduction)):
            raise TypeError("selector can only be a first, first_n, last, last_n, "
                            "max, max_n, min or min_n reduction")
        super().__init__(lookup_column)
        self.selector = selector
        self.columns = (selector.column, lookup_column)

    @property
    def out_dtype(self):
        return np.int64 if self.column is None else np.float64

    def planes(self):
        return ()

    def validate(self, schema):
        if self.column is not None:
            super().validate(schema)
        self.selector.validate(schema)
        if self.column is not None and self.column == self.selector.column:
            raise ValueError("where and its contained reduction cannot use the same column")

    def _key(self):
        return ("where", self.selector._key(), self.column)

    def __repr__(self):
        return f"where(selector={self.selector!r}, lookup_column={self.column!r})"


class by(Reduction):
    """Apply ``reduction`` separately per category (reductions.py:696-823)."""

    def __init__(self, cat_column, reduction=None):
        super().__init__()
        if reduction is None:
            reduction = count()
        if isinstance(cat_column, CategoryPreprocess):
            self.categorizer = cat_column
        elif isinstance(cat_column, str):
            self.categorizer = category_codes(cat_column)
        else:
            raise TypeError("first argument must be a column name or a CategoryPreprocess instance")
        if isinstance(reduction, (by, summary)):
            raise NotImplementedError("by() of a categorical or summary reduction")
        self.column = self.categorizer.column
        self.columns = (self.categorizer.column, getattr(reduction, "column", None))
        self.reduction = reduction

    @property
    def cat_column(self):
        return self.columns[0]

    @property
    def val_column(self):
        return self.columns[1]

    def validate(self, schema):
        self.categorizer.validate(schema)
        self.reduction.validate(schema)

    def is_categorical(self):
        return True

    def planes(self):
        return self.reduction.planes()

    @property
    def finalize_mode(self):
        return self.reduction.finalize_mode

    @property
    def out_dtype(self):
        return self.reduction.out_dtype

    @property
    def needs_m2(self):
        return self.reduction.needs_m2

    def _key(self):
        return ("by", self.categorizer._key(), self.reduction._key())

    def __repr__(self):
        return f"by(column={self.column!r}, reduction={self.reduction!r})"


class count_cat(by):
    """Alias for ``by(column, count())`` (reductions.py:1263-1278)."""

    def __init__(self, column):
        super().__init__(column, count())

    def __repr__(self):
        return f"count_cat(column={self.column!r})"


class summary:
    """A collection of named reductions computed in one call; the result is a Dataset
    with the names sorted (reductions.py:2169-2233)."""

    def __init__(self, **kwargs):
        if not kwargs:
            raise ValueError("summary needs at least one reduction")
        ks, vs = zip(*sorted(kwargs.items()))
        self.keys = ks
        self.values = vs

    def validate(self, schema):
        for v in self.values:
            v.validate(schema)

    def is_categorical(self):
        return builtins_any(v.is_categorical() for v in self.values)

    def __repr__(self):
        pairs = ", ".join(f"{k}={v!r}" for k, v in zip(self.keys, self.values))
        return f"summary({pairs})"


def builtins_any(it):
    for v in it:
        if v:
            return True
    return False


def leaves(agg):
    """``[(name or None, reduction)]`` of an aggregation (compiler.py:271-277)."""
    if isinstance(agg, summary):
        return list(zip(agg.keys, agg.values))
    return [(None, agg)]

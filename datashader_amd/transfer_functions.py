"""``tf.shade`` on the device: aggregate -> RGBA image (``transfer_functions/__init__.py``).

``shade(agg, cmap, color_key, how, alpha, min_alpha, span, name, color_baseline,
rescale_discrete_levels)`` takes the reference's arguments and raises its errors
(``transfer_functions/__init__.py:616-747``).  The colouring itself -- mask, offset, clip,
the transform (``linear``, ``log``, ``cbrt``, ``eq_hist``), the span of the transformed
values, the colour / alpha interpolation and the categorical colour mix -- runs as the
``dsx_shade`` kernels of libdsx on the plane where it lies: a device aggregate stays on
the device and its image is a device tensor; a host (numpy) aggregate is uploaded and
its image returned as numpy, as the reference returns numpy images.

Callable ``how`` and callable (matplotlib) ``cmap`` run user numpy code on the host and
are not supported by this backend (NotImplementedError).
"""
from __future__ import annotations

from collections.abc import Iterator

import numpy as np

from . import _abi
from . import _device as dev
from . import xrlite
from .colors import Sets1to3, rgb

try:  # real xarray objects are accepted as input when xarray is importable
    import xarray as _xr
except Exception:  # pragma: no cover - xarray is not part of this image
    _xr = None

__all__ = ["Image", "shade"]


class Image(xrlite.DataArray):
    """An RGBA image: uint32 pixels ``r | g << 8 | b << 16 | a << 24``
    (``transfer_functions/__init__.py:30-83``)."""

    border = 1

    def to_pil(self, origin="lower"):
        from PIL.Image import fromarray
        data = self.values
        arr = np.flipud(data) if origin == "lower" else data
        if arr.ndim == 2 and arr.dtype == np.uint32:
            arr = arr.view(np.uint8).reshape(arr.shape + (4,))
        return fromarray(np.ascontiguousarray(arr))

    def to_bytesio(self, format="png", origin="lower"):
        from io import BytesIO
        fp = BytesIO()
        self.to_pil(origin).save(fp, format)
        fp.seek(0)
        return fp


def _is_dataarray(agg):
    return isinstance(agg, xrlite.DataArray) or (_xr is not None and isinstance(agg, _xr.DataArray))


def _coords_of(agg):
    """{name: values} of an xrlite or xarray DataArray, in its coordinate order."""
    return {k: np.asarray(c.values) for k, c in agg.coords.items()}


# dtype of the plane -> (ABI code, numpy dtype the kernels read)
_NP_CODES = {
    np.dtype(np.bool_): _abi.BOOL, np.dtype(np.uint32): _abi.U32,
    np.dtype(np.int32): _abi.I32, np.dtype(np.int64): _abi.I64,
    np.dtype(np.float32): _abi.F32, np.dtype(np.float64): _abi.F64,
}
_NP_WIDEN = {np.dtype(np.uint8): np.uint32, np.dtype(np.uint16): np.uint32,
             np.dtype(np.int8): np.int32, np.dtype(np.int16): np.int32}


def _device_plane(data):
    """(device tensor, ABI dtype code, numpy dtype of the values, came-from-host)."""
    t = dev.torch()
    dev.require_gpu()
    if isinstance(data, np.ndarray):
        a = data
        if a.dtype in _NP_WIDEN:  # same values and the same masking rule
            a = a.astype(_NP_WIDEN[a.dtype])
        if a.dtype not in _NP_CODES:
            raise TypeError(f"cannot shade an aggregate of dtype {data.dtype}")
        code = _NP_CODES[a.dtype]
        a = np.ascontiguousarray(a)
        host = t.from_numpy(a.view(np.uint8) if a.dtype == np.bool_ else
                            (a.view(np.int32) if a.dtype == np.uint32 else a))
        dt = host.to("cuda")
        if a.dtype == np.bool_:
            dt = dt.view(t.bool)
        elif a.dtype == np.uint32:
            dt = dt.view(t.uint32)
        return dt, code, a.dtype, True
    tt = data
    m = {t.bool: (_abi.BOOL, np.bool_), t.uint32: (_abi.U32, np.uint32),
         t.int32: (_abi.I32, np.int32), t.int64: (_abi.I64, np.int64),
         t.float32: (_abi.F32, np.float32), t.float64: (_abi.F64, np.float64)}
    if tt.dtype in (t.uint8, t.int8, t.int16):
        wide = tt.to(t.int32)
        tt = wide.view(t.uint32) if tt.dtype == t.uint8 else wide
    if tt.dtype not in m:
        raise TypeError(f"cannot shade an aggregate of dtype {tt.dtype}")
    code, npdt = m[tt.dtype]
    if not tt.is_cuda:
        tt = tt.to("cuda")
    return tt.contiguous(), code, np.dtype(npdt), False


_workspaces = {}
_palettes = {}


def _workspace(lib, desc, device):
    need = int(lib.dsx_shade_workspace_size(desc))
    if need == 0:
        raise ValueError("invalid shade descriptor")
    t = dev.torch()
    key = (str(device), t.cuda.current_stream().cuda_stream)
    ws = _workspaces.get(key)
    if ws is None or ws.numel() < need:
        ws = t.empty(need, dtype=t.uint8, device=device)
        _workspaces[key] = ws
    return ws


def _palette(colors, device):
    """Device uint8 [n, 3] of the (r, g, b) triples (cached per palette)."""
    key = (str(device), tuple(colors))
    p = _palettes.get(key)
    if p is None:
        t = dev.torch()
        flat = [int(c) for trip in colors for c in trip] or [0, 0, 0]
        p = t.tensor(flat, dtype=t.uint8).to(device)
        if len(_palettes) > 256:
            _palettes.clear()
        _palettes[key] = p
    return p


def _run(desc, device):
    lib = _abi.load()
    ws = None
    if desc.kind != _abi.SHADE_COLOR_KEY:
        ws = _workspace(lib, desc, device)
    _abi.check(lib.dsx_shade(desc, dev.ptr(ws), 0 if ws is None else ws.numel(),
                             dev.stream_handle()), "dsx_shade")


def _how_code(how):
    if callable(how):
        raise NotImplementedError(
            "callable `how` runs numpy code on the host; the MI355X backend shades on the "
            "device with 'eq_hist', 'log', 'cbrt' or 'linear'")
    if how not in _abi.HOW_CODES:
        raise ValueError(f"Unknown interpolation method: {how}")
    return _abi.HOW_CODES[how]


def _check_span(span, how):
    if span is not None and how == "eq_hist":
        raise ValueError("span is not (yet) valid to use with eq_hist")


def _span_fields(desc, span, how):
    if span is None:
        return
    _check_span(span, how)
    lo, hi = span[0], span[1]
    desc.has_span = 1
    desc.span_lo = float(lo)
    desc.span_hi = float(hi)
    desc.span_delta = float(hi - lo)


def _new_image(tensor, H, W, to_host, dims, coords, name):
    t = dev.torch()
    img = tensor.view(t.uint32).reshape(H, W)
    data = img.cpu().numpy() if to_host else img
    return Image(data, coords=coords, dims=dims, name=name)


def _interpolate(agg, cmap, how, alpha, span, min_alpha, name, rescale_discrete_levels):
    """``_interpolate`` (transfer_functions/__init__.py:251-357) on the device."""
    if agg.ndim != 2:
        raise ValueError("agg must be 2D")
    how_code = _how_code(how)
    if isinstance(cmap, Iterator):
        cmap = list(cmap)
    if isinstance(cmap, tuple) and len(cmap) and isinstance(cmap[0], str):
        cmap = list(cmap)
    if isinstance(cmap, list):
        colors = [rgb(c) for c in cmap]
        if not colors:
            raise ValueError("not enough values to unpack (expected 3, got 0)")
        if len(colors) > 4096:
            raise ValueError("the device colormap takes at most 4096 colours")
        kind = _abi.SHADE_CMAP
    elif isinstance(cmap, (str, tuple)):
        colors = [rgb(cmap)]
        kind = _abi.SHADE_COLOR
    elif callable(cmap):
        raise NotImplementedError(
            "matplotlib colormaps run on the host; pass a list of colours instead")
    else:
        raise TypeError("Expected `cmap` of `matplotlib.colors.Colormap`, "
                        f"`list`, `str`, or `tuple`; got: '{type(cmap)}'")
    _check_span(span, how)
    data, code, _, to_host = _device_plane(agg.data)
    H, W = (int(s) for s in agg.shape)
    t = dev.torch()
    out = t.empty(H * W, dtype=t.int32, device=data.device)
    d = _abi.ShadeDesc()
    d.kind, d.dtype, d.H, d.W, d.ncat = kind, code, H, W, 1
    d.data, d.out = dev.ptr(data), dev.ptr(out)
    d.how = how_code
    pal = _palette(colors, data.device)
    d.ncolors, d.colors = len(colors), dev.ptr(pal)
    d.alpha, d.min_alpha = float(alpha), float(min_alpha)
    d.nalpha = len(np.arange(min_alpha, alpha + 1)) if kind == _abi.SHADE_COLOR else 1
    if kind == _abi.SHADE_COLOR and d.nalpha < 1:
        raise ValueError("array of sample points is empty")
    _span_fields(d, span, how)
    d.rescale_discrete_levels = int(bool(rescale_discrete_levels))
    _run(d, data.device)
    return _new_image(out, H, W, to_host, list(agg.dims), _coords_of(agg), name)


def _colorize(agg, color_key, how, alpha, span, min_alpha, name, color_baseline,
              rescale_discrete_levels):
    """``_colorize`` + ``_interpolate_alpha`` (transfer_functions/__init__.py:360-538)."""
    if not agg.ndim == 3:
        raise ValueError("agg must be 3D")
    dims = list(agg.dims)
    coords = _coords_of(agg)
    cats = list(coords.get(dims[-1], []))
    out_coords = {dims[1]: coords.get(dims[1]), dims[0]: coords.get(dims[0])}
    out_coords = {k: v for k, v in out_coords.items() if v is not None}
    H, W = int(agg.shape[0]), int(agg.shape[1])
    if not len(cats):
        data = np.zeros((H, W), dtype=np.uint32)
        return Image(data, dims=dims[:-1], coords=out_coords, name=name)
    if color_key is None:
        raise ValueError("Color key must be provided, with at least as many " +
                         "colors as there are categorical fields")
    if not isinstance(color_key, dict):
        color_key = dict(zip(cats, color_key))
    if len(color_key) < len(cats):
        raise ValueError(f"Insufficient colors provided ({len(color_key)}) for the categorical "
                         f"fields available ({len(cats)})")
    colors = [rgb(color_key[c]) for c in cats]
    how_code = _how_code(how)
    _check_span(span, how)
    data, code, npdt, to_host = _device_plane(agg.data)
    if code == _abi.BOOL:
        raise TypeError("numpy boolean subtract, the `-` operator, is not supported")
    if (color_baseline is not None and npdt.kind in "iu"
            and isinstance(color_baseline, (float, np.floating))):
        raise TypeError(f"Cannot cast ufunc 'subtract' output from dtype('float64') to "
                        f"dtype('{npdt}') with casting rule 'same_kind'")
    t = dev.torch()
    out = t.empty(H * W, dtype=t.int32, device=data.device)
    d = _abi.ShadeDesc()
    d.kind, d.dtype, d.H, d.W, d.ncat = _abi.SHADE_CATEGORICAL, code, H, W, len(cats)
    d.data, d.out = dev.ptr(data), dev.ptr(out)
    d.how = how_code
    pal = _palette(colors, data.device)
    d.ncolors, d.colors = len(colors), dev.ptr(pal)
    d.alpha, d.min_alpha = float(alpha), float(min_alpha)
    _span_fields(d, span, how)
    d.rescale_discrete_levels = int(bool(rescale_discrete_levels))
    if color_baseline is not None:
        d.has_baseline, d.baseline = 1, float(color_baseline)
    _run(d, data.device)
    return _new_image(out, H, W, to_host, dims[:-1], out_coords, name)


def _key_value(k):
    """A colour-key key as the float it compares equal to (NaN: never equal)."""
    if isinstance(k, (bool, np.bool_, int, float, np.integer, np.floating)):
        return float(k)
    return float("nan")


def _apply_discrete_colorkey(agg, color_key, alpha, name, color_baseline):
    """``_apply_discrete_colorkey`` (transfer_functions/__init__.py:541-613)."""
    if not agg.ndim == 2:
        raise ValueError("agg must be 2D")
    if (color_key is None) or (not isinstance(color_key, dict)):
        raise ValueError("Color key must be provided as a dictionary")
    cats = list(color_key.keys())
    colors = [rgb(color_key[c]) for c in cats]
    data, code, _, to_host = _device_plane(agg.data)
    H, W = (int(s) for s in agg.shape)
    t = dev.torch()
    out = t.empty(H * W, dtype=t.int32, device=data.device)
    keys = t.tensor([_key_value(k) for k in cats] or [float("nan")], dtype=t.float64,
                    device=data.device)
    d = _abi.ShadeDesc()
    d.kind, d.dtype, d.H, d.W, d.ncat = _abi.SHADE_COLOR_KEY, code, H, W, 1
    d.data, d.out = dev.ptr(data), dev.ptr(out)
    pal = _palette(colors, data.device)
    d.ncolors, d.colors, d.keys = len(colors), dev.ptr(pal), dev.ptr(keys)
    d.alpha = float(alpha)
    if color_baseline is not None:
        d.has_baseline, d.baseline = 1, float(color_baseline)
    _run(d, data.device)
    return _new_image(out, H, W, to_host, list(agg.dims), _coords_of(agg), name)


def shade(agg, cmap=["lightblue", "darkblue"], color_key=Sets1to3,  # noqa: B006
          how="eq_hist", alpha=255, min_alpha=40, span=None, name=None,
          color_baseline=None, rescale_discrete_levels=False):
    """Colour an aggregate (``transfer_functions/__init__.py:616-747``).

    2D aggregates are colour-mapped through ``cmap`` (a list of colours, or one colour
    whose alpha follows the value), or mapped category by category through a ``dict``
    ``color_key``; 3D aggregates mix the ``color_key`` colours of their categories by
    value, with alpha from the total.  ``how`` is ``'eq_hist'``, ``'log'``, ``'cbrt'`` or
    ``'linear'``; ``span`` fixes the normalisation range.
    """
    if not _is_dataarray(agg):
        raise TypeError("agg must be instance of DataArray")
    name = agg.name if name is None else name
    if not ((0 <= min_alpha <= 255) and (0 <= alpha <= 255)):
        raise ValueError(f"min_alpha ({min_alpha}) and alpha ({alpha}) must be between 0 and 255")
    if rescale_discrete_levels and how != "eq_hist":
        rescale_discrete_levels = False
    if agg.ndim == 2:
        if color_key is not None and isinstance(color_key, dict):
            return _apply_discrete_colorkey(agg, color_key, alpha, name, color_baseline)
        return _interpolate(agg, cmap, how, alpha, span, min_alpha, name,
                            rescale_discrete_levels)
    if agg.ndim == 3:
        return _colorize(agg, color_key, how, alpha, span, min_alpha, name, color_baseline,
                         rescale_discrete_levels)
    raise ValueError("agg must use 2D or 3D coordinates")

"""The C ABI: libdsx.so loads, exports every function include/dsx.h declares, the ctypes
structures match the header's layout, and descriptor validation fails with the
documented status codes before any kernel is launched.  CPU only."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from datashader_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dsx.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dsx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    lib = _abi.load()
    declared = header_functions()
    assert set(declared) == set(_abi.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dsx_[a-z0-9_]+)", out))
    missing = set(declared) - exported
    assert not missing, missing
    for f in declared:
        assert hasattr(lib, f)
    assert lib.dsx_abi_version() == _abi.ABI_VERSION


STRUCTS = {
    "dsx_planes": _abi.Planes, "dsx_view": _abi.View, "dsx_categorizer": _abi.Categorizer,
    "dsx_points_desc": _abi.PointsDesc, "dsx_lines_desc": _abi.LinesDesc,
    "dsx_finalize_desc": _abi.FinalizeDesc, "dsx_grid_index_desc": _abi.GridIndexDesc,
    "dsx_gather_desc": _abi.GatherDesc, "dsx_grid_ref": _abi.GridRef,
    "dsx_shade_desc": _abi.ShadeDesc, "dsx_combine_desc": _abi.CombineDesc,
    "dsx_rowindex_desc": _abi.RowIndexDesc, "dsx_take_desc": _abi.TakeDesc,
    "dsx_nselect_desc": _abi.NSelectDesc, "dsx_batch_req": _abi.BatchReq,
    "dsx_batch_config": _abi.BatchConfig,
}


def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {}
    for line in out.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_status_strings():
    lib = _abi.load()
    for code in range(-7, 1):
        assert lib.dsx_strerror(code)
    with pytest.raises(_abi.DsxError, match="unsupported dtype"):
        _abi.check(-2, "x")


def _view(W=4, H=4):
    v = _abi.View()
    v.sx, v.tx, v.sy, v.ty = 1.0, 0.0, 1.0, 0.0
    v.xmin, v.xmax, v.ymin, v.ymax = 0.0, 4.0, 0.0, 4.0
    v.width, v.height = W, H
    return v


def test_points_descriptor_validation():
    """Invalid descriptors are rejected on the host, before any HIP call."""
    lib = _abi.load()
    pl = _abi.Planes()
    d = _abi.PointsDesc()
    d.n = 10
    ws, wsz = ctypes.c_void_p(0x1000), 1 << 40

    def call(desc, view=None, planes=pl, mean=None, work=ws, size=wsz):
        return lib.dsx_points_agg(desc, view or _view(), planes, mean, work, size, None)

    assert lib.dsx_points_agg(None, _view(), pl, None, ws, wsz, None) == -1
    assert call(d) == -1                                            # x / y NULL
    d.x = d.y = ctypes.c_void_p(0x1000)
    d.xy_dtype = 7
    assert call(d) == -2                                            # coordinates must be float
    d.xy_dtype = _abi.F32
    assert call(d, _view(0, 4)) == -3                               # empty canvas
    d.n = -1
    assert call(d) == -3
    d.n = 10
    assert call(d) == -6                                            # no plane requested
    pl.countv = ctypes.c_void_p(0x2000)
    assert call(d) == -6                                            # countv without a value
    pl = _abi.Planes()
    pl.count = ctypes.c_void_p(0x2000)
    assert call(d, planes=pl, work=None) == -4                      # workspace missing
    assert call(d, planes=pl, size=16) == -4                        # workspace too small
    assert lib.dsx_points_workspace_size(d, _view(), pl) > 0
    d.strategy = 3
    assert call(d, planes=pl) == -7                                 # unknown strategy
    assert lib.dsx_points_workspace_size(d, _view(), pl) == 0
    # AUTO needs the workspace of either strategy (their buffers share the bytes)
    d.n = 50_000_000
    sizes = {}
    for s in (_abi.STRATEGY_AUTO, _abi.STRATEGY_WINDOW, _abi.STRATEGY_TILES):
        d.strategy = s
        sizes[s] = lib.dsx_points_workspace_size(d, _view(1024, 1024), pl)
    assert sizes[_abi.STRATEGY_AUTO] == max(sizes[_abi.STRATEGY_WINDOW], sizes[_abi.STRATEGY_TILES])
    assert sizes[_abi.STRATEGY_TILES] < sizes[_abi.STRATEGY_WINDOW]
    d.strategy, d.n = _abi.STRATEGY_AUTO, 10
    d.val, d.val_dtype = ctypes.c_void_p(0x3000), _abi.F32
    assert call(d, planes=pl, mean=ctypes.c_void_p(0x5000)) == -6   # m2 mode takes sum only
    d.cat.kind = _abi.CAT_CODES
    assert call(d, planes=pl) == -1                                 # categorizer column NULL
    d.cat.col, d.cat.dtype, d.cat.ncat = ctypes.c_void_p(0x4000), _abi.F32, 3
    assert call(d, planes=pl) == -2                                 # codes must be integers
    d.cat.kind, d.cat.dtype, d.cat.modulo = _abi.CAT_MODULO, _abi.I64, 0
    assert call(d, planes=pl) == -3                                 # modulo 0
    d.cat.kind = 9
    assert call(d, planes=pl) == -7
    assert lib.dsx_points_workspace_size(d, _view(), pl) == 0


def test_lines_and_misc_validation():
    lib = _abi.load()
    pl = _abi.Planes()
    d = _abi.LinesDesc()
    d.nlines, d.nverts, d.x_dtype, d.y_dtype = 1, 3, _abi.F64, _abi.F64
    d.axis = 2
    assert lib.dsx_lines_agg(d, _view(), pl, None) == -3
    d.axis = 0
    d.x_dtype = 5
    assert lib.dsx_lines_agg(d, _view(), pl, None) == -2
    d.x_dtype = _abi.F64
    assert lib.dsx_lines_agg(d, _view(), pl, None) == -1            # x / y NULL
    d.nverts = 1
    assert lib.dsx_lines_agg(d, _view(), pl, None) == 0             # no segment
    pl.sum = ctypes.c_void_p(0x2000)
    assert lib.dsx_lines_agg(d, _view(), pl, None) == -6            # sum without a value
    assert lib.dsx_fill(None, _abi.FILL_ZERO32, 5, None) == -1
    assert lib.dsx_fill(ctypes.c_void_p(0x1000), 99, 5, None) == -7
    assert lib.dsx_fill(None, _abi.FILL_ZERO32, 0, None) == 0
    f = _abi.FinalizeDesc()
    f.mode, f.ncells, f.out = 42, 4, ctypes.c_void_p(0x1000)
    assert lib.dsx_finalize(f, None) == -7
    f.mode = _abi.FIN_MEAN
    assert lib.dsx_finalize(f, None) == -1                          # countv / sum NULL
    f.mode, f.key, f.key_bytes = _abi.FIN_MIN, ctypes.c_void_p(0x1000), 3
    assert lib.dsx_finalize(f, None) == -2
    assert lib.dsx_bounds(None, _abi.F32, 4, ctypes.c_void_p(0x1000), None, 0, None) == -1
    assert lib.dsx_bounds(ctypes.c_void_p(0x1000), 9, 4, ctypes.c_void_p(0x1000), None, 0,
                          None) == -2


def test_serving_validation():
    """Host-side validation of the grid index and gather descriptors (nothing launched)."""
    lib = _abi.load()
    g = _abi.GridIndexDesc()
    assert lib.dsx_grid_index_build(None, None, 0, None) == -1
    g.n, g.grid, g.dtype = 0, 16, _abi.F32
    assert lib.dsx_grid_index_build(g, None, 0, None) == -3          # no rows
    g.n = 10
    g.dtype = _abi.I32
    assert lib.dsx_grid_index_build(g, None, 0, None) == -2
    g.dtype = _abi.F32
    assert lib.dsx_grid_index_build(g, None, 0, None) == -1          # columns NULL
    for f in ("x", "y", "out_x", "out_y", "cell_start"):
        setattr(g, f, ctypes.c_void_p(0x1000))
    g.val = ctypes.c_void_p(0x1000)
    assert lib.dsx_grid_index_build(g, None, 0, None) == -1          # val without out_val
    g.val = None
    g.x0, g.x1, g.y0, g.y1 = 0.0, 0.0, 0.0, 1.0
    assert lib.dsx_grid_index_build(g, None, 0, None) == -3          # empty extent
    g.x1, g.grid = 1.0, 5000
    assert lib.dsx_grid_index_build(g, None, 0, None) == -3          # grid too fine
    g.grid = 16
    assert lib.dsx_grid_index_build(g, None, 0, None) == -4          # no workspace
    d = _abi.GatherDesc()
    assert lib.dsx_gather_rows(None, None) == -1
    d.ncols = 0
    assert lib.dsx_gather_rows(d, None) == -3
    d.ncols, d.nitems = 1, 0
    assert lib.dsx_gather_rows(d, None) == 0                        # nothing to copy
    d.nitems = 2
    assert lib.dsx_gather_rows(d, None) == -1                       # items NULL
    d.items = ctypes.c_void_p(0x1000)
    d.src[0] = d.dst[0] = 0x2000
    d.elem_bytes[0] = 2
    assert lib.dsx_gather_rows(d, None) == -2


def test_indexed_validation():
    lib = _abi.load()
    d = _abi.PointsDesc()
    d.n, d.xy_dtype = 10, _abi.F32
    d.x = d.y = ctypes.c_void_p(0x1000)
    pl = _abi.Planes()
    pl.count = ctypes.c_void_p(0x2000)
    v = _view()
    g = _abi.GridRef()
    g.grid, g.x0, g.x1, g.y0, g.y1 = 64, 0.0, 1.0, 0.0, 1.0

    def call(**kw):
        return lib.dsx_points_agg_indexed(d, kw.get("view", v), pl, None, kw.get("g", g),
                                          None, 0, None)
    assert call() == -1                                             # cell_start NULL
    g.cell_start = ctypes.c_void_p(0x3000)
    assert call() == -4                                             # no workspace
    g.grid = 2048
    assert call() == -3                                             # lattice too fine
    g.grid = 64
    v.xlog = 1
    assert call() == -7                                             # log axis
    v.xlog = 0
    d.cat.kind, d.cat.dtype, d.cat.ncat, d.cat.col = _abi.CAT_CODES, _abi.I8, 4, 0x4000
    assert call() == -7                                             # categorizer
    d.cat.kind, d.cat.ncat = _abi.CAT_NONE, 1
    pl.sum = ctypes.c_void_p(0x5000)
    assert call() == -6                                             # sum without a value
    assert lib.dsx_points_indexed_workspace_size(d, v, pl, g) == 0

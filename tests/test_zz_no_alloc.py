"""dsx_bounds_noalloc (the autorange bounds without a workspace, for callers that cannot
allocate) called through the C ABI gives exactly what dsx_bounds gives, on two calls in a
row on one GPU, the second over a smaller column than the first (the array inside the
library keeps the first call's partials, so stale partials would show)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bounds_with_workspace(lib, col, dtype):
    n = col.numel()
    ws = int(lib.dsx_bounds_workspace_size(n))
    work = torch.empty(max(ws, 16), dtype=torch.uint8, device=col.device)
    out = torch.full((2,), 7.0, dtype=torch.float64, device=col.device)
    st = lib.dsx_bounds(ctypes.c_void_p(col.data_ptr()), dtype, n,
                        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(work.data_ptr()), ws,
                        None)
    assert st == 0, st
    return out


def _bounds_noalloc(lib, col, dtype):
    out = torch.full((2,), 7.0, dtype=torch.float64, device=col.device)
    st = lib.dsx_bounds_noalloc(ctypes.c_void_p(col.data_ptr()), dtype, col.numel(),
                                ctypes.c_void_p(out.data_ptr()), None)
    assert st == 0, st
    return out


@pytest.mark.parametrize("tdtype,code", [(torch.float32, 1), (torch.float64, 2)])
def test_noalloc_bounds_equal_workspace_bounds(tdtype, code):
    from datashader_amd import _abi
    lib = _abi.load()
    torch.cuda.set_device(0)
    rng = np.random.default_rng(11)
    big = rng.normal(0.0, 3.0, 5_000_003)
    big[rng.integers(0, big.size, 1000)] = np.nan
    big[17] = -40.0     # the extremes sit in the first call only: a second call that
    big[-5] = 55.0      # read a stale partial of the first would return them
    small = rng.uniform(-1.0, 2.0, 70_001)
    small[::13] = np.nan
    for host in (big, small):
        col = torch.from_numpy(host).to(tdtype).cuda()
        print(f"noalloc bounds: dtype {tdtype}, n {col.numel()}", flush=True)
        want = _bounds_with_workspace(lib, col, code)
        got = _bounds_noalloc(lib, col, code)
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        np.testing.assert_array_equal(g, w)
        ref = host.astype(np.float32 if tdtype == torch.float32 else np.float64)
        np.testing.assert_array_equal(g, [np.nanmin(ref), np.nanmax(ref)])


def test_noalloc_bounds_all_nan_and_empty():
    from datashader_amd import _abi
    lib = _abi.load()
    torch.cuda.set_device(0)
    col = torch.full((1000,), float("nan"), dtype=torch.float32, device="cuda")
    got = _bounds_noalloc(lib, col, 1).cpu().numpy()
    assert got[0] == np.inf and got[1] == -np.inf
    empty = torch.empty(0, dtype=torch.float64, device="cuda")
    got = _bounds_noalloc(lib, empty, 2).cpu().numpy()
    assert got[0] == np.inf and got[1] == -np.inf


def test_noalloc_bounds_rejects_bad_arguments():
    from datashader_amd import _abi
    lib = _abi.load()
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    col = torch.zeros(10, dtype=torch.float32, device="cuda")
    assert lib.dsx_bounds_noalloc(ctypes.c_void_p(col.data_ptr()), 1, 10, None, None) == -1
    assert lib.dsx_bounds_noalloc(ctypes.c_void_p(col.data_ptr()), 6, 10,
                                  ctypes.c_void_p(out.data_ptr()), None) == -2
    assert lib.dsx_bounds_noalloc(ctypes.c_void_p(col.data_ptr()), 1, -1,
                                  ctypes.c_void_p(out.data_ptr()), None) == -3

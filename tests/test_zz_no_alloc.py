"""The no-allocation cell sort (wsim_cell_sort_noalloc) against the
allocating grid build (wsim_grid_create + wsim_grid_build): sorted keys,
sorted indices and cell_start bit for bit, on two calls in a row, the
second with fewer particles, on one stream; and the size limit is refused.

Both go through the C ABI (ctypes).  The first call of the process looks up
the device address of the scratch array; the second reuses it, and its
smaller input must not see anything of the first call's larger one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import water_simulation_amd as ws
from water_simulation_amd import _ffi

pytestmark = pytest.mark.gpu


def cloud(n, seed, dims, cell):
    rng = np.random.default_rng(seed)
    ext = np.array(dims, np.float64) * cell
    x = rng.uniform(-0.05 * ext, 1.05 * ext, (n, 3)).astype(np.float32)  # some outside the grid
    x[: n // 10] = x[n // 10: 2 * (n // 10)]                            # duplicates
    if n > 20:
        x[5, 1] = np.nan
    return x


def via_grid(x, origin, dims, cell):
    g = ws.CellGrid(origin, dims, cell)
    buf = ws.DeviceBuffer.from_array(x)
    g.build(buf.ptr, len(x))
    ws.device_synchronize()
    n, ncs = len(x), int(np.prod(dims)) + 1
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    cs = np.empty(ncs, np.uint32)
    ws.memcpy_d2h(keys, g.sorted_keys())
    ws.memcpy_d2h(perm, g.sorted_indices())
    ws.memcpy_d2h(cs, g.cell_start())
    return keys, perm, cs


def via_noalloc(x, origin, dims, cell, outs, stream):
    n, ncs = len(x), int(np.prod(dims)) + 1
    buf = ws.DeviceBuffer.from_array(x)
    kd, pd, cd = outs
    st = _ffi.lib().wsim_cell_sort_noalloc(buf.ptr, n, _ffi.F3(*origin), _ffi.I3(*dims),
                                           float(cell), kd.ptr, pd.ptr, cd.ptr, stream)
    _ffi.check(st)
    _ffi.check(_ffi.lib().wsim_stream_synchronize(stream))
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    cs = np.empty(ncs, np.uint32)
    ws.memcpy_d2h(keys, kd.ptr)
    ws.memcpy_d2h(perm, pd.ptr)
    ws.memcpy_d2h(cs, cd.ptr)
    return keys, perm, cs


def test_noalloc_cell_sort_matches_grid_build_twice():
    origin, dims, cell = (-0.01, -0.01, -0.01), (82, 92, 72), 0.01
    ncs = int(np.prod(dims)) + 1
    n1, n2 = 1_500_000, 70_001
    # outputs sized for the first (larger) call and reused by the second
    outs = (ws.DeviceBuffer(4 * n1), ws.DeviceBuffer(4 * n1), ws.DeviceBuffer(4 * ncs))
    s = C.c_void_p()
    _ffi.check(_ffi.lib().wsim_stream_create(C.byref(s)))
    try:
        for n, seed in ((n1, 1), (n2, 2)):
            x = cloud(n, seed, dims, cell)
            want = via_grid(x, origin, dims, cell)
            got = via_noalloc(x, origin, dims, cell, outs, s.value)
            for name, g, w in zip(("sorted keys", "sorted indices", "cell_start"), got, want):
                bad = np.nonzero(g != w)[0]
                assert bad.size == 0, f"n={n} {name}: {bad.size} differ, first at {bad[0]}"
            # the stable (key, original index) order itself
            keys_orig = np.empty(n, np.uint32)
            keys_orig[got[1]] = got[0]
            assert np.array_equal(got[1], np.lexsort((np.arange(n), keys_orig)))
    finally:
        _ffi.lib().wsim_stream_destroy(s.value)


def test_noalloc_cell_sort_refuses_too_many_particles():
    origin, dims, cell = (0.0, 0.0, 0.0), (4, 4, 4), 0.25
    n = (1 << 21) + 1
    st = _ffi.lib().wsim_cell_sort_noalloc(None, n, _ffi.F3(*origin), _ffi.I3(*dims), cell,
                                           None, None, None, None)
    assert st == 3  # WSIM_ERR_INVALID_ARGUMENT, before any launch
    assert "exceed the maximum" in _ffi.lib().wsim_last_error().decode()

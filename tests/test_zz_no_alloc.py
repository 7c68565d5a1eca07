"""The allocation-free grid entry point (aleator_gpu_grid_build_device_noalloc; no counterpart in
the reference) called through the C ABI against the allocating one
(aleator_gpu_grid_build_device): the same kernels on the same inputs, so the tables agree bit for
bit -- on two calls in a row, the second with a smaller input -- and a build whose scratch would
exceed the fixed 32 MiB array is refused with ALEATOR_INVALID_ARGUMENT before any device work."""
import ctypes as C

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework

pytestmark = pytest.mark.gpu

ENTRY = ("aleator_gpu_grid_build_device", "aleator_gpu_grid_build_device_noalloc")


def _lib():
    lib = C.CDLL(aa.library_path())
    for name in ENTRY:
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_int32, C.c_int32,
                      C.POINTER(C.c_int32), C.c_int32, C.c_double, C.c_void_p]
    lib.aleator_gpu_last_error.restype = C.c_char_p
    return lib


def _build(lib, name, fw, grid_m, guests, coulomb, counts, precision):
    lat = (C.c_double * 9)(*np.ascontiguousarray(grid_m, dtype=np.float64).ravel())
    g = (C.c_uint32 * len(guests))(*guests)
    n = (C.c_int32 * 3)(*counts)
    tables = len(guests) + (1 if coulomb else 0)
    buf = aa.DeviceBuffer(tables * int(np.prod(counts)))
    rc = getattr(lib, name)(fw.handle, C.byref(lat), g, len(guests), int(coulomb), n, precision, C.c_double(1e5),
                            buf.ptr)
    out = buf.to_host() if rc == 0 else None
    buf.free()
    return rc, out


def test_noalloc_grid_equals_the_allocating_entry_point(ctx):
    print("no-alloc grid: c1 three tables (131^3, MIXED) then 0.3 A (FP64)", flush=True)
    lib = _lib()
    prim, _ = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    _, lj = ol.c1_system()
    fw = device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))
    for spacing, precision in ((0.15, aa.MIXED), (0.3, aa.FP64)):
        counts = ol.grid_counts(prim.m, spacing)
        rc0, ref = _build(lib, ENTRY[0], fw, prim.m, [4, 5], True, counts, precision)
        rc1, got = _build(lib, ENTRY[1], fw, prim.m, [4, 5], True, counts, precision)
        assert rc0 == 0 and rc1 == 0, lib.aleator_gpu_last_error()
        print(f"  {counts} precision {precision}: {got.size} values, bitwise equal {np.array_equal(got, ref)}")
        assert np.array_equal(got, ref)


def test_noalloc_grid_refuses_builds_above_its_limit(ctx):
    lib = _lib()
    prim, _ = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    _, lj = ol.c1_system()
    fw = device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))
    # the reciprocal table alone is 17 x 600 x 600 x 16 B = 98 MB > 32 MiB
    rc, _ = _build(lib, ENTRY[1], fw, prim.m, [4], True, (64, 600, 600), aa.MIXED)
    assert rc == 1  # ALEATOR_INVALID_ARGUMENT
    assert b"allocation-free limit" in lib.aleator_gpu_last_error()

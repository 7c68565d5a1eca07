"""Device FrameworkEnergyGrid::interpolate (framework_energy_grid.cc:251-284) against the oracle's
restatement (itself equal to the reference's interpolate through the cache-file tests), and a
device-built grid written in cache format v1 and loaded back bit for bit, whose interpolate
equals the oracle's on the loaded table.  That files written by this writer load in the
reference's own loader bit for bit is tested on the CPU (test_grid_cache.py), where the compiled
reference exists."""
import ctypes as C
import tempfile

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework
from oracle_lib import i32p, ptr
from test_grid_cache import key, system

pytestmark = pytest.mark.gpu


def oracle_interpolate(m, counts, table, pos):
    lat = ol.lattice(m)
    n = np.ascontiguousarray(counts, dtype=np.int32)
    t = np.ascontiguousarray(table)
    return np.array([ol.oracle().orc_grid_interpolate(C.byref(lat), ptr(n, i32p), ptr(t), ptr(p.copy()))
                     for p in pos])


@pytest.mark.parametrize("triclinic", [False, True])
def test_device_interpolate_equals_oracle(ctx, triclinic):
    print(f"device interpolate, triclinic={triclinic}", flush=True)
    fw, lj = system(triclinic=triclinic)
    counts = ol.grid_counts(fw.m, 0.6)
    table = ol.grid_build(lj, fw, [4], counts)[0]
    rng = np.random.default_rng(7)
    pos = np.concatenate([rng.uniform(-20.0, 35.0, size=(3000, 3)),
                          (np.indices((3, 3, 3)).reshape(3, -1).T / 3.0) @ fw.m])  # some on nodes exactly
    got = aa.grid_interpolate(ctx, fw.m, counts, table, pos)
    assert np.array_equal(got, oracle_interpolate(fw.m, counts, table, pos))


def test_device_grid_round_trips_through_the_cache_file(ctx):
    print("device grid -> cache v1 -> loader", flush=True)
    fw, lj = system(triclinic=False)  # the device requires rc <= L_perp / 2
    dev = device_framework(ctx, fw, lj)
    counts = aa.grid_point_counts(fw.m, 0.7)
    tables = dev.grid([4, 5], counts, precision=aa.FP64)
    k = key(fw, lj, [4, 5], 0.7)
    rng = np.random.default_rng(8)
    pos = rng.uniform(-10.0, 25.0, size=(1000, 3))
    with tempfile.TemporaryDirectory() as d:
        aa.grid_cache_serialize(k.file_path(d), k, counts, [4, 5], tables)
        n, sp, loaded = aa.grid_cache_load(d, k)
        assert n == tuple(counts) and list(sp) == [4, 5]
        assert np.array_equal(loaded, tables)
        for slot in range(2):
            got = aa.grid_interpolate(ctx, fw.m, counts, loaded[slot], pos)
            assert np.array_equal(got, oracle_interpolate(fw.m, counts, loaded[slot], pos))


def test_interpolate_errors(ctx):
    with pytest.raises(ValueError):
        aa.grid_interpolate(ctx, np.diag([10.0, 10.0, 10.0]), (0, 4, 4), np.zeros(16), np.zeros((2, 3)))
    with pytest.raises(ValueError):
        aa.grid_interpolate(ctx, np.zeros((3, 3)), (4, 4, 4), np.zeros(64), np.zeros((2, 3)))

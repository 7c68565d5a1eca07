"""The energy-grid cache file, format v1, against the compiled reference (oracle/_ref):
the reference's own key hash, reference-written files loading here bit for bit, files written
here loading in the reference's FrameworkEnergyGrid::loadFromCache (checked through its
interpolate, which must equal the oracle's interpolate of the tables written), and the
loader's error semantics (framework_energy_grid.cc:357-408; test_framework_energy_grid_cache.cc
covers the same cases in the reference).  Host code only: no device needed."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from oracle_lib import i32p, ptr, u32p

pytestmark = pytest.mark.skipif(not ol.have_ref(), reason="oracle/_ref not built")


def system(trunc=ol.SHIFTED, triclinic=True):
    m = ol.cell_from_params(12.0, 13.0, 12.5, 60.0, 70.0, 80.0) if triclinic else np.diag([12.0] * 3)
    fw = ol.synth_framework(m, 70, seed=5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    return fw, ol.LJ(eps, sig, 5.5, trunc)


def key(fw, lj, guests, spacing):
    return aa.GridCacheKey(fw.m, fw.x, fw.y, fw.z, fw.species, lj.eps, lj.sigma, lj.cutoff, lj.truncation, guests,
                           spacing)


def ref_args(fw, lj, guests, spacing):
    g = np.ascontiguousarray(guests, dtype=np.uint32)
    keep = (np.ascontiguousarray(fw.m), g)
    return keep, (C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff), lj.truncation,
                  lj.mixing, ptr(keep[0]), C.c_uint64(fw.n), ptr(fw.x), ptr(fw.y), ptr(fw.z),
                  ptr(fw.species, u32p), ptr(g, u32p), len(g), C.c_double(spacing))


def ref_hash(fw, lj, guests, spacing):
    keep, args = ref_args(fw, lj, guests, spacing)
    h = C.c_uint64()
    assert ol.ref().ref_grid_cache_hash(*args, C.byref(h)) == 0
    return h.value


def ref_load_interpolate(fw, lj, guests, spacing, directory, guest, pos):
    keep, args = ref_args(fw, lj, guests, spacing)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(len(pos))
    counts = np.zeros(3, dtype=np.int32)
    hit = C.c_int()
    rc = ol.ref().ref_grid_cache_load_interpolate(*args, directory.encode(), C.c_uint32(guest), ptr(pos),
                                                 C.c_uint64(len(pos)), ptr(out), ptr(counts, i32p), C.byref(hit))
    return rc, hit.value, tuple(counts), out


@pytest.mark.parametrize("guests", [[4], [5, 4, 5], [4, 5, 0]])
@pytest.mark.parametrize("trunc", [ol.TRUNCATED, ol.LFS])
def test_key_hash_and_file_name_equal_the_reference(guests, trunc):
    fw, lj = system(trunc)
    k = key(fw, lj, guests, 0.7)
    assert k.hash() == ref_hash(fw, lj, guests, 0.7)
    assert os.path.basename(k.file_path("/x")) == f"framework_energy_grid_v1_{ref_hash(fw, lj, guests, 0.7):016x}.bin"
    # any change of a key field changes the hash
    assert key(fw, lj, guests, 0.71).hash() != k.hash()
    lj2 = ol.LJ(lj.eps.copy(), lj.sigma, lj.cutoff, lj.truncation)
    lj2.eps[0] += 1e-12
    assert key(fw, lj2, guests, 0.7).hash() != k.hash()


def test_reference_written_file_loads_here_bit_for_bit():
    fw, lj = system()
    with tempfile.TemporaryDirectory() as d:
        g = np.array([4, 5], dtype=np.uint32)
        counts = np.zeros(3, dtype=np.int32)
        cap = 64 ** 3 * 2
        out = np.zeros(cap)
        rc = ol.ref().ref_grid_build(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                                     lj.truncation, lj.mixing, ptr(np.ascontiguousarray(fw.m)), C.c_uint64(fw.n),
                                     ptr(fw.x), ptr(fw.y), ptr(fw.z), ptr(fw.species, u32p), ptr(g, u32p), 2,
                                     C.c_double(0.9), d.encode(), ptr(counts, i32p), ptr(out), C.c_uint64(cap))
        assert rc == 0
        n = int(np.prod(counts))
        loaded = aa.grid_cache_load(d, key(fw, lj, [4, 5], 0.9))
        assert loaded is not None
        npts, species, tables = loaded
        assert npts == tuple(counts)
        assert list(species) == [4, 5]
        assert np.array_equal(tables.reshape(2, -1), out[:2 * n].reshape(2, n))
        # and the oracle built the same tables
        assert np.array_equal(tables, ol.grid_build(lj, fw, [4, 5], npts))


def test_file_written_here_loads_in_the_reference():
    fw, lj = system()
    counts = ol.grid_counts(fw.m, 0.8)
    tables = ol.grid_build(lj, fw, [4, 5], counts)
    rng = np.random.default_rng(3)
    pos = rng.uniform(-15.0, 30.0, size=(300, 3))
    k = key(fw, lj, [5, 4], 0.8)
    with tempfile.TemporaryDirectory() as d:
        aa.grid_cache_serialize(k.file_path(d), k, counts, [5, 4], tables[::-1].copy())
        for slot, guest in enumerate([4, 5]):
            rc, hit, rcounts, got = ref_load_interpolate(fw, lj, [5, 4], 0.8, d, guest, pos)
            assert rc == 0 and hit == 1 and rcounts == tuple(counts)
            lat = ol.lattice(fw.m)
            n = np.ascontiguousarray(counts, dtype=np.int32)
            t = np.ascontiguousarray(tables[slot])
            expect = np.array([ol.oracle().orc_grid_interpolate(C.byref(lat), ptr(n, i32p), ptr(t), ptr(p.copy()))
                               for p in pos])
            assert np.array_equal(got, expect)
        # round trip through this loader
        npts, species, back = aa.grid_cache_load(d, k)
        assert npts == tuple(counts) and list(species) == [4, 5] and np.array_equal(back, tables)


def test_loader_error_semantics():
    fw, lj = system()
    counts = (5, 6, 7)
    tables = np.arange(2 * 210, dtype=np.float64).reshape(2, 7, 6, 5)
    k = key(fw, lj, [4, 5], 0.8)
    with tempfile.TemporaryDirectory() as d:
        assert aa.grid_cache_load(d, k) is None  # a miss is not an error
        path = k.file_path(d)
        aa.grid_cache_serialize(path, k, counts, [4, 5], tables)
        raw = open(path, "rb").read()
        # version mismatch
        open(path, "wb").write((2).to_bytes(4, "little") + raw[4:])
        with pytest.raises(RuntimeError, match="format version"):
            aa.grid_cache_load(d, k)
        assert ref_load_interpolate(fw, lj, [4, 5], 0.8, d, 4, np.zeros((1, 3)))[0] != 0
        # truncated
        open(path, "wb").write(raw[:-9])
        with pytest.raises(RuntimeError, match="truncated"):
            aa.grid_cache_load(d, k)
        assert ref_load_interpolate(fw, lj, [4, 5], 0.8, d, 4, np.zeros((1, 3)))[0] != 0
        # a foreign key under this key's file name (a collision or a tampered directory)
        other = key(fw, lj, [4, 5], 0.81)
        aa.grid_cache_serialize(path, other, counts, [4, 5], tables)
        with pytest.raises(RuntimeError, match="stored key"):
            aa.grid_cache_load(d, k)
        assert ref_load_interpolate(fw, lj, [4, 5], 0.8, d, 4, np.zeros((1, 3)))[0] != 0
    with pytest.raises(ValueError):
        aa.grid_cache_serialize("/tmp/x.bin", k, counts, [4, 4], tables)  # duplicate species

"""Bit-for-bit agreement of the oracle with the reference's own hot-path code, compiled from
its sources into oracle/_ref (oracle/Makefile target `ref`).  Skipped where that build is
absent."""
import ctypes as C
import tempfile

import numpy as np
import pytest

import oracle_lib as ol
from oracle_lib import i32p, i64p, ptr, u32p

pytestmark = pytest.mark.skipif(not ol.have_ref(), reason="oracle/_ref not built (needs the reference sources)")


def arr(m):
    return np.ascontiguousarray(m, dtype=np.float64)


def test_philox_streams_bitwise():
    r = ol.ref()
    for seed, stream in ((0, 0), (42, 3), (2 ** 63 + 5, 2 ** 40 + 1)):
        out = np.zeros(777)
        r.ref_uniforms(C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(777), ptr(out))
        assert np.array_equal(out, ol.uniforms(seed, stream, 0, 777))


@pytest.mark.parametrize("cell", [(24, 24, 24, 60, 60, 60), (11, 13, 17, 80, 100, 115), (20, 20, 20, 90, 90, 90)])
def test_minimum_image_bitwise(cell):
    m = ol.cell_from_params(*cell)
    rng = np.random.default_rng(0)
    a, b = rng.uniform(-30, 30, (400, 3)), rng.uniform(-30, 30, (400, 3))
    out = np.zeros((400, 3))
    assert ol.ref().ref_minimum_image(ptr(arr(m)), C.c_uint64(400), ptr(a), ptr(b), ptr(out)) == 0
    lat = ol.lattice(m)
    for i in range(400):
        d = np.zeros(3)
        ol.oracle().orc_minimum_image(C.byref(lat), ptr(a[i].copy()), ptr(b[i].copy()), ptr(d))
        assert np.array_equal(d, out[i])


def ref_grid(lj, fw, guests, spacing):
    g = np.array(sorted(set(guests)), dtype=np.uint32)
    counts = np.zeros(3, dtype=np.int32)
    cap = 64 ** 3 * len(g)
    out = np.zeros(cap)
    with tempfile.TemporaryDirectory() as d:
        rc = ol.ref().ref_grid_build(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                                     lj.truncation, lj.mixing, ptr(arr(fw.m)), C.c_uint64(fw.n), ptr(fw.x), ptr(fw.y),
                                     ptr(fw.z), ptr(fw.species, u32p), ptr(g, u32p), len(g), C.c_double(spacing),
                                     d.encode(), ptr(counts, i32p), ptr(out), C.c_uint64(cap))
    assert rc == 0
    n = int(np.prod(counts))
    return tuple(int(v) for v in counts), out[: n * len(g)].reshape(len(g), counts[2], counts[1], counts[0])


@pytest.mark.parametrize("trunc", [ol.TRUNCATED, ol.SHIFTED, ol.LFS])
def test_grid_build_bitwise_cubic(trunc):
    fw, lj = ol.c0_system(60, a=12.0)
    lj.cutoff, lj.truncation = 5.9, trunc
    counts, ref = ref_grid(lj, fw, [4, 1], 0.8)
    assert counts == ol.grid_counts(fw.m, 0.8)
    assert np.array_equal(ol.grid_build(lj, fw, [4, 1], counts), ref)


def test_grid_build_bitwise_triclinic():
    m = ol.cell_from_params(12.0, 13.0, 12.5, 60.0, 70.0, 80.0)
    fw = ol.synth_framework(m, 90, seed=5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 4.5)
    counts, ref = ref_grid(lj, fw, [5, 4], 0.7)
    assert np.array_equal(ol.grid_build(lj, fw, [4, 5], counts), ref)


def test_grid_nodes_bitwise_on_c1_supercell():
    prim, lj = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    counts = ol.grid_counts(prim.m, 0.15)
    nodes = np.array([0, 4321, 1234567, 2248090], dtype=np.int64)
    for guest in (4, 5):
        out = np.zeros(len(nodes))
        assert ol.ref().ref_grid_nodes_lj(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(12.0), 0, 0,
                                          ptr(arr(sc.m)), C.c_uint64(sc.n), ptr(sc.x), ptr(sc.y), ptr(sc.z),
                                          ptr(sc.species, u32p), ptr(arr(prim.m)),
                                          ptr(np.array(counts, dtype=np.int32), i32p), C.c_uint32(guest),
                                          ptr(nodes, i64p), C.c_uint64(len(nodes)), C.c_double(1e5), ptr(out)) == 0
        assert np.array_equal(out, ol.grid_nodes_lj(lj, sc, prim.m, counts, guest, nodes))
    real, rec = np.zeros(len(nodes)), np.zeros(len(nodes))
    assert ol.ref().ref_coulomb_nodes(C.c_double(0.265), C.c_double(12.0), 8, ptr(arr(sc.m)), C.c_uint64(sc.n),
                                      ptr(sc.x), ptr(sc.y), ptr(sc.z), ptr(sc.q), ptr(arr(prim.m)),
                                      ptr(np.array(counts, dtype=np.int32), i32p), ptr(nodes, i64p),
                                      C.c_uint64(len(nodes)), ptr(real), ptr(rec)) == 0
    st = ol.EwaldState(sc, 0.265, 8)
    oreal, orec = ol.coulomb_nodes(st, 12.0, sc, prim.m, counts, nodes)
    assert np.array_equal(real, oreal)
    # the reference reaches the reciprocal potential through two proposeInsertion calls
    assert np.all(np.abs(rec - orec) <= 1e-12 * np.abs(orec))


def test_reciprocal_terms_bitwise():
    m = ol.cell_from_params(24, 24, 24, 60, 60, 60) * 2
    k, w, n = np.zeros((3000, 3)), np.zeros(3000), C.c_uint64()
    assert ol.ref().ref_reciprocal_terms(ptr(arr(m)), C.c_double(0.265), 8, C.c_uint64(3000), ptr(k), ptr(w),
                                         C.byref(n)) == 0
    fw = ol.synth_framework(ol.cell_from_params(24, 24, 24, 60, 60, 60), 10, charge_sigma=0.5)
    fw.m = m
    st = ol.EwaldState(fw, 0.265, 8)
    ok, ow, _ = st.kvectors()
    assert n.value == st.nk == 2204
    assert np.array_equal(k[:2204], ok) and np.array_equal(w[:2204], ow)


def ref_state(fw, alpha, kmax):
    h = ol.ref().ref_ewald_state_create(C.c_double(alpha), kmax, ptr(arr(fw.m)), C.c_uint64(fw.n), ptr(fw.x),
                                        ptr(fw.y), ptr(fw.z), ptr(fw.q))
    assert h
    return h


def test_ewald_state_bitwise_sequence():
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    r = ol.ref()
    h = ref_state(fw, 0.35, 6)
    st = ol.EwaldState(fw, 0.35, 6)
    raw = np.zeros(2 * st.nk + 2)
    rng = np.random.default_rng(3)
    mols = []
    try:
        for step in range(25):
            c = rng.uniform(0, 15, 3)
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            mol = np.array([[*(c - 1.16 * u), -0.35], [*c, 0.70], [*(c + 1.16 * u), -0.35]])
            if mols and step % 3 == 2:
                old = mols.pop(0)
                assert r.ref_ewald_state_propose(h, ptr(old), 3, None, 0) == st.propose(old, [])
                r.ref_ewald_state_commit(h, ptr(old), 3, None, 0)
                st.commit(old, [])
            elif mols and step % 3 == 1:
                old = mols[-1]
                new = old.copy()
                new[:, :3] += 0.7
                assert r.ref_ewald_state_propose(h, ptr(old), 3, ptr(new), 3) == st.propose(old, new)
                r.ref_ewald_state_commit(h, ptr(old), 3, ptr(new), 3)
                st.commit(old, new)
                mols[-1] = new
            else:
                assert r.ref_ewald_state_propose(h, None, 0, ptr(mol), 3) == st.propose([], mol)
                r.ref_ewald_state_commit(h, None, 0, ptr(mol), 3)
                st.commit([], mol)
                mols.append(mol)
            r.ref_ewald_state_raw(h, ptr(raw), C.c_uint64(raw.size))
            assert np.array_equal(raw, st.raw())
            assert r.ref_ewald_state_committed(h) == st.committed()
        allp = np.concatenate(mols)
        x = np.concatenate([fw.x, allp[:, 0]]); y = np.concatenate([fw.y, allp[:, 1]])
        z = np.concatenate([fw.z, allp[:, 2]]); q = np.concatenate([fw.q, allp[:, 3]])
        off = np.arange(0, 3 * len(mols) + 1, 3, dtype=np.uint64)
        sites = np.arange(fw.n, fw.n + 3 * len(mols), dtype=np.uint64)
        full = ol.Framework(fw.m, x, y, z, q, np.zeros(len(x), dtype=np.uint32))
        molecules = [list(range(fw.n + 3 * i, fw.n + 3 * i + 3)) for i in range(len(mols))]
        assert r.ref_ewald_state_recompute(h, C.c_uint64(len(x)), ptr(x), ptr(y), ptr(z), ptr(q), ptr(off, ol.u64p),
                                           ptr(sites, ol.u64p), C.c_uint64(len(mols))) == st.recompute(full, molecules)
        r.ref_ewald_state_resync(h, C.c_uint64(len(x)), ptr(x), ptr(y), ptr(z), ptr(q), ptr(off, ol.u64p),
                                 ptr(sites, ol.u64p), C.c_uint64(len(mols)))
        st.resync(full, molecules)
        r.ref_ewald_state_raw(h, ptr(raw), C.c_uint64(raw.size))
        assert np.array_equal(raw, st.raw())
    finally:
        r.ref_ewald_state_destroy(h)


def ref_widom(lj, fw, sites, charged, alpha, rc_real, kmax, seed, stream, n, T):
    sx = np.array([s[:3] for s in sites], dtype=np.float64).ravel()
    ss = np.array([s[3] for s in sites], dtype=np.uint32)
    sq = np.array([s[4] for s in sites], dtype=np.float64)
    kh = C.c_double()
    assert ol.ref().ref_widom(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                              lj.truncation, lj.mixing, ptr(arr(fw.m)), C.c_uint64(fw.n), ptr(fw.x), ptr(fw.y),
                              ptr(fw.z), ptr(fw.q), ptr(fw.species, u32p), len(sites), ptr(sx), ptr(ss, u32p),
                              ptr(sq), charged, C.c_double(alpha), C.c_double(rc_real), kmax, C.c_uint64(seed),
                              C.c_uint64(stream), C.c_uint64(n), C.c_double(T), C.byref(kh)) == 0
    return kh.value


def test_widom_henry_coefficient_bitwise_methane():
    fw, lj = ol.c0_system()
    kh, _ = ol.widom(lj, fw, [(0.0, 0.0, 0.0, 4, 0.0)], seed=1, stream=1, n_trials=200, temperature=298.0)
    assert kh == ref_widom(lj, fw, [(0.0, 0.0, 0.0, 4, 0.0)], 0, 0.0, 0.0, 1, 1, 1, 200, 298.0)


def test_widom_henry_coefficient_bitwise_charged_co2():
    m = ol.cell_from_params(20.0, 21.0, 22.0, 70.0, 80.0, 95.0)
    fw = ol.synth_framework(m, 250, seed=7, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 8.0)
    st = ol.EwaldState(fw, 0.35, 6)
    kh, _ = ol.widom(lj, fw, ol.co2_sites(), seed=2, stream=3, n_trials=60, temperature=298.0, ewald=st, rc_real=8.0)
    assert kh == ref_widom(lj, fw, ol.co2_sites(), 1, 0.35, 8.0, 6, 2, 3, 60, 298.0)


def test_tail_and_pair_energy_bitwise():
    fw, lj = ol.c0_system()
    out = C.c_double()
    for idx in (0, 17, 423):
        assert ol.ref().ref_lj_particle_energy(C.c_uint32(5), ptr(lj.eps), ptr(lj.sigma), C.c_double(12.8), 0, 0,
                                               ptr(arr(fw.m)), C.c_uint64(fw.n), ptr(fw.x), ptr(fw.y), ptr(fw.z),
                                               ptr(fw.species, u32p), C.c_uint64(idx), C.byref(out)) == 0
        l, keep = ol.lj_struct(lj)
        p, keep2 = ol.particles(fw)
        assert out.value == ol.oracle().orc_lj_particle_energy(C.byref(l), C.c_size_t(idx), C.byref(p),
                                                               C.byref(ol.lattice(fw.m)), None, C.c_size_t(0))
    assert ol.ref().ref_lj_tail(C.c_uint32(5), ptr(lj.eps), ptr(lj.sigma), C.c_double(12.8), 0, 0, ptr(arr(fw.m)),
                                C.c_uint64(fw.n), ptr(fw.species, u32p), C.byref(out)) == 0
    l, keep = ol.lj_struct(lj)
    p, keep2 = ol.particles(fw)
    assert out.value == ol.oracle().orc_lj_tail_energy(C.byref(l), C.byref(p), C.byref(ol.lattice(fw.m)))


@pytest.mark.parametrize("fugacity", [1e6, 1e7, 1e8])
def test_gcmc_methane_trajectory_bitwise(fugacity):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    g = ol.Gcmc(lj, fw, sites, fugacity, 300.0, 5, 0)
    g.run(4000)
    n_ref, parts, _ = ol.ref_gcmc_run(lj, fw, sites, fugacity, 300.0, 5, 0, 4000)
    assert g.n_molecules == n_ref
    assert all(np.array_equal(a, b) for a, b in zip(g.particles(), parts))


def test_gcmc_charged_co2_trajectory_bitwise():
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    g = ol.Gcmc(lj, fw, ol.co2_sites(), 1e6, 298.0, 3, 1, ewald=(0.35, 7.5, 5))
    g.run(1500)
    n_ref, parts, e_ref = ol.ref_gcmc_run(lj, fw, ol.co2_sites(), 1e6, 298.0, 3, 1, 1500, ewald=(0.35, 7.5, 5))
    assert g.n_molecules == n_ref and n_ref > 0
    assert all(np.array_equal(a, b) for a, b in zip(g.particles(), parts))
    assert g.electrostatic_energy() == e_ref


@pytest.mark.parametrize("fugacity", [1e7, 1e8])
def test_gcmc_methane_with_energy_grid_trajectory_bitwise(fugacity, tmp_path):
    """The engine given a FrameworkEnergyGrid (framework LJ interpolated, guests scanned): the
    oracle with the reference's own tables follows the reference's trajectory bit for bit."""
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    n_ref, parts, _ = ol.ref_gcmc_run_grid(lj, fw, sites, fugacity, 300.0, 5, 0, 3000, 0.5, [4], tmp_path)
    counts, tables = ref_grid(lj, fw, [4], 0.5)
    g = ol.Gcmc(lj, fw, sites, fugacity, 300.0, 5, 0)
    g.set_grid(counts, [4], tables)
    g.run(3000)
    assert g.n_molecules == n_ref and n_ref > 0
    assert all(np.array_equal(a, b) for a, b in zip(g.particles(), parts))
    direct = ol.Gcmc(lj, fw, sites, fugacity, 300.0, 5, 0)
    direct.run(3000)
    assert not all(np.array_equal(a, b) for a, b in zip(direct.particles(), parts))  # the grid changed the run


def test_gcmc_charged_co2_with_energy_grid_trajectory_bitwise(tmp_path):
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    n_ref, parts, e_ref = ol.ref_gcmc_run_grid(lj, fw, ol.co2_sites(), 1e6, 298.0, 3, 1, 1000, 0.6, [4, 5], tmp_path,
                                               ewald=(0.35, 7.5, 5))
    counts, tables = ref_grid(lj, fw, [4, 5], 0.6)
    g = ol.Gcmc(lj, fw, ol.co2_sites(), 1e6, 298.0, 3, 1, ewald=(0.35, 7.5, 5))
    g.set_grid(counts, [4, 5], tables)
    g.run(1000)
    assert g.n_molecules == n_ref and n_ref > 0
    assert all(np.array_equal(a, b) for a, b in zip(g.particles(), parts))
    assert g.electrostatic_energy() == e_ref


def test_gcmc_grid_missing_species_is_out_of_range():
    fw, lj = ol.c0_system(40, a=14.0)
    g = ol.Gcmc(lj, fw, [(0.0, 0.0, 0.0, 4, 0.0)], 1e7, 300.0, 5, 0)
    with pytest.raises(IndexError):
        g.set_grid((4, 4, 4), [3], np.zeros((1, 4, 4, 4)))


def test_peng_robinson_bitwise():
    out = C.c_double()
    for bar in (0.1, 1.1525, 30.0):
        assert ol.ref().ref_pr_fugacity(C.c_double(304.1282), C.c_double(7377300.0), C.c_double(0.22394),
                                        C.c_double(298.0), C.c_double(bar * 1e5), C.byref(out)) == 0
        assert out.value == ol.oracle().orc_pr_fugacity(C.c_double(304.1282), C.c_double(7377300.0),
                                                        C.c_double(0.22394), C.c_double(298.0), C.c_double(bar * 1e5))

"""Pins the FP64 oracle (oracle/oracle.c) to published data, closed forms and brute force,
independently of the compiled reference: the anchors the reference's own validation suite
uses (tests/validation/*.cc), restated."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from oracle_lib import ptr

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ----------------------------------------------------------------- Philox
def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    ol.oracle().orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def test_philox_random123_known_answers():
    # test_philox_random123_vectors.cc:22-44
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_stream_layout_and_uniform_range():
    # nextU64 = out0 << 32 | out1 of counter (call lo, hi, stream lo, hi), key (seed lo, hi)
    seed, stream = 0x123456789ABCDEF0, 0x0FEDCBA987654321
    out = philox([5, 0, stream & 0xFFFFFFFF, stream >> 32], [seed & 0xFFFFFFFF, seed >> 32])
    u64 = (out[0] << 32) | out[1]
    assert ol.uniforms(seed, stream, 5, 1)[0] == (u64 >> 11) * 2.0 ** -53
    a, b = ol.uniforms(42, 0, 0, 8), ol.uniforms(42, 0, 0, 8)
    assert np.array_equal(a, b)
    assert ol.uniforms(42, 1, 0, 1)[0] != a[0]
    u = ol.uniforms(7, 3, 0, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * 9.13e-4


# ----------------------------------------------------------------- lattice
def test_minimum_image_matches_brute_force_on_rhombohedral_cell():
    m = ol.cell_from_params(10.0, 10.0, 10.0, 60.0, 60.0, 60.0)
    lat = ol.lattice(m)
    rng = np.random.default_rng(1)
    shifts = np.array([[i, j, k] for i in range(-7, 8) for j in range(-7, 8) for k in range(-7, 8)]) @ m
    for _ in range(300):
        a, b = rng.uniform(-15, 15, 3), rng.uniform(-15, 15, 3)
        d = np.zeros(3)
        ol.oracle().orc_minimum_image(C.byref(lat), ptr(a.copy()), ptr(b.copy()), ptr(d))
        brute = np.min(np.linalg.norm(b - a + shifts, axis=1))
        assert abs(np.linalg.norm(d) - brute) <= 1e-12 * max(brute, 1.0)


def test_frac_cart_round_trip_and_widths():
    m = ol.cell_from_params(11.0, 13.0, 17.0, 80.0, 100.0, 115.0)
    lat = ol.lattice(m)
    f = np.array([0.1, 0.7, 0.35])
    c, f2 = np.zeros(3), np.zeros(3)
    ol.oracle().orc_frac_to_cart(C.byref(lat), ptr(f), ptr(c))
    ol.oracle().orc_cart_to_frac(C.byref(lat), ptr(c), ptr(f2))
    assert np.allclose(f, f2, rtol=0, atol=1e-14)
    vol = abs(np.linalg.det(m))
    for axis in range(3):
        o = [m[(axis + 1) % 3], m[(axis + 2) % 3]]
        expect = vol / np.linalg.norm(np.cross(o[0], o[1]))
        assert abs(ol.oracle().orc_perpendicular_width(C.byref(lat), axis) - expect) <= 1e-12 * expect


def test_supercell_replication_counts():
    # n_i * L_perp_i > 2 rc, smallest such n (supercell.cc:7-20)
    assert ol.min_supercell(ol.cell_from_params(24, 24, 24, 60, 60, 60), 12.0) == (2, 2, 2)
    assert ol.min_supercell(ol.cubic(25.83), 12.8) == (1, 1, 1)
    assert ol.min_supercell(ol.cubic(10.0), 12.8) == (3, 3, 3)


# ----------------------------------------------------------------- Lennard-Jones
NIST_LJ = {  # test_lennard_jones_nist.cc:122-155 (U_pair, W_pair, U_LRC)
    "lj_sample_config_periodic1.txt": [(3.0, ol.TRUNCATED, -4.3515e+03, -5.6867e+02, -1.9849e+02),
                                       (4.0, ol.TRUNCATED, -4.4675e+03, -1.2639e+03, -8.3769e+01),
                                       (3.0, ol.LFS, -3.8709e+03, 3.1754e+02, None)],
    "lj_sample_config_periodic2.txt": [(3.0, ol.TRUNCATED, -6.9000e+02, -5.6846e+02, -2.4230e+01),
                                       (4.0, ol.TRUNCATED, -7.0460e+02, -6.5599e+02, -1.0226e+01),
                                       (3.0, ol.LFS, -6.2012e+02, -4.4533e+02, None)],
    "lj_sample_config_periodic3.txt": [(3.0, ol.TRUNCATED, -1.1467e+03, -1.1649e+03, -4.9622e+01),
                                       (4.0, ol.TRUNCATED, -1.1754e+03, -1.3371e+03, -2.0942e+01),
                                       (3.0, ol.LFS, -1.0210e+03, -9.3578e+02, None)],
    "lj_sample_config_periodic4.txt": [(3.0, ol.TRUNCATED, -1.6790e+01, -4.6249e+01, -5.4517e-01),
                                       (4.0, ol.TRUNCATED, -1.7060e+01, -4.7869e+01, -2.3008e-01),
                                       (3.0, ol.LFS, -1.5001e+01, -4.3096e+01, None)],
}


def load_nist(path, with_type=False):
    """NIST SRSW sample file: box lengths, entry count (particles for LJ, molecules of three
    sites for SPC/E), then one row per site."""
    with open(path) as f:
        lx, ly, lz = (float(v) for v in f.readline().split())
        n = int(f.readline()) * (3 if with_type else 1)
        rows = [f.readline().split() for _ in range(n)]
    xyz = np.array([[float(r[1]), float(r[2]), float(r[3])] for r in rows])
    types = [r[4] for r in rows] if with_type else None
    return np.diag([lx, ly, lz]), xyz, types


@pytest.mark.parametrize("name", sorted(NIST_LJ))
def test_lennard_jones_nist_reference(name):
    m, xyz, _ = load_nist(os.path.join(GOLDEN, "nist_lj", name))
    fw = ol.Framework(m, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), np.zeros(len(xyz)),
                      np.zeros(len(xyz), dtype=np.uint32))
    lat = ol.lattice(m)
    p, keep = ol.particles(fw)
    for rc, trunc, upair, wpair, ulrc in NIST_LJ[name]:
        lj, keep2 = ol.lj_struct(ol.LJ(np.array([1.0]), np.array([1.0]), rc, trunc))
        u = ol.oracle().orc_lj_total_energy(C.byref(lj), C.byref(p), C.byref(lat))
        w = ol.oracle().orc_lj_total_virial(C.byref(lj), C.byref(p), C.byref(lat))
        assert abs(u - upair) / abs(upair) < 5e-5
        assert abs(w - wpair) / abs(wpair) < 5e-5
        if ulrc is not None:
            t = ol.oracle().orc_lj_tail_energy(C.byref(lj), C.byref(p), C.byref(lat))
            assert abs(t - ulrc) / abs(ulrc) < 5e-5


def test_mixing_rules_closed_form():
    # lennard_jones.cc:49-68: same species exact; LB arithmetic sigma; geometric sigma
    eps, sig = np.array([52.8, 148.0]), np.array([3.43, 3.73])
    for mixing, s_expect in ((0, 0.5 * (3.43 + 3.73)), (1, np.sqrt(3.43 * 3.73))):
        lj, keep = ol.lj_struct(ol.LJ(eps, sig, 12.0, mixing=mixing))
        e, s = C.c_double(), C.c_double()
        ol.oracle().orc_lj_mixed(C.byref(lj), 0, 1, C.byref(e), C.byref(s))
        assert e.value == np.sqrt(52.8 * 148.0) and s.value == s_expect
        ol.oracle().orc_lj_mixed(C.byref(lj), 1, 1, C.byref(e), C.byref(s))
        assert e.value == 148.0 and s.value == 3.73


def simpson(f, a, b, n):
    x = np.linspace(a, b, n + 1)
    y = f(x)
    return (b - a) / n / 3.0 * (y[0] + y[-1] + 4 * y[1:-1:2].sum() + 2 * y[2:-1:2].sum())


def test_tail_correction_matches_quadrature():
    # U_LRC = (2 pi / V) sum_ab N_a N_b int_rc^inf r^2 V_ab(r) dr, with u = 1/r
    # (test_lennard_jones_tail_correction.cc:103-190)
    eps, sig, rc = np.array([52.8, 30.2, 148.0]), np.array([3.43, 3.12, 3.73]), 12.8
    species = np.array([0] * 30 + [1] * 20 + [2] * 5, dtype=np.uint32)
    m = ol.cubic(30.0)
    fw = ol.Framework(m, np.zeros(55), np.zeros(55), np.zeros(55), np.zeros(55), species)
    lj, keep = ol.lj_struct(ol.LJ(eps, sig, rc))
    p, keep2 = ol.particles(fw)
    lat = ol.lattice(m)
    got = ol.oracle().orc_lj_tail_energy(C.byref(lj), C.byref(p), C.byref(lat))
    counts = np.bincount(species, minlength=3)
    total = 0.0
    for a in range(3):
        for b in range(3):
            e = eps[a] if a == b else np.sqrt(eps[a] * eps[b])
            s = sig[a] if a == b else 0.5 * (sig[a] + sig[b])
            integrand = lambda u: 4 * e * (s ** 12 * u ** 8 - s ** 6 * u ** 2)
            total += counts[a] * counts[b] * simpson(integrand, 0.0, 1.0 / rc, 2000)
    expect = 2 * np.pi / 27000.0 * total
    assert abs(got - expect) <= 1e-9 * abs(expect)


def test_supercell_lj_energy_invariance():
    fw, lj = ol.c0_system(120, a=13.0)
    lj.cutoff = 6.0
    l1, k1 = ol.lj_struct(lj)
    p, keep = ol.particles(fw)
    e1 = ol.oracle().orc_lj_total_energy(C.byref(l1), C.byref(p), C.byref(ol.lattice(fw.m)))
    sc = ol.replicate(fw, (2, 2, 2))
    p2, keep2 = ol.particles(sc)
    e8 = ol.oracle().orc_lj_total_energy(C.byref(l1), C.byref(p2), C.byref(ol.lattice(sc.m)))
    assert abs(e8 - 8 * e1) <= 1e-10 * abs(8 * e1)


# ----------------------------------------------------------------- Ewald
def rock_salt(a=1.0, nrep=3):
    na = [(0, 0, 0), (a, a, 0), (a, 0, a), (0, a, a)]
    cl = [(a, 0, 0), (0, a, 0), (0, 0, a), (a, a, a)]
    pos, q = [], []
    for ox in range(nrep):
        for oy in range(nrep):
            for oz in range(nrep):
                d = 2 * a * np.array([ox, oy, oz])
                for s in na:
                    pos.append(np.array(s) + d); q.append(1.0)
                for s in cl:
                    pos.append(np.array(s) + d); q.append(-1.0)
    pos = np.array(pos)
    return ol.Framework(ol.cubic(2 * a * nrep), pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(),
                        np.array(q), np.zeros(len(q), dtype=np.uint32))


def ewald_total(fw, alpha, rc, kmax, excl=None):
    p, keep = ol.particles(fw)
    b = ol.OrcBreakdown()
    ex = np.ascontiguousarray(excl if excl is not None else np.zeros((0, 2)), dtype=np.uintp).ravel()
    rc_ = ol.oracle().orc_ewald_breakdown_energy(C.c_double(alpha), C.c_double(rc), C.c_int(kmax), C.byref(p),
                                                 C.byref(ol.lattice(fw.m)),
                                                 ex.ctypes.data_as(ol.szp) if len(ex) else None,
                                                 C.c_size_t(len(ex) // 2), C.byref(b))
    assert rc_ == 0
    return b


def test_madelung_constant():
    # test_ewald_madelung.cc:108-127: a = 1, 3x3x3, alpha 1.4, rc 2.9, kmax 12, 1e-6
    fw = rock_salt()
    b = ewald_total(fw, 1.4, 2.9, 12)
    total = b.real + b.reciprocal + b.self + b.exclusion
    madelung = -total / ol.KCOULOMB * 1.0 / 108
    assert abs(madelung - 1.747564594633) / 1.747564594633 < 1e-6


def test_alpha_independence():
    # test_ewald_alpha_independence.cc:68-110
    fw = rock_salt()
    e = []
    for alpha in (1.6, 1.9, 2.2, 2.5):
        b = ewald_total(fw, alpha, 2.9, 20)
        e.append(b.real + b.reciprocal + b.self)
    assert max(abs(v - e[0]) / abs(e[0]) for v in e) < 1e-6


SPCE = {  # test_ewald_spce_water.cc:136-146 (E_real, E_fourier, E_self, E_intra)
    "spce_sample_config_periodic1.txt": (-5.58889e+05, 6.27009e+03, -2.84469e+06, 2.80999e+06),
    "spce_sample_config_periodic2.txt": (-1.19295e+06, 6.03495e+03, -5.68938e+06, 5.61998e+06),
    "spce_sample_config_periodic3.txt": (-1.96297e+06, 5.24461e+03, -8.53407e+06, 8.42998e+06),
    "spce_sample_config_periodic4.txt": (-3.57226e+06, 7.58785e+03, -1.42235e+07, 1.41483e+07),
}


@pytest.mark.parametrize("name", sorted(SPCE))
def test_spce_water_ewald_terms(name):
    m, xyz, types = load_nist(os.path.join(GOLDEN, "nist_spce", name), with_type=True)
    q = np.array([-2 * 0.42380 if t == "O" else 0.42380 for t in types])
    fw = ol.Framework(m, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), q, np.zeros(len(q), dtype=np.uint32))
    n_mol = len(q) // 3
    excl = np.array([(3 * i + a, 3 * i + b) for i in range(n_mol) for a, b in ((0, 1), (0, 2), (1, 2))])
    b = ewald_total(fw, 5.6 / m[0, 0], 10.0, 5, excl)
    for got, ref in zip((b.real, b.reciprocal, b.self, b.exclusion), SPCE[name]):
        assert abs(got - ref) / abs(ref) < 5e-4


def test_kset_counts():
    # closed-form counts of n^2 < kmax^2 + 2, n != 0 (test_gcmc_co2_irmof1_isotherm.cc:208-213)
    for kmax, expect in ((6, 948), (8, 2204), (5, 0)):
        ks = ol.OrcKset()
        ol.oracle().orc_reciprocal_terms(C.byref(ol.lattice(ol.cubic(20.0))), C.c_double(0.3), C.c_int(kmax),
                                         C.byref(ks))
        n = ks.n
        brute = sum(1 for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(-kmax, kmax + 1)
                    if (a, b, c) != (0, 0, 0) and a * a + b * b + c * c < kmax * kmax + 2)
        ol.oracle().orc_kset_free(C.byref(ks))
        assert n == brute
        if expect:
            assert n == expect


def charged_box():
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    return ol.synth_framework(m, 120, seed=4, stream=0, charge_sigma=0.5)


def co2_at(c, u):
    u = np.asarray(u, dtype=float) / np.linalg.norm(u)
    c = np.asarray(c, dtype=float)
    return np.array([[*(c - 1.16 * u), -0.35], [*c, 0.70], [*(c + 1.16 * u), -0.35]])


def test_state_delta_equals_breakdown_difference():
    fw = charged_box()
    alpha, kmax = 0.35, 6
    st = ol.EwaldState(fw, alpha, kmax)
    mol = co2_at([3.0, 4.0, 5.0], [1, 2, 0.5])
    d = st.propose([], mol)
    before = ewald_total(fw, alpha, 1e-3, kmax)
    after_fw = ol.Framework(fw.m, np.concatenate([fw.x, mol[:, 0]]), np.concatenate([fw.y, mol[:, 1]]),
                            np.concatenate([fw.z, mol[:, 2]]), np.concatenate([fw.q, mol[:, 3]]),
                            np.zeros(fw.n + 3, dtype=np.uint32))
    n = fw.n
    after = ewald_total(after_fw, alpha, 1e-3, kmax, np.array([(n, n + 1), (n, n + 2), (n + 1, n + 2)]))
    expect = (after.reciprocal + after.self + after.exclusion) - (before.reciprocal + before.self)
    assert abs(d - expect) <= 1e-10 * abs(after.reciprocal)
    assert abs(st.committed() - (before.reciprocal + before.self)) <= 1e-10 * abs(before.self)


def test_state_antisymmetry_rollback_and_drift():
    # test_ewald_incremental_state.cc:137-430
    fw = charged_box()
    st = ol.EwaldState(fw, 0.35, 6)
    rng = np.random.default_rng(2)
    mols = []
    for step in range(30):
        mol = co2_at(rng.uniform(0, 15, 3), rng.normal(size=3))
        raw = st.raw()
        d_ins = st.propose([], mol)
        assert np.array_equal(raw, st.raw())  # propose never mutates
        st.commit([], mol)
        d_del = st.propose(mol, [])
        assert abs(d_ins + d_del) <= 1e-9 * max(abs(d_ins), 1.0)
        mols.append(mol)
    allp = np.concatenate(mols)
    full = ol.Framework(fw.m, np.concatenate([fw.x, allp[:, 0]]), np.concatenate([fw.y, allp[:, 1]]),
                        np.concatenate([fw.z, allp[:, 2]]), np.concatenate([fw.q, allp[:, 3]]),
                        np.zeros(fw.n + len(allp), dtype=np.uint32))
    molecules = [list(range(fw.n + 3 * i, fw.n + 3 * i + 3)) for i in range(len(mols))]
    drift = abs(st.recompute(full, molecules) - st.committed())
    assert drift <= 1e-10 * abs(st.committed())


def test_coulomb_potential_is_the_linear_term_of_propose_insertion():
    # phi_recip(r) = (proposeInsertion(+1 at r) - proposeInsertion(-1 at r)) / 2
    fw = charged_box()
    st = ol.EwaldState(fw, 0.35, 6)
    counts = (12, 13, 14)
    nodes = np.array([0, 77, 1000, 2183], dtype=np.int64)
    real, rec = ol.coulomb_nodes(st, 8.0, fw, fw.m, counts, nodes)
    for i, node in enumerate(nodes):
        ix, iy, iz = node % counts[0], (node // counts[0]) % counts[1], node // (counts[0] * counts[1])
        c = np.array([ix / counts[0], iy / counts[1], iz / counts[2]]) @ fw.m
        plus = st.propose([], [[*c, 1.0]])
        minus = st.propose([], [[*c, -1.0]])
        assert abs(rec[i] - 0.5 * (plus - minus)) <= 1e-12 * max(abs(rec[i]), 1.0)


def test_grid_interpolation_exact_on_nodes_and_linear_fields():
    m = ol.cell_from_params(10.0, 11.0, 12.0, 80.0, 95.0, 100.0)
    n = np.array([7, 8, 9], dtype=np.int32)
    lat = ol.lattice(m)
    # a field linear in fractional coordinates within a cell, sampled at nodes
    f = lambda i, j, k: 3.0 + 2.0 * i - 1.5 * j + 0.5 * k
    table = np.array([[[f(i, j, k) for i in range(7)] for j in range(8)] for k in range(9)], dtype=np.float64)
    for (i, j, k) in ((0, 0, 0), (3, 5, 7), (6, 7, 8)):
        pos = np.array([i / 7, j / 8, k / 9]) @ m
        v = ol.oracle().orc_grid_interpolate(C.byref(lat), ptr(n, ol.i32p), ptr(table.ravel()), ptr(pos))
        assert abs(v - f(i, j, k)) <= 1e-12 * abs(f(i, j, k))
    pos = np.array([2.5 / 7, 3.25 / 8, 4.75 / 9]) @ m
    v = ol.oracle().orc_grid_interpolate(C.byref(lat), ptr(n, ol.i32p), ptr(table.ravel()), ptr(pos))
    assert abs(v - f(2.5, 3.25, 4.75)) <= 1e-12


# ----------------------------------------------------------------- synthetic frameworks
def test_synthetic_framework_invariants():
    m = ol.cell_from_params(24.0, 24.0, 24.0, 60.0, 60.0, 60.0)
    fw = ol.synth_framework(m, 600, seed=1, stream=0, charge_sigma=0.5)
    assert list(np.bincount(fw.species)) == [270, 180, 102, 48]
    assert abs(fw.q.sum()) < 1e-8 and 0.4 < fw.q.std() < 0.6
    lat = ol.lattice(m)
    dmin = np.inf
    for i in range(0, 600, 7):
        for j in range(i):
            d = np.zeros(3)
            ol.oracle().orc_minimum_image(C.byref(lat), ptr(np.array([fw.x[i], fw.y[i], fw.z[i]])),
                                          ptr(np.array([fw.x[j], fw.y[j], fw.z[j]])), ptr(d))
            dmin = min(dmin, np.linalg.norm(d))
    assert dmin >= 1.0
    assert list(np.bincount(ol.c0_system()[0].species)) == [191, 127, 72, 34]


def test_peng_robinson_ideal_limit():
    # fugacity coefficient -> 1 as P -> 0 (test_peng_robinson.cc:37-60)
    f = ol.oracle().orc_pr_fugacity(C.c_double(304.1282), C.c_double(7377300.0), C.c_double(0.22394),
                                    C.c_double(298.0), C.c_double(1.0))
    assert abs(f / 1.0 - 1.0) < 1e-6

"""ctypes bindings for the oracle (oracle/liboracle.so) and the compiled reference
(oracle/_ref/libaleator_ref.so).  Test infrastructure only: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libaleator_ref.so")

dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)
szp = C.POINTER(C.c_size_t)

KCOULOMB = 167100.9468982874
TRUNCATED, SHIFTED, LFS = 0, 1, 2


def ptr(a, t=dp):
    return a.ctypes.data_as(t) if a is not None else None


class OrcLattice(C.Structure):
    _fields_ = [("m", C.c_double * 9), ("orthorhombic", C.c_int), ("reduced", C.c_double * 9)]


class OrcParticles(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", dp), ("y", dp), ("z", dp), ("q", dp), ("species", u32p)]


class OrcLJ(C.Structure):
    _fields_ = [("epsilon", dp), ("sigma", dp), ("n_species", C.c_uint32), ("cutoff", C.c_double),
                ("truncation", C.c_int), ("mixing", C.c_int)]


class OrcKset(C.Structure):
    _fields_ = [("n", C.c_size_t), ("k", dp), ("weight", dp), ("nvec", i32p)]


class OrcEwaldState(C.Structure):
    _fields_ = [("lat", OrcLattice), ("alpha", C.c_double), ("terms", OrcKset), ("volume", C.c_double),
                ("s_re", dp), ("s_im", dp), ("self_energy", C.c_double),
                ("exclusion_energy", C.c_double)]


class OrcSite(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("charge", C.c_double)]


class OrcMolSite(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("lj_species", C.c_uint32),
                ("charge", C.c_double)]


class OrcBreakdown(C.Structure):
    _fields_ = [("real", C.c_double), ("reciprocal", C.c_double), ("self", C.c_double),
                ("exclusion", C.c_double), ("surface", C.c_double)]


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _oracle = C.CDLL(ORACLE_SO)
        _oracle.orc_lj_particle_energy.restype = C.c_double
        _oracle.orc_lj_tail_energy.restype = C.c_double
        _oracle.orc_lj_total_energy.restype = C.c_double
        _oracle.orc_lj_total_virial.restype = C.c_double
        _oracle.orc_lj_pair_energy.restype = C.c_double
        _oracle.orc_lattice_volume.restype = C.c_double
        _oracle.orc_perpendicular_width.restype = C.c_double
        _oracle.orc_ewald_real_particle_energy.restype = C.c_double
        _oracle.orc_ewald_state_committed.restype = C.c_double
        _oracle.orc_ewald_state_delta.restype = C.c_double
        _oracle.orc_ewald_state_recompute.restype = C.c_double
        _oracle.orc_grid_interpolate.restype = C.c_double
        _oracle.orc_widom.restype = C.c_double
        _oracle.orc_rng_next_uniform.restype = C.c_double
        _oracle.orc_rng_next_u64.restype = C.c_uint64
        _oracle.orc_mlp_loss_grad.restype = C.c_double
        _oracle.orc_pr_fugacity.restype = C.c_double
    return _oracle


def pr_fugacity(tc, pc, omega, temperature, pressure) -> float:
    """PengRobinson::fugacityPascal restated (oracle.c, peng_robinson.cc:10-110)."""
    return oracle().orc_pr_fugacity(C.c_double(tc), C.c_double(pc), C.c_double(omega), C.c_double(temperature),
                                    C.c_double(pressure))


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"{REF_SO} missing: run `make -C oracle ref` where the reference exists")
        _ref = C.CDLL(REF_SO)
        _ref.ref_ewald_state_create.restype = C.c_void_p
        _ref.ref_ewald_state_destroy.argtypes = [C.c_void_p]
        for f in ("ref_ewald_state_propose", "ref_ewald_state_committed", "ref_ewald_state_recompute"):
            getattr(_ref, f).restype = C.c_double
        _ref.ref_ewald_state_propose.argtypes = [C.c_void_p, dp, C.c_int, dp, C.c_int]
        _ref.ref_ewald_state_commit.argtypes = [C.c_void_p, dp, C.c_int, dp, C.c_int]
        _ref.ref_ewald_state_committed.argtypes = [C.c_void_p]
        _ref.ref_ewald_state_raw.argtypes = [C.c_void_p, dp, C.c_uint64]
        _ref.ref_ewald_state_raw.restype = C.c_uint64
        _ref.ref_ewald_state_resync.argtypes = [C.c_void_p, C.c_uint64, dp, dp, dp, dp, u64p, u64p,
                                                C.c_uint64]
        _ref.ref_ewald_state_recompute.argtypes = _ref.ref_ewald_state_resync.argtypes
    return _ref


# --------------------------------------------------------------------------- model objects
@dataclass
class Framework:
    m: np.ndarray          # (3,3) row vectors
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray
    species: np.ndarray    # uint32

    @property
    def n(self):
        return len(self.x)


@dataclass
class LJ:
    eps: np.ndarray
    sigma: np.ndarray
    cutoff: float
    truncation: int = TRUNCATED
    mixing: int = 0


def lattice(m) -> OrcLattice:
    lat = OrcLattice()
    mm = np.ascontiguousarray(m, dtype=np.float64)
    oracle().orc_lattice_init(C.byref(lat), ptr(mm))
    return lat


def particles(fw: Framework) -> tuple[OrcParticles, tuple]:
    keep = tuple(np.ascontiguousarray(a) for a in (fw.x, fw.y, fw.z, fw.q))
    sp = np.ascontiguousarray(fw.species, dtype=np.uint32)
    p = OrcParticles(fw.n, ptr(keep[0]), ptr(keep[1]), ptr(keep[2]), ptr(keep[3]), ptr(sp, u32p))
    return p, keep + (sp,)


def lj_struct(lj: LJ) -> tuple[OrcLJ, tuple]:
    e = np.ascontiguousarray(lj.eps, dtype=np.float64)
    s = np.ascontiguousarray(lj.sigma, dtype=np.float64)
    return OrcLJ(ptr(e), ptr(s), len(e), lj.cutoff, lj.truncation, lj.mixing), (e, s)


def synth_framework(m, n_atoms, seed=1, stream=0, min_dist=1.0,
                    proportions=(0.45, 0.30, 0.17, 0.08), charge_sigma=0.0, charge_stream=2):
    m = np.ascontiguousarray(m, dtype=np.float64)
    x, y, z, q = (np.zeros(n_atoms) for _ in range(4))
    sp = np.zeros(n_atoms, dtype=np.uint32)
    props = np.ascontiguousarray(proportions, dtype=np.float64)
    rc = oracle().orc_synth_framework(ptr(m), C.c_size_t(n_atoms), C.c_uint64(seed), C.c_uint64(stream),
                                      C.c_double(min_dist), ptr(props), C.c_uint32(len(props)),
                                      C.c_double(charge_sigma), C.c_uint64(charge_stream),
                                      ptr(x), ptr(y), ptr(z), ptr(q), ptr(sp, u32p))
    if rc != 0:
        raise RuntimeError("synthetic framework placement failed")
    return Framework(m, x, y, z, q, sp)


def cubic(a):
    return np.diag([a, a, a]).astype(np.float64)


def cell_from_params(a, b, c, alpha, beta, gamma):
    """Row-vector lattice matrix from edges (A) and angles (degrees)."""
    al, be, ga = np.radians([alpha, beta, gamma])
    ca, cb, cg, sg = np.cos(al), np.cos(be), np.cos(ga), np.sin(ga)
    cy = (ca - cb * cg) / sg
    cz2 = 1.0 - cb * cb - cy * cy
    return np.array([[a, 0, 0], [b * cg, b * sg, 0], [c * cb, c * cy, c * np.sqrt(cz2)]], dtype=np.float64)


# --------------------------------------------------------------------------- oracle wrappers
def uniforms(seed, stream, first, n):
    out = np.zeros(n)
    oracle().orc_uniforms(C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(first), C.c_uint64(n), ptr(out))
    return out


def min_supercell(m, rc):
    n = (C.c_int * 3)()
    lat = lattice(m)
    oracle().orc_min_supercell_replication(C.byref(lat), C.c_double(rc), n)
    return tuple(n)


def replicate(fw: Framework, reps):
    lat = lattice(fw.m)
    p, keep = particles(fw)
    out = OrcParticles()
    om = np.zeros((3, 3))
    oracle().orc_replicate_supercell(C.byref(p), C.byref(lat), reps[0], reps[1], reps[2], C.byref(out), ptr(om))
    n = out.n
    res = Framework(om, np.ctypeslib.as_array(out.x, (n,)).copy(), np.ctypeslib.as_array(out.y, (n,)).copy(),
                    np.ctypeslib.as_array(out.z, (n,)).copy(), np.ctypeslib.as_array(out.q, (n,)).copy(),
                    np.ctypeslib.as_array(out.species, (n,)).copy())
    oracle().orc_particles_free(C.byref(out))
    return res


def grid_counts(m, spacing):
    n = np.zeros(3, dtype=np.int32)
    lat = lattice(m)
    oracle().orc_grid_point_counts(C.byref(lat), C.c_double(spacing), ptr(n, i32p))
    return tuple(int(v) for v in n)


def grid_build(lj: LJ, fw: Framework, guests, counts, cap=1e5):
    lat = lattice(fw.m)
    p, keep = particles(fw)
    l, keep2 = lj_struct(lj)
    g = np.ascontiguousarray(sorted(set(guests)), dtype=np.uint32)
    n = np.ascontiguousarray(counts, dtype=np.int32)
    out = np.zeros((len(g), n[2], n[1], n[0]))
    oracle().orc_grid_build(C.byref(l), C.byref(p), C.byref(lat), ptr(g, u32p), len(g), ptr(n, i32p),
                            C.c_double(cap), ptr(out))
    return out


def grid_nodes_lj(lj: LJ, fw: Framework, grid_m, counts, guest, nodes, cap=1e5):
    lat = lattice(fw.m)
    p, keep = particles(fw)
    l, keep2 = lj_struct(lj)
    gm = np.ascontiguousarray(grid_m, dtype=np.float64)
    n = np.ascontiguousarray(counts, dtype=np.int32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    out = np.zeros(len(nodes))
    oracle().orc_grid_nodes_lj(C.byref(l), C.byref(p), C.byref(lat), ptr(gm), ptr(n, i32p), C.c_uint32(guest),
                               ptr(nodes, i64p), C.c_size_t(len(nodes)), C.c_double(cap), ptr(out))
    return out


class EwaldState:
    """Oracle EwaldIncrementalState over a framework (restates ewald_incremental_state.cc)."""

    def __init__(self, fw: Framework, alpha, kmax):
        self.st = OrcEwaldState()
        lat = lattice(fw.m)
        p, self._keep = particles(fw)
        if oracle().orc_ewald_state_init(C.byref(self.st), C.c_double(alpha), C.c_int(kmax), C.byref(p),
                                         C.byref(lat)) != 0:
            raise RuntimeError("ewald state init failed")

    def __del__(self):
        try:
            oracle().orc_ewald_state_free(C.byref(self.st))
        except Exception:
            pass

    @property
    def nk(self):
        return self.st.terms.n

    @staticmethod
    def _sites(s):
        s = np.asarray(s, dtype=np.float64).reshape(-1, 4)
        arr = (OrcSite * max(len(s), 1))()
        for i, r in enumerate(s):
            arr[i] = OrcSite(*r)
        return arr, len(s)

    def propose(self, removed, added):
        r, nr = self._sites(removed)
        a, na = self._sites(added)
        return oracle().orc_ewald_state_delta(C.byref(self.st), r, C.c_size_t(nr), a, C.c_size_t(na),
                                              None, None, None, None, None)

    def commit(self, removed, added):
        r, nr = self._sites(removed)
        a, na = self._sites(added)
        oracle().orc_ewald_state_commit(C.byref(self.st), r, C.c_size_t(nr), a, C.c_size_t(na))

    def committed(self):
        return oracle().orc_ewald_state_committed(C.byref(self.st))

    def raw(self):
        out = np.zeros(2 * self.nk + 2)
        oracle().orc_ewald_state_raw(C.byref(self.st), ptr(out))
        return out

    def _mol(self, fw, molecules):
        off = np.zeros(len(molecules) + 1, dtype=np.uint64)
        flat = []
        for i, m in enumerate(molecules):
            flat.extend(m)
            off[i + 1] = len(flat)
        flat = np.ascontiguousarray(flat if flat else [0], dtype=np.uint64)
        return off, flat

    def resync(self, fw: Framework, molecules):
        p, keep = particles(fw)
        off, flat = self._mol(fw, molecules)
        oracle().orc_ewald_state_resync(C.byref(self.st), C.byref(p), ptr(off, szp), ptr(flat, szp),
                                        C.c_size_t(len(molecules)))

    def recompute(self, fw: Framework, molecules):
        p, keep = particles(fw)
        off, flat = self._mol(fw, molecules)
        return oracle().orc_ewald_state_recompute(C.byref(self.st), C.byref(p), ptr(off, szp),
                                                  ptr(flat, szp), C.c_size_t(len(molecules)))

    def kvectors(self):
        n = self.nk
        return (np.ctypeslib.as_array(self.st.terms.k, (n, 3)).copy(),
                np.ctypeslib.as_array(self.st.terms.weight, (n,)).copy(),
                np.ctypeslib.as_array(self.st.terms.nvec, (n, 3)).copy())

    def structure_factor(self):
        n = self.nk
        return (np.ctypeslib.as_array(self.st.s_re, (n,)).copy(),
                np.ctypeslib.as_array(self.st.s_im, (n,)).copy())


def coulomb_nodes(st: EwaldState, rc_real, fw: Framework, grid_m, counts, nodes):
    p, keep = particles(fw)
    gm = np.ascontiguousarray(grid_m, dtype=np.float64)
    n = np.ascontiguousarray(counts, dtype=np.int32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    real = np.zeros(len(nodes))
    rec = np.zeros(len(nodes))
    oracle().orc_grid_nodes_coulomb(C.byref(st.st), C.c_double(rc_real), C.byref(p), ptr(gm), ptr(n, i32p),
                                    ptr(nodes, i64p), C.c_size_t(len(nodes)), ptr(real), ptr(rec))
    return real, rec


def widom(lj: LJ, fw: Framework, sites, seed, stream, n_trials, temperature, ewald: EwaldState | None = None,
          rc_real=0.0, first_call=0):
    """sites: list of (x, y, z, lj_species, charge).  Returns (K_H, dU per trial)."""
    lat = lattice(fw.m)
    p, keep = particles(fw)
    l, keep2 = lj_struct(lj)
    arr = (OrcMolSite * len(sites))()
    for i, s in enumerate(sites):
        arr[i] = OrcMolSite(s[0], s[1], s[2], int(s[3]), s[4])
    du = np.zeros(n_trials)
    kh = oracle().orc_widom(C.byref(l), C.byref(p), C.byref(lat), arr, C.c_size_t(len(sites)),
                            C.byref(ewald.st) if ewald is not None else None, C.c_double(rc_real),
                            C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(first_call), C.c_uint64(n_trials),
                            C.c_double(temperature), ptr(du))
    return kh, du


# --------------------------------------------------------------------------- standard workloads
C0_TYPES = ((52.8, 3.43), (30.2, 3.12), (22.1, 2.57), (62.4, 2.46))
METHANE = (148.0, 3.73)
CO2_C = (27.0, 2.80)
CO2_O = (79.0, 3.05)


def c0_system(n_atoms=424, a=25.83):
    fw = synth_framework(cubic(a), n_atoms, seed=1, stream=0)
    eps = np.array([t[0] for t in C0_TYPES] + [METHANE[0]])
    sig = np.array([t[1] for t in C0_TYPES] + [METHANE[1]])
    return fw, LJ(eps, sig, 12.8, TRUNCATED)


def co2_sites():
    """Rigid linear CO2 along z: O(-), C, O(+); species 4 = C, 5 = O."""
    return [(0.0, 0.0, -1.16, 5, -0.35), (0.0, 0.0, 0.0, 4, 0.70), (0.0, 0.0, 1.16, 5, -0.35)]


def c1_system(n_atoms=600):
    m = cell_from_params(24.0, 24.0, 24.0, 60.0, 60.0, 60.0)
    fw = synth_framework(m, n_atoms, seed=1, stream=0, charge_sigma=0.5, charge_stream=2)
    eps = np.array([t[0] for t in C0_TYPES] + [CO2_C[0], CO2_O[0]])
    sig = np.array([t[1] for t in C0_TYPES] + [CO2_C[1], CO2_O[1]])
    return fw, LJ(eps, sig, 12.0, TRUNCATED)


# --------------------------------------------------------------------------- GCMC move loop
class OrcRng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("ctr", C.c_uint32 * 4), ("call_index", C.c_uint64)]


class OrcGcmc(C.Structure):
    _fields_ = [("lj", OrcLJ), ("eps", dp), ("sig", dp), ("lat", OrcLattice), ("n", C.c_size_t),
                ("cap", C.c_size_t), ("n_fw", C.c_size_t), ("x", dp), ("y", dp), ("z", dp), ("q", dp),
                ("sp", u32p), ("n_mol", C.c_size_t), ("mol_cap", C.c_size_t), ("ns", C.c_size_t),
                ("mol", szp), ("sites", C.c_void_p), ("fugacity_internal", C.c_double),
                ("temperature", C.c_double), ("max_translation", C.c_double), ("max_rotation", C.c_double),
                ("charged", C.c_int), ("alpha", C.c_double), ("rc_real", C.c_double), ("ew", OrcEwaldState),
                ("moves_since_resync", C.c_size_t), ("rng", OrcRng),
                ("attempted", C.c_uint64 * 4), ("accepted", C.c_uint64 * 4),
                ("prop_cdf", dp), ("prop_fac", dp), ("prop_nb", C.c_int * 3), ("prop_reps", C.c_int * 3),
                ("prop_nbins", C.c_size_t), ("prop_inv", C.c_double * 9),
                ("grid_tables", dp), ("grid_species", u32p), ("grid_n_tables", C.c_int32), ("grid_n", C.c_int32 * 3)]


class Gcmc:
    """Oracle GCMC chain (restates monte_carlo_engine.cc run/attempt*)."""

    def __init__(self, lj: LJ, fw: Framework, sites, fugacity_pa, temperature, seed, stream, ewald=None,
                 max_translation=1.5, max_rotation=0.5):
        self.g = OrcGcmc()
        l, self._k1 = lj_struct(lj)
        p, self._k2 = particles(fw)
        arr = (OrcMolSite * len(sites))()
        for i, s in enumerate(sites):
            arr[i] = OrcMolSite(s[0], s[1], s[2], int(s[3]), s[4])
        alpha, rc_real, kmax = ewald if ewald is not None else (1.0, 1.0, 1)
        rc = oracle().orc_gcmc_init(C.byref(self.g), C.byref(l), C.byref(p), C.byref(lattice(fw.m)), arr,
                                    C.c_size_t(len(sites)), C.c_double(fugacity_pa), C.c_double(temperature),
                                    C.c_double(max_translation), C.c_double(max_rotation), C.c_double(alpha),
                                    C.c_double(rc_real), C.c_int(kmax), C.c_int(1 if ewald is not None else 0),
                                    C.c_uint64(seed), C.c_uint64(stream))
        if rc != 0:
            raise RuntimeError("gcmc init failed")
        oracle().orc_gcmc_electrostatic_energy.restype = C.c_double

    def __del__(self):
        try:
            oracle().orc_gcmc_free(C.byref(self.g))
        except Exception:
            pass

    def set_proposal(self, w, nb, reps, mix):
        """Biased insertion from bin weights w (nb bins over the cell / reps); None = uniform."""
        self._w = None if w is None else np.ascontiguousarray(w, dtype=np.float64).ravel()
        nb_a = (C.c_int * 3)(*[int(v) for v in nb])
        reps_a = (C.c_int * 3)(*[int(v) for v in reps])
        rc = oracle().orc_gcmc_set_proposal(C.byref(self.g), None if w is None else ptr(self._w), nb_a, reps_a,
                                            C.c_double(mix))
        if rc != 0:
            raise ValueError("invalid proposal")

    def set_grid(self, n_points, species, tables):
        """Framework LJ from these interpolated tables ([n_tables][n2][n1][n0], species ids), as
        the reference engine given a FrameworkEnergyGrid; tables None = direct scan."""
        if tables is None:
            assert oracle().orc_gcmc_set_grid(C.byref(self.g), None, None, 0, None) == 0
            return
        n = np.ascontiguousarray(n_points, dtype=np.int32)
        sp = np.ascontiguousarray(species, dtype=np.uint32)
        t = np.ascontiguousarray(tables, dtype=np.float64)
        if oracle().orc_gcmc_set_grid(C.byref(self.g), ptr(n, i32p), ptr(sp, u32p), C.c_int32(len(sp)), ptr(t)) != 0:
            raise IndexError("no grid table for a site species")

    def run(self, steps, trace=False):
        tr = np.zeros(steps) if trace else None
        oracle().orc_gcmc_run(C.byref(self.g), C.c_uint64(steps), ptr(tr) if trace else None)
        return tr

    @property
    def n_molecules(self):
        return self.g.n_mol

    def particles(self):
        n = self.g.n
        return tuple(np.ctypeslib.as_array(a, (n,)).copy() for a in (self.g.x, self.g.y, self.g.z, self.g.q))

    def molecule_sites(self):
        n = self.g.n_mol * self.g.ns
        return np.ctypeslib.as_array(self.g.mol, (n,)).copy().reshape(self.g.n_mol, self.g.ns) if n else \
            np.zeros((0, self.g.ns), dtype=np.uint64)

    def electrostatic_energy(self):
        return oracle().orc_gcmc_electrostatic_energy(C.byref(self.g))

    @property
    def counts(self):
        return list(self.g.attempted), list(self.g.accepted)


def ref_gcmc_run(lj: LJ, fw: Framework, sites, fugacity_pa, temperature, seed, stream, steps, ewald=None,
                 max_translation=1.5, max_rotation=0.5):
    """The reference engine (oracle/_ref) run for `steps`: (molecule count, (x, y, z, q) of all
    particles, committed Ewald energy)."""
    sx = np.array([s[:3] for s in sites], dtype=np.float64).ravel()
    ss = np.array([s[3] for s in sites], dtype=np.uint32)
    sq = np.array([s[4] for s in sites], dtype=np.float64)
    cap = fw.n + 3 * len(sites) * 2000
    out = [np.zeros(cap) for _ in range(4)]
    npart, nmol, e = C.c_uint64(), C.c_uint64(), C.c_double()
    alpha, rc_real, kmax = ewald if ewald is not None else (1.0, 1.0, 1)
    rc = ref().ref_gcmc_run(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                            lj.truncation, lj.mixing, ptr(np.ascontiguousarray(fw.m)), C.c_uint64(fw.n), ptr(fw.x),
                            ptr(fw.y), ptr(fw.z), ptr(fw.q), ptr(fw.species, u32p), len(sites), ptr(sx),
                            ptr(ss, u32p), ptr(sq), C.c_double(fugacity_pa), C.c_double(temperature),
                            C.c_double(max_translation), C.c_double(max_rotation), 1 if ewald is not None else 0,
                            C.c_double(alpha), C.c_double(rc_real), kmax, C.c_uint64(seed), C.c_uint64(stream),
                            C.c_uint64(steps), C.c_uint64(cap), ptr(out[0]), ptr(out[1]), ptr(out[2]), ptr(out[3]),
                            C.byref(npart), C.byref(nmol), C.byref(e))
    if rc != 0:
        raise RuntimeError("reference engine failed")
    n = npart.value
    return nmol.value, tuple(a[:n] for a in out), e.value


# ---------------------------------------------------------------- learned insertion density network
MLP_PARAMS = 64 * 256 + 256 + 3 * (256 * 256 + 256) + 256 + 1


def mlp_frequencies() -> np.ndarray:
    out = np.zeros(96, dtype=np.int8)
    oracle().orc_mlp_frequencies(out.ctypes.data_as(C.c_void_p))
    return out.reshape(32, 3)


def mlp_forward(params, frac, bf16: bool) -> np.ndarray:
    p = np.ascontiguousarray(params, dtype=np.float64)
    f = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
    y = np.zeros(len(f))
    oracle().orc_mlp_forward(ptr(p), ptr(f), C.c_size_t(len(f)), C.c_int(1 if bf16 else 0), ptr(y))
    return y


def mlp_loss_grad(params, frac, target, bf16: bool):
    p = np.ascontiguousarray(params, dtype=np.float64)
    f = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(target, dtype=np.float64)
    g = np.zeros(MLP_PARAMS)
    loss = oracle().orc_mlp_loss_grad(ptr(p), ptr(f), ptr(t), C.c_size_t(len(f)), C.c_int(1 if bf16 else 0), ptr(g))
    return loss, g


def mlp_offsets():
    """(name, start, size) of each packed parameter tensor."""
    sizes = [("W1", 64 * 256), ("b1", 256)] + sum([[(f"W{l}", 256 * 256), (f"b{l}", 256)] for l in (2, 3, 4)], []) + \
        [("w5", 256), ("b5", 1)]
    out, o = [], 0
    for name, n in sizes:
        out.append((name, o, n))
        o += n
    return out


def mlp_he_init(seed: int) -> np.ndarray:
    """He-uniform weights from Philox(seed, 0) at call index = parameter index, zero biases
    (the initialisation aleator_gpu_mlp_init documents)."""
    p = np.zeros(MLP_PARAMS)
    u = uniforms(seed, 0, 0, MLP_PARAMS)
    for name, o, n in mlp_offsets():
        if name.startswith("b"):
            continue
        fan_in = 64 if name == "W1" else 256
        a = np.sqrt(3.0 / fan_in) if name == "w5" else np.sqrt(6.0 / fan_in)
        p[o:o + n] = (2.0 * u[o:o + n] - 1.0) * a
    return p.astype(np.float32)


# ---------------------------------------------------------------- CIF reader of the reference (oracle/_ref)
def ref_read_cif(path):
    """io::readCif of the compiled reference: None on io::CifParseError, else a dict with the
    lattice, per-atom arrays, species symbols and source labels."""
    path = os.fsencode(str(path))
    lat = np.zeros(9)
    n = C.c_uint64()
    slen, llen = C.c_uint64(), C.c_uint64()
    st = ref().ref_read_cif(path, ptr(lat), C.byref(n), C.c_uint64(0), None, None, None, None, None, None, None,
                            None, C.c_uint64(0), C.byref(slen), None, C.c_uint64(0), C.byref(llen))
    if st == 1:
        return None
    if st != 0:
        raise RuntimeError("reference readCif raised a non-CIF exception")
    k = n.value
    arrs = [np.zeros(k) for _ in range(6)]
    sp = np.zeros(k, dtype=np.uint32)
    sym = C.create_string_buffer(slen.value + 1)
    lab = C.create_string_buffer(llen.value + 1)
    st = ref().ref_read_cif(path, ptr(lat), C.byref(n), C.c_uint64(k), *[ptr(a) for a in arrs], ptr(sp, u32p), sym,
                            C.c_uint64(slen.value + 1), C.byref(slen), lab, C.c_uint64(llen.value + 1), C.byref(llen))
    assert st == 0
    names = ("x", "y", "z", "mass", "charge", "occupancy")
    out = dict(zip(names, arrs))
    out.update(lattice=lat.reshape(3, 3), species=sp,
               species_symbols=sym.value.decode().split("\n") if slen.value else [],
               source_labels=lab.value.decode().split("\n") if llen.value else [])
    return out


def ref_cif_document(text: str):
    """io::parseCifDocument of the compiled reference, flattened (see ref_shim.cc): None on
    io::CifParseError."""
    b = text.encode()
    n = C.c_uint64()
    st = ref().ref_cif_document(b, C.c_uint64(len(b)), None, C.c_uint64(0), C.byref(n))
    if st == 1:
        return None
    assert st == 0
    buf = C.create_string_buffer(n.value + 1)
    ref().ref_cif_document(b, C.c_uint64(len(b)), buf, C.c_uint64(n.value + 1), C.byref(n))
    return buf.raw[:n.value].decode()


def ref_try_cif_number(token: str):
    v, has = C.c_double(), C.c_int()
    assert ref().ref_cif_try_number(token.encode(), C.byref(v), C.byref(has)) == 0
    return v.value if has.value else None


def ref_symmetry_operation(text: str):
    """(matrix, translation), or None on io::CifParseError."""
    m, t = np.zeros(9), np.zeros(3)
    st = ref().ref_symmetry_operation(text.encode(), ptr(m), ptr(t))
    return None if st == 1 else (m.reshape(3, 3), t)


def ref_element_from_label(label: str):
    buf = C.create_string_buffer(4)
    found = C.c_int()
    assert ref().ref_element_from_label(label.encode(), buf, C.byref(found)) == 0
    return buf.value.decode() if found.value else None


def ref_gcmc_run_grid(lj: LJ, fw: Framework, sites, fugacity_pa, temperature, seed, stream, steps, spacing, guests,
                      cache_dir, ewald=None, max_translation=1.5, max_rotation=0.5):
    """The reference engine with its FrameworkEnergyGrid (built through loadOrBuildFrameworkEnergyGrid
    in cache_dir): (molecule count, (x, y, z, q) of all particles, committed Ewald energy)."""
    sx = np.array([s[:3] for s in sites], dtype=np.float64).ravel()
    ss = np.array([s[3] for s in sites], dtype=np.uint32)
    sq = np.array([s[4] for s in sites], dtype=np.float64)
    g = np.ascontiguousarray(sorted(set(guests)), dtype=np.uint32)
    cap = fw.n + 3 * len(sites) * 2000
    out = [np.zeros(cap) for _ in range(4)]
    npart, nmol, e = C.c_uint64(), C.c_uint64(), C.c_double()
    alpha, rc_real, kmax = ewald if ewald is not None else (1.0, 1.0, 1)
    rc = ref().ref_gcmc_run_grid(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                                 lj.truncation, lj.mixing, ptr(np.ascontiguousarray(fw.m)), C.c_uint64(fw.n),
                                 ptr(fw.x), ptr(fw.y), ptr(fw.z), ptr(fw.q), ptr(fw.species, u32p), len(sites), ptr(sx),
                                 ptr(ss, u32p), ptr(sq), C.c_double(fugacity_pa), C.c_double(temperature),
                                 C.c_double(max_translation), C.c_double(max_rotation), 1 if ewald is not None else 0,
                                 C.c_double(alpha), C.c_double(rc_real), kmax, C.c_uint64(seed), C.c_uint64(stream),
                                 C.c_uint64(steps), C.c_uint64(cap), ptr(out[0]), ptr(out[1]), ptr(out[2]), ptr(out[3]),
                                 C.byref(npart), C.byref(nmol), C.byref(e), C.c_double(spacing),
                                 os.fsencode(str(cache_dir)), ptr(g, u32p), C.c_int(len(g)))
    if rc != 0:
        raise RuntimeError("reference engine failed")
    n = npart.value
    return nmol.value, tuple(a[:n] for a in out), e.value


def ref_grid_cache_hash(fw: Framework, lj: LJ, guests, spacing) -> int:
    """The compiled reference's makeFrameworkEnergyGridCacheKey(...).hash() for these inputs."""
    g = np.ascontiguousarray(guests, dtype=np.uint32)
    m = np.ascontiguousarray(fw.m)
    h = C.c_uint64()
    rc = ref().ref_grid_cache_hash(C.c_uint32(len(lj.eps)), ptr(lj.eps), ptr(lj.sigma), C.c_double(lj.cutoff),
                                   lj.truncation, lj.mixing, ptr(m), C.c_uint64(fw.n), ptr(fw.x), ptr(fw.y), ptr(fw.z),
                                   ptr(fw.species, u32p), ptr(g, u32p), len(g), C.c_double(spacing), C.byref(h))
    if rc != 0:
        raise RuntimeError("reference key hash failed")
    return h.value

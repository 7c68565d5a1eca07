"""aleator_amd -- MI355X (gfx950) adsorption hot path behind the C ABI of include/aleator_gpu.h.

Python binding (ctypes) of the in-tree HIP library aleator_amd/lib/libaleator_gpu.so.  The
library is the only compute path: if it is missing, importing the binding raises instead of
falling back to anything else.  Build it with `make -C aleator_amd` (or __graft_entry__.build()).

Objects mirror the reference's operator interfaces:
  Framework.grid(...)        FrameworkEnergyGrid constructor (framework_energy_grid.cc:193-249)
  Framework.widom(...)       MonteCarloEngine::widomInsertionHenryCoefficient (monte_carlo_engine.cc:850-917)
  EwaldState                 forcefield::EwaldIncrementalState (ewald_incremental_state.hpp:70-171)
Errors raise the Python analogue of the reference's C++ exception (ValueError for
std::invalid_argument, IndexError for std::out_of_range, RuntimeError otherwise).
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
import weakref

import numpy as np

__all__ = ["Context", "Framework", "EwaldState", "grid_point_counts", "synth_framework", "library_path",
           "FP64", "MIXED", "TRUNCATED", "SHIFTED", "LINEAR_FORCE_SHIFTED", "NotImplementedInReference",
           "read_cif", "parse_cif", "Structure", "CifDocument", "CifParseError"]

FP64, MIXED = 0, 1
TRUNCATED, SHIFTED, LINEAR_FORCE_SHIFTED = 0, 1, 2
LORENTZ_BERTHELOT, GEOMETRIC = 0, 1

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libaleator_gpu.so")


def library_path() -> str:
    return _LIB_PATH


class NotImplementedInReference(RuntimeError):
    """aleator::NotImplemented (src/core/exceptions.hpp:13-17)."""


class DeviceError(RuntimeError):
    """HIP runtime failure."""


class CifParseError(RuntimeError):
    """io::CifParseError (src/io/cif_error.hpp:8-16)."""


_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)


class _Lattice(C.Structure):
    _fields_ = [("m", C.c_double * 9)]


class _Particles(C.Structure):
    _fields_ = [("x", _dp), ("y", _dp), ("z", _dp), ("charge", _dp), ("species", _u32p), ("n", C.c_uint64)]


class _LJ(C.Structure):
    _fields_ = [("epsilon", _dp), ("sigma", _dp), ("n_species", C.c_uint32), ("cutoff", C.c_double),
                ("truncation", C.c_int32), ("mixing", C.c_int32)]


class _Ewald(C.Structure):
    _fields_ = [("alpha", C.c_double), ("real_space_cutoff", C.c_double), ("kmax", C.c_int32)]


class _ChargedSite(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("charge", C.c_double)]


class _MolSite(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("lj_species", C.c_uint32),
                ("charge", C.c_double)]


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is missing: build the HIP library with `make -C aleator_amd`")
    lib = C.CDLL(_LIB_PATH)
    lib.aleator_gpu_last_error.restype = C.c_char_p
    lib.aleator_gpu_stream.restype = C.c_void_p
    lib.aleator_gpu_stream.argtypes = [C.c_void_p]
    _lib = lib
    return lib


def _check(status: int):
    if status == 0:
        return
    msg = _load().aleator_gpu_last_error().decode(errors="replace")
    raise {1: ValueError, 2: IndexError, 3: RuntimeError, 4: NotImplementedInReference,
           5: DeviceError, 6: CifParseError}.get(status, RuntimeError)(msg)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _lattice(m) -> _Lattice:
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    return _Lattice((C.c_double * 9)(*m))


def grid_point_counts(lattice, spacing: float) -> tuple[int, int, int]:
    """max(1, lround(L_perp/spacing)) per axis (framework_energy_grid.cc:207-211)."""
    n = np.zeros(3, dtype=np.int32)
    lat = _lattice(lattice)
    _check(_load().aleator_gpu_grid_point_counts(C.byref(lat), C.c_double(spacing), _p(n, _i32p)))
    return tuple(int(v) for v in n)


class _CacheKey(C.Structure):
    _fields_ = [("lattice", C.POINTER(_Lattice)), ("atoms", C.POINTER(_Particles)), ("lj", C.POINTER(_LJ)),
                ("guest_species", _u32p), ("n_guest", C.c_int32), ("spacing", C.c_double)]


class GridCacheKey:
    """makeFrameworkEnergyGridCacheKey's inputs (framework_energy_grid.cc:171-191): Cartesian
    atoms and species as given, lattice, per-species (epsilon, sigma), cutoff, truncation,
    guest species, spacing."""

    def __init__(self, lattice, x, y, z, species, epsilon, sigma, cutoff, truncation, guests, spacing):
        self._keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, epsilon, sigma)]
        self._sp = np.ascontiguousarray(species, dtype=np.uint32)
        self._g = np.ascontiguousarray(guests, dtype=np.uint32)
        self._lat = _lattice(lattice)
        xx, yy, zz, eps, sig = self._keep
        self._p = _Particles(_p(xx), _p(yy), _p(zz), None, _p(self._sp, _u32p), len(xx))
        self._lj = _LJ(_p(eps), _p(sig), len(eps), float(cutoff), int(truncation), 0)
        self.key = _CacheKey(C.pointer(self._lat), C.pointer(self._p), C.pointer(self._lj),
                             _p(self._g, _u32p) if len(self._g) else None, len(self._g), float(spacing))

    def hash(self) -> int:
        h = C.c_uint64()
        _check(_load().aleator_grid_cache_key_hash(C.byref(self.key), C.byref(h)))
        return h.value

    def file_path(self, directory: str) -> str:
        buf = C.create_string_buffer(4096)
        _check(_load().aleator_grid_cache_file_path(directory.encode(), C.byref(self.key), buf, C.c_uint64(4096)))
        return buf.value.decode()


def grid_cache_serialize(path: str, key: GridCacheKey, n_points, species, tables):
    """FrameworkEnergyGrid::serialize: tables [n_tables][nz][ny][nx], one per species."""
    tables = np.ascontiguousarray(tables, dtype=np.float64)
    sp = np.ascontiguousarray(species, dtype=np.uint32)
    n = np.ascontiguousarray(n_points, dtype=np.int32)
    _check(_load().aleator_grid_cache_serialize(path.encode(), C.byref(key.key), _p(n, _i32p), _p(sp, _u32p),
                                                C.c_int32(len(sp)), _p(tables)))


def grid_cache_load(directory: str, key: GridCacheKey):
    """FrameworkEnergyGrid::loadFromCache: None on a miss, else (n_points, species, tables
    [n_tables][nz][ny][nx]); RuntimeError for a mismatching or corrupt file."""
    n = np.zeros(3, dtype=np.int32)
    nt, hit = C.c_int32(), C.c_int32()
    _check(_load().aleator_grid_cache_load(directory.encode(), C.byref(key.key), _p(n, _i32p), None, C.byref(nt),
                                           None, C.c_uint64(0), C.byref(hit)))
    if not hit.value:
        return None
    sp = np.zeros(max(nt.value, 1), dtype=np.uint32)
    out = np.zeros((nt.value, n[2], n[1], n[0]))
    _check(_load().aleator_grid_cache_load(directory.encode(), C.byref(key.key), _p(n, _i32p), _p(sp, _u32p),
                                           C.byref(nt), _p(out), C.c_uint64(out.size), C.byref(hit)))
    return tuple(int(v) for v in n), sp[:nt.value].copy(), out


# ---------------------------------------------------------------- CIF structure files
class Structure:
    """io::StructureData (io/structure_io.hpp:18-37): lattice rows, Cartesian positions, masses,
    charges, species indices into species_symbols, occupancies and source labels, per atom."""

    def __init__(self, handle):
        lib = _load()
        try:
            n, ns = C.c_uint64(), C.c_uint32()
            _check(lib.aleator_structure_size(handle, C.byref(n), C.byref(ns)))
            lat = _Lattice()
            _check(lib.aleator_structure_lattice(handle, C.byref(lat)))
            self.lattice = np.array(lat.m[:]).reshape(3, 3)
            k = n.value
            self.x, self.y, self.z, self.mass, self.charge, self.occupancy = (np.zeros(k) for _ in range(6))
            self.species = np.zeros(k, dtype=np.uint32)
            _check(lib.aleator_structure_atoms(handle, _p(self.x), _p(self.y), _p(self.z), _p(self.mass),
                                               _p(self.charge), _p(self.species, _u32p), _p(self.occupancy)))
            out = C.c_char_p()
            self.species_symbols = []
            for i in range(ns.value):
                _check(lib.aleator_structure_species_symbol(handle, C.c_uint32(i), C.byref(out)))
                self.species_symbols.append(out.value.decode())
            self.source_labels = []
            for i in range(k):
                _check(lib.aleator_structure_source_label(handle, C.c_uint64(i), C.byref(out)))
                self.source_labels.append(out.value.decode())
        finally:
            lib.aleator_structure_destroy(handle)

    def __len__(self):
        return len(self.x)

    def positions(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


def read_cif(path) -> Structure:
    """io::readCif (io/structure_io.cc:173-250); CifParseError on malformed input."""
    h = C.c_void_p()
    _check(_load().aleator_read_cif(os.fsencode(path), C.byref(h)))
    return Structure(h)


def parse_cif(text: str) -> Structure:
    """read_cif from CIF text in memory."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    h = C.c_void_p()
    _check(_load().aleator_parse_cif(b, C.c_uint64(len(b)), C.byref(h)))
    return Structure(h)


class CifDocument:
    """io::CifDocument (io/cif_document.hpp:26-42) of parseCifDocument(text): block_name, tags
    (lowercased tag -> raw value) and loops (list of (columns, rows))."""

    def __init__(self, text: str):
        lib = _load()
        b = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        _check(lib.aleator_cif_document_parse(b, C.c_uint64(len(b)), C.byref(h)))
        self._h = h
        try:
            out = C.c_char_p()
            _check(lib.aleator_cif_document_block_name(h, C.byref(out)))
            self.block_name = out.value.decode()
            nt = C.c_uint64()
            _check(lib.aleator_cif_document_num_tags(h, C.byref(nt)))
            self.tags = {}
            val = C.c_char_p()
            for i in range(nt.value):
                _check(lib.aleator_cif_document_tag(h, C.c_uint64(i), C.byref(out), C.byref(val)))
                self.tags[out.value.decode()] = val.value.decode()
            nl = C.c_uint64()
            _check(lib.aleator_cif_document_num_loops(h, C.byref(nl)))
            self.loops = []
            for li in range(nl.value):
                nc, nr = C.c_uint64(), C.c_uint64()
                _check(lib.aleator_cif_document_loop_shape(h, C.c_uint64(li), C.byref(nc), C.byref(nr)))
                cols = []
                for c in range(nc.value):
                    _check(lib.aleator_cif_document_loop_column(h, C.c_uint64(li), C.c_uint64(c), C.byref(out)))
                    cols.append(out.value.decode())
                rows = []
                for r in range(nr.value):
                    row = []
                    for c in range(nc.value):
                        _check(lib.aleator_cif_document_loop_value(h, C.c_uint64(li), C.c_uint64(r), C.c_uint64(c),
                                                                   C.byref(out)))
                        row.append(out.value.decode())
                    rows.append(row)
                self.loops.append((cols, rows))
        except Exception:
            self.close()
            raise

    def find_tag(self, tag: str):
        """CifDocument::findTag: the raw value, or None."""
        out = C.c_char_p()
        _check(_load().aleator_cif_document_find_tag(self._h, tag.encode(), C.byref(out)))
        return None if out.value is None else out.value.decode()

    def find_loop(self, columns):
        """CifDocument::findLoopWithColumns: index into loops of the first loop with every column, or None."""
        arr = (C.c_char_p * max(len(columns), 1))(*[c.encode() for c in columns])
        idx = C.c_int64()
        _check(_load().aleator_cif_document_find_loop(self._h, arr, C.c_uint32(len(columns)), C.byref(idx)))
        return None if idx.value < 0 else idx.value

    def close(self):
        if self._h:
            _load().aleator_cif_document_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def try_parse_cif_number(token: str):
    """tryParseCifNumber (cif_document.cc:166-183): float, or None for '?', '.' or non-numbers."""
    v, has = C.c_double(), C.c_int32()
    _check(_load().aleator_cif_try_parse_number(token.encode(), C.byref(v), C.byref(has)))
    return v.value if has.value else None


def parse_cif_number(token: str) -> float:
    """parseCifNumber (cif_document.cc:158-164): CifParseError when there is no number."""
    v = C.c_double()
    _check(_load().aleator_cif_parse_number(token.encode(), C.byref(v)))
    return v.value


def parse_symmetry_operation(text: str):
    """parseSymmetryOperation (symmetry_operation.cc:118-150): (matrix 3x3, translation 3)."""
    m, t = np.zeros(9), np.zeros(3)
    _check(_load().aleator_parse_symmetry_operation(text.encode(), _p(m), _p(t)))
    return m.reshape(3, 3), t


def element_from_label(label: str):
    """elementFromLabel (element_table.cc:81-99): canonical symbol, or None."""
    buf = C.create_string_buffer(4)
    found = C.c_int32()
    _check(_load().aleator_element_from_label(label.encode(), buf, C.byref(found)))
    return buf.value.decode() if found.value else None


def atomic_mass(symbol: str):
    """atomicMass (element_table.cc:67-79): amu, or None for an unknown (case-exact) symbol."""
    m, found = C.c_double(), C.c_int32()
    _check(_load().aleator_atomic_mass(symbol.encode(), C.byref(m), C.byref(found)))
    return m.value if found.value else None


def philox_uniforms(seed: int, stream: int, first_call_index: int, n: int) -> np.ndarray:
    """Philox4x32Rng(seed, stream) uniforms at call indices first .. first+n-1."""
    out = np.zeros(n)
    _check(_load().aleator_philox_uniforms(C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(first_call_index),
                                           C.c_uint64(n), _p(out)))
    return out


def supercell_replication(lattice, cutoff: float) -> tuple[int, int, int]:
    """minimumSupercellReplication (supercell.cc:7-20)."""
    n = np.zeros(3, dtype=np.int32)
    lat = _lattice(lattice)
    _check(_load().aleator_supercell_replication(C.byref(lat), C.c_double(cutoff), _p(n, _i32p)))
    return tuple(int(v) for v in n)


def replicate_supercell(lattice, x, y, z, charge, species, reps):
    """replicateSupercell (supercell.cc:22-60): returns (lattice, x, y, z, charge, species)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, charge)]
    sp = np.ascontiguousarray(species, dtype=np.uint32)
    n = len(keep[0]) * int(np.prod(reps))
    out = [np.zeros(n) for _ in range(4)]
    osp = np.zeros(n, dtype=np.uint32)
    parts = _Particles(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(sp, _u32p), len(keep[0]))
    lat = _lattice(lattice)
    olat = _Lattice()
    _check(_load().aleator_replicate_supercell(C.byref(parts), C.byref(lat), C.c_int32(reps[0]), C.c_int32(reps[1]),
                                               C.c_int32(reps[2]), _p(out[0]), _p(out[1]), _p(out[2]), _p(out[3]),
                                               _p(osp, _u32p), C.byref(olat)))
    return np.array(olat.m[:]).reshape(3, 3), out[0], out[1], out[2], out[3], osp


def peng_robinson_fugacity(tc, pc, omega, temperature, pressure) -> float:
    """PengRobinson::fugacityPascal (peng_robinson.cc:10-110)."""
    out = C.c_double()
    _check(_load().aleator_peng_robinson_fugacity(C.c_double(tc), C.c_double(pc), C.c_double(omega),
                                                  C.c_double(temperature), C.c_double(pressure), C.byref(out)))
    return out.value


def lj_tail_insertion_delta(epsilon, sigma, cutoff, lattice, framework_species, sites, mixing=0) -> float:
    """tailEnergyCorrection(framework + molecule) - tailEnergyCorrection(framework)."""
    eps = np.ascontiguousarray(epsilon, dtype=np.float64)
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    lj = _LJ(_p(eps), _p(sig), len(eps), cutoff, 0, mixing)
    sp = np.ascontiguousarray(framework_species, dtype=np.uint32)
    arr = Framework._sites(sites)
    lat = _lattice(lattice)
    out = C.c_double()
    _check(_load().aleator_lj_tail_insertion_delta(C.byref(lj), C.byref(lat), _p(sp, _u32p), C.c_uint64(len(sp)),
                                                   arr, C.c_int32(len(sites)), C.byref(out)))
    return out.value


def erf_kernel_poly(alpha: float, rmax: float, tol: float, min_degree: int = 0):
    """The device's erf(alpha r)/r polynomial: (coefficients lowest first, xs) with
    erf(alpha r)/r = sum_k c_k x^k, x = r^2 xs - 1, for r in [0, rmax]."""
    coef = np.zeros(29)
    xs = C.c_double()
    deg = C.c_int32()
    _check(_load().aleator_erf_kernel_poly(C.c_double(alpha), C.c_double(rmax), C.c_double(tol),
                                           C.c_int32(min_degree), _p(coef), C.byref(xs), C.byref(deg)))
    return coef[:deg.value + 1].copy(), xs.value


def synth_framework(lattice, n_atoms: int, seed=1, stream=0, min_dist=1.0,
                    proportions=(0.45, 0.30, 0.17, 0.08), charge_sigma=0.0, charge_stream=2):
    """Synthetic framework (no reference counterpart): returns (x, y, z, charge, species)."""
    x, y, z, q = (np.zeros(n_atoms) for _ in range(4))
    sp = np.zeros(n_atoms, dtype=np.uint32)
    props = np.ascontiguousarray(proportions, dtype=np.float64)
    lat = _lattice(lattice)
    _check(_load().aleator_synth_framework(C.byref(lat), C.c_uint64(n_atoms), C.c_uint64(seed),
                                           C.c_uint64(stream), C.c_double(min_dist), _p(props),
                                           C.c_uint32(len(props)), C.c_double(charge_sigma),
                                           C.c_uint64(charge_stream), _p(x), _p(y), _p(z), _p(q), _p(sp, _u32p)))
    return x, y, z, q, sp


_creation = itertools.count()


class Context:
    """One GPU and its HIP stream (no reference counterpart).  Closing it first closes every
    object still open on it (frameworks, plans, chains, ...), newest first, so that no native
    handle outlives the context it points into."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        self._owned = []  # (creation number, weak reference) of objects created on this context
        _check(_load().aleator_gpu_init(C.c_int32(device), C.byref(self._h)))

    def _adopt(self, obj):
        self._owned = [(n, r) for n, r in self._owned if r() is not None]
        self._owned.append((next(_creation), weakref.ref(obj)))

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return _load().aleator_gpu_stream(self._h)

    def timing(self, enable: bool) -> tuple[float, int]:
        """HIP-event time (ms) and count of dominant-kernel launches since the last call;
        then enables or disables recording."""
        ms = C.c_double()
        n = C.c_uint64()
        _check(_load().aleator_gpu_timing(self._h, C.c_int32(1 if enable else 0), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self._h:
            for _, r in sorted(self._owned, key=lambda e: e[0], reverse=True):
                obj = r()
                if obj is not None:
                    obj.close()
            self._owned = []
            _load().aleator_gpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Framework:
    """Device-resident rigid framework with its LJ table and (optionally) Ewald S(k)."""

    def __init__(self, ctx: Context, lattice, x, y, z, species, epsilon, sigma, cutoff,
                 truncation=TRUNCATED, mixing=LORENTZ_BERTHELOT, charge=None, ewald=None):
        self.ctx = ctx
        self.lattice = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        sp = np.ascontiguousarray(species, dtype=np.uint32)
        q = np.ascontiguousarray(charge, dtype=np.float64) if charge is not None else None
        eps = np.ascontiguousarray(epsilon, dtype=np.float64)
        sig = np.ascontiguousarray(sigma, dtype=np.float64)
        parts = _Particles(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(q) if q is not None else None,
                           _p(sp, _u32p), len(keep[0]))
        lj = _LJ(_p(eps), _p(sig), len(eps), cutoff, truncation, mixing)
        ew = _Ewald(*ewald) if ewald is not None else None
        lat = _lattice(self.lattice)
        self._h = C.c_void_p()
        _check(_load().aleator_gpu_framework_create(ctx.handle, C.byref(lat), C.byref(parts), C.byref(lj),
                                                    C.byref(ew) if ew is not None else None, C.byref(self._h)))
        ctx._adopt(self)
        self.n_species = len(eps)

    @property
    def handle(self):
        return self._h

    @property
    def num_k(self) -> int:
        n = C.c_uint64()
        _check(_load().aleator_gpu_framework_num_k(self._h, C.byref(n)))
        return n.value

    def close(self):
        if self._h:
            _load().aleator_gpu_framework_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_lj_tail(self, enable: bool = True):
        """Add the analytic LJ tail of one insertion to grid tables and Widom energies
        (aleator_gpu_framework_set_lj_tail; off by default, as in the reference)."""
        _check(_load().aleator_gpu_framework_set_lj_tail(self._h, C.c_int32(1 if enable else 0)))

    def nonfinite_count(self) -> int:
        """Non-finite table values written by this framework's grid builds so far
        (aleator_gpu_framework_nonfinite_count)."""
        n = C.c_uint64()
        _check(_load().aleator_gpu_framework_nonfinite_count(self._h, C.byref(n)))
        return n.value

    def grid(self, guests, n_points, coulomb=False, precision=FP64, cap=1.0e5, grid_lattice=None):
        """Energy tables [len(guests) (+1 Coulomb)][nz][ny][nx] in K (Coulomb: K/e)."""
        g = np.ascontiguousarray(guests, dtype=np.uint32)
        n = np.ascontiguousarray(n_points, dtype=np.int32)
        out = np.zeros((len(g) + (1 if coulomb else 0), n[2], n[1], n[0]))
        gl = _lattice(grid_lattice) if grid_lattice is not None else None
        _check(_load().aleator_gpu_grid_build(self._h, C.byref(gl) if gl is not None else None,
                                              _p(g, _u32p) if len(g) else None, C.c_int32(len(g)),
                                              C.c_int32(1 if coulomb else 0), _p(n, _i32p), C.c_int32(precision),
                                              C.c_double(cap), _p(out)))
        return out

    def grid_device(self, guests, n_points, out_ptr: int, coulomb=False, precision=FP64, cap=1.0e5,
                    grid_lattice=None):
        """Device-output variant: writes to device pointer `out_ptr` on the context stream."""
        g = np.ascontiguousarray(guests, dtype=np.uint32)
        n = np.ascontiguousarray(n_points, dtype=np.int32)
        gl = _lattice(grid_lattice) if grid_lattice is not None else None
        _check(_load().aleator_gpu_grid_build_device(self._h, C.byref(gl) if gl is not None else None,
                                                     _p(g, _u32p) if len(g) else None, C.c_int32(len(g)),
                                                     C.c_int32(1 if coulomb else 0), _p(n, _i32p),
                                                     C.c_int32(precision), C.c_double(cap), C.c_void_p(out_ptr)))

    @staticmethod
    def _sites(sites):
        arr = (_MolSite * len(sites))()
        for i, s in enumerate(sites):
            arr[i] = _MolSite(float(s[0]), float(s[1]), float(s[2]), int(s[3]), float(s[4]))
        return arr

    def widom(self, sites, n_trials, temperature, seed=1, stream=0, first_call_index=0, precision=FP64,
              return_du=True):
        """Returns (sum_t exp(-dU_t/T), dU per trial or None); K_H = sum / (n_trials * T)."""
        arr = self._sites(sites)
        du = np.zeros(n_trials) if return_du else None
        s = C.c_double()
        _check(_load().aleator_gpu_widom(self._h, arr, C.c_int32(len(sites)), C.c_uint64(seed), C.c_uint64(stream),
                                         C.c_uint64(first_call_index), C.c_uint64(n_trials),
                                         C.c_double(temperature), C.c_int32(precision),
                                         _p(du) if du is not None else None, C.byref(s)))
        return s.value, du

    def widom_device(self, sites, n_trials, temperature, du_ptr: int, sum_ptr: int, seed=1, stream=0,
                     first_call_index=0, precision=FP64):
        arr = self._sites(sites)
        _check(_load().aleator_gpu_widom_device(self._h, arr, C.c_int32(len(sites)), C.c_uint64(seed),
                                                C.c_uint64(stream), C.c_uint64(first_call_index),
                                                C.c_uint64(n_trials), C.c_double(temperature), C.c_int32(precision),
                                                C.c_void_p(du_ptr) if du_ptr else None, C.c_void_p(sum_ptr)))

    def henry_coefficient(self, sites, n_trials, temperature, **kw):
        """widomInsertionHenryCoefficient: (sum exp(-dU/T) / n_trials) / T."""
        s, _ = self.widom(sites, n_trials, temperature, return_du=False, **kw)
        return (s / n_trials) / temperature


class GridPlan:
    """Reusable (batched, optionally graph-captured, optionally autotuned) grid build over device
    outputs.  outputs: device pointers, one per framework, each sized (n_guest + coulomb) * nodes
    doubles."""

    def __init__(self, ctx: Context, frameworks, guests, n_points, outputs, coulomb=False, precision=FP64,
                 cap=1.0e5, grid_lattices=None, tile=-1, use_graph=False):
        n = len(frameworks)
        fws = (C.c_void_p * n)(*[f.handle.value for f in frameworks])
        g = np.ascontiguousarray(guests, dtype=np.uint32)
        pts = np.ascontiguousarray(n_points, dtype=np.int32).reshape(n, 3)
        outs = (C.c_void_p * n)(*[int(o) for o in outputs])
        lats = None
        if grid_lattices is not None:
            lats = (_Lattice * n)(*[_lattice(m) for m in grid_lattices])
        self._keep = (frameworks, g, pts, outs, lats)
        self._h = C.c_void_p()
        _check(_load().aleator_gpu_grid_plan_create(ctx.handle, C.c_int32(n), fws, lats,
                                                    _p(g, _u32p) if len(g) else None, C.c_int32(len(g)),
                                                    C.c_int32(1 if coulomb else 0), _p(pts, _i32p),
                                                    C.c_int32(precision), C.c_double(cap), outs, C.c_int32(tile),
                                                    C.c_int32(1 if use_graph else 0), C.byref(self._h)))
        ctx._adopt(self)

    def execute(self):
        _check(_load().aleator_gpu_grid_plan_execute(self._h))

    def tile(self):
        t = C.c_int32()
        ms = (C.c_float * 4)()
        _check(_load().aleator_gpu_grid_plan_tile(self._h, C.byref(t), ms))
        return t.value, list(ms)

    def atoms_per_pass(self) -> int:
        """Framework atoms staged per pass of the LDS list build, tuned with the tile shape."""
        n = C.c_int32()
        _check(_load().aleator_gpu_grid_plan_atoms_per_pass(self._h, C.byref(n)))
        return n.value

    def erf_degree(self) -> tuple:
        """Degrees of the Coulomb erf polynomials: (inner entries, shell entries); 0 without one."""
        d, f = C.c_int32(), C.c_int32()
        _check(_load().aleator_gpu_grid_plan_erf_degree(self._h, C.byref(d), C.byref(f)))
        return d.value, f.value

    def close(self):
        if self._h:
            _load().aleator_gpu_grid_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A plain hipMalloc'd buffer for device-output entry points (no torch needed)."""

    def __init__(self, n_doubles: int):
        self.nbytes = int(n_doubles) * 8
        self.ptr = C.c_void_p()
        if _hip_rt().hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 8))) != 0:
            raise DeviceError("hipMalloc failed")

    @classmethod
    def from_array(cls, a) -> "DeviceBuffer":
        """A device copy of a host array (its bytes, any dtype)."""
        a = np.ascontiguousarray(a)
        buf = cls((a.nbytes + 7) // 8)
        if a.nbytes and _hip_rt().hipMemcpy(buf.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) != 0:
            raise DeviceError("hipMemcpy failed")
        return buf

    def to_host(self, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.nbytes // 8)
        _hip_rt().hipDeviceSynchronize()
        if _hip_rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) != 0:
            raise DeviceError("hipMemcpy failed")
        return out.view(dtype)

    def free(self):
        if self.ptr:
            _hip_rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __int__(self):
        return self.ptr.value


class _ChainDesc(C.Structure):
    _fields_ = [("fugacity_pa", C.c_double), ("temperature", C.c_double), ("seed", C.c_uint64),
                ("stream", C.c_uint64)]


class _ChainStats(C.Structure):
    _fields_ = [("n_molecules", C.c_int32), ("error", C.c_int32), ("steps", C.c_uint64),
                ("sum_molecules", C.c_double), ("attempted", C.c_uint64 * 4), ("accepted", C.c_uint64 * 4)]


_hip = None


def _hip_rt():
    """HIP runtime entry points for the few device buffers the binding itself needs."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class Gcmc:
    """Batched single-species GCMC chains (MonteCarloEngine::run, monte_carlo_engine.cc:405-671),
    one persistent launch advancing every chain.  chains: iterable of
    (fugacity_pa, temperature, seed, stream)."""

    def __init__(self, fw: Framework, sites, chains, max_translation=1.5, max_rotation=0.5, max_molecules=512):
        self.fw = fw
        self.ns = len(sites)
        self.max_molecules = int(max_molecules)
        chains = list(chains)
        self.n_chains = len(chains)
        arr = Framework._sites(sites)
        desc = (_ChainDesc * len(chains))(*[_ChainDesc(float(f), float(t), int(s), int(st)) for f, t, s, st in chains])
        self._h = C.c_void_p()
        _check(_load().aleator_gpu_gcmc_create(fw.handle, arr, C.c_int32(len(sites)), desc, C.c_int32(len(chains)),
                                               C.c_double(max_translation), C.c_double(max_rotation),
                                               C.c_int32(max_molecules), C.byref(self._h)))
        fw.ctx._adopt(self)

    def close(self):
        if self._h:
            _load().aleator_gpu_gcmc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n_steps: int, trace: bool = False):
        """Advances every chain by n_steps; with trace, returns int32 [n_chains][n_steps] molecule
        counts after each step.  Raises RuntimeError when a chain stopped because an accepted
        insertion would exceed max_molecules (the chain keeps its state before that step; its
        trace rows from there on read -1)."""
        if not trace:
            _check(_load().aleator_gpu_gcmc_run(self._h, C.c_uint64(n_steps), None))
            self._raise_on_error()
            return None
        hip = _hip_rt()
        nbytes = self.n_chains * n_steps * 4
        d = C.c_void_p()
        if hip.hipMalloc(C.byref(d), C.c_size_t(nbytes)) != 0:
            raise DeviceError("hipMalloc failed")
        try:
            _check(_load().aleator_gpu_gcmc_run(self._h, C.c_uint64(n_steps), d))
            out = np.zeros((self.n_chains, n_steps), dtype=np.int32)
            if hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, C.c_size_t(nbytes), 2) != 0:  # device to host
                raise DeviceError("hipMemcpy failed")
        finally:
            hip.hipFree(d)
        self._raise_on_error(out)
        return out

    def _raise_on_error(self, trace=None):
        bad = [c for c, s in enumerate(self.stats()) if s["error"]]
        if bad:
            err = RuntimeError(f"gcmc: chain(s) {bad} stopped: an accepted insertion would exceed max_molecules "
                               f"({self.max_molecules})")
            err.trace = trace  # with trace: rows from the stopping step on read -1
            raise err

    def stats(self):
        out = (_ChainStats * self.n_chains)()
        _check(_load().aleator_gpu_gcmc_stats(self._h, out))
        return [{"n_molecules": s.n_molecules, "error": s.error, "steps": s.steps, "sum_molecules": s.sum_molecules,
                 "attempted": list(s.attempted), "accepted": list(s.accepted)} for s in out]

    def reset_stats(self):
        _check(_load().aleator_gpu_gcmc_reset_stats(self._h))

    def use_graph(self, enable: bool = True):
        """Replay run() (without trace) from a captured hipGraph (aleator_gpu_gcmc_use_graph)."""
        _check(_load().aleator_gpu_gcmc_use_graph(self._h, C.c_int32(1 if enable else 0)))

    def set_proposal(self, weights, nb, reps, uniform_mix=0.05):
        """Biased insertion from bin weights [nb0][nb1][nb2] over the cell divided by reps
        (aleator_gpu_gcmc_set_proposal); None restores uniform insertion."""
        nb_a = (C.c_int32 * 3)(*[int(v) for v in nb])
        reps_a = (C.c_int32 * 3)(*[int(v) for v in reps])
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
        if w is not None and w.size != int(np.prod(nb)):
            raise ValueError("weights must hold nb[0]*nb[1]*nb[2] entries")
        _check(_load().aleator_gpu_gcmc_set_proposal(self._h, None if w is None else _p(w), nb_a, reps_a,
                                                     C.c_double(uniform_mix)))

    def set_energy_grid(self, n_points, species, tables):
        """Framework LJ from energy-grid tables (aleator_gpu_gcmc_set_energy_grid): tables is a
        host array [len(species)][n2][n1][n0], or a device pointer (DeviceBuffer or int) to one;
        None restores the direct framework scan.  IndexError when a site's species has no table
        (the reference's std::out_of_range)."""
        if tables is None:
            _check(_load().aleator_gpu_gcmc_set_energy_grid(self._h, None, None, C.c_int32(0), None))
            return
        n = (C.c_int32 * 3)(*[int(v) for v in n_points])
        sp = np.ascontiguousarray(species, dtype=np.uint32)
        if isinstance(tables, np.ndarray):
            t = np.ascontiguousarray(tables, dtype=np.float64)
            if t.size != sp.size * int(np.prod([int(v) for v in n_points])):
                raise ValueError("tables must hold len(species) * n0 * n1 * n2 values")
            _check(_load().aleator_gpu_gcmc_set_energy_grid_host(self._h, n, _p(sp, _u32p), C.c_int32(sp.size), _p(t)))
        else:
            _check(_load().aleator_gpu_gcmc_set_energy_grid(self._h, n, _p(sp, _u32p), C.c_int32(sp.size),
                                                            C.cast(C.c_void_p(int(tables)), _dp)))

    def set_guest_cells(self, mode: int):
        """Guest-guest candidates through guest cells (aleator_gpu_gcmc_set_guest_cells): 0 full
        scan (the default), 1 cells when every axis has at least 4 and a step starts with 1024 or
        more guest sites, 2 cells in every step whenever the guest sites fit (IndexError above
        4096)."""
        _check(_load().aleator_gpu_gcmc_set_guest_cells(self._h, C.c_int32(int(mode))))

    def guest_cells(self):
        """Guest cells per lattice axis in use; (0, 0, 0) for the full scan."""
        n = (C.c_int32 * 3)()
        _check(_load().aleator_gpu_gcmc_guest_cells(self._h, n))
        return tuple(int(v) for v in n)

    def positions(self, chain: int) -> np.ndarray:
        n = C.c_int32()
        _check(_load().aleator_gpu_gcmc_positions(self._h, C.c_int32(chain), None, C.c_uint64(0), C.byref(n)))
        out = np.zeros((n.value, self.ns, 3))
        _check(_load().aleator_gpu_gcmc_positions(self._h, C.c_int32(chain), _p(out), C.c_uint64(out.size),
                                                  C.byref(n)))
        return out

    def electrostatic_energy(self, chain: int) -> float:
        out = C.c_double()
        _check(_load().aleator_gpu_gcmc_electrostatic_energy(self._h, C.c_int32(chain), C.byref(out)))
        return out.value

    def resync_check(self):
        """Per chain: the largest relative drift |fresh - committed| / max(1, |fresh|) of the Ewald
        energy seen at a resync (aleator_gpu_gcmc_resync_check), and the number of resyncs."""
        drift = np.zeros(self.n_chains)
        count = np.zeros(self.n_chains, dtype=np.uint64)
        _check(_load().aleator_gpu_gcmc_resync_check(self._h, _p(drift), _p(count, C.POINTER(C.c_uint64))))
        return drift, count

    def save_chain(self, chain: int) -> bytes:
        """One chain's state as the checkpoint record of aleator_gpu_gcmc_save_chain."""
        n = C.c_uint64()
        _check(_load().aleator_gpu_gcmc_checkpoint_size(self._h, C.c_int32(chain), C.byref(n)))
        buf = C.create_string_buffer(n.value)
        written = C.c_uint64()
        _check(_load().aleator_gpu_gcmc_save_chain(self._h, C.c_int32(chain), buf, n, C.byref(written)))
        return buf.raw[:written.value]

    def load_chain(self, chain: int, record: bytes):
        """Restores one chain from a save_chain record (aleator_gpu_gcmc_load_chain)."""
        _check(_load().aleator_gpu_gcmc_load_chain(self._h, C.c_int32(chain), C.c_char_p(bytes(record)),
                                                   C.c_uint64(len(record))))


def grid_interpolate_device(ctx: "Context", lattice, n_points, table_ptr: int, positions_ptr: int, n: int,
                            out_ptr: int):
    """FrameworkEnergyGrid::interpolate of one device table at n device positions [n][3]."""
    lat = _lattice(lattice)
    npts = np.ascontiguousarray(n_points, dtype=np.int32)
    _check(_load().aleator_gpu_grid_interpolate(ctx.handle, C.byref(lat), _p(npts, _i32p), C.c_void_p(int(table_ptr)),
                                                C.c_void_p(int(positions_ptr)), C.c_uint64(n),
                                                C.c_void_p(int(out_ptr))))


def grid_interpolate(ctx: "Context", lattice, n_points, table, positions) -> np.ndarray:
    """Host convenience of grid_interpolate_device: table [nz][ny][nx], positions [n][3]."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    d_t = DeviceBuffer.from_array(np.ascontiguousarray(table, dtype=np.float64))
    d_p = DeviceBuffer.from_array(pos)
    d_o = DeviceBuffer(len(pos))
    grid_interpolate_device(ctx, lattice, n_points, int(d_t), int(d_p), len(pos), int(d_o))
    return d_o.to_host()[:len(pos)].copy()


class Mlp:
    """Fourier-feature MLP (64 -> 4 x 256 SiLU -> 1) trained and evaluated on the device with
    BF16 MFMA GEMMs (aleator_gpu_mlp_*).  Row inputs are fractional positions, double [n][3]."""

    def __init__(self, ctx: Context, max_rows: int = 4096, seed: int | None = 0):
        self.ctx = ctx
        self.max_rows = int(max_rows)
        self._h = C.c_void_p()
        _check(_load().aleator_gpu_mlp_create(ctx.handle, C.c_int32(self.max_rows), C.byref(self._h)))
        ctx._adopt(self)
        if seed is not None:
            self.init(seed)

    @staticmethod
    def num_params() -> int:
        f = _load().aleator_gpu_mlp_num_params
        f.restype = C.c_int64
        return int(f())

    @staticmethod
    def frequencies() -> np.ndarray:
        out = np.zeros(96, dtype=np.int8)
        _check(_load().aleator_gpu_mlp_frequencies(out.ctypes.data_as(C.c_void_p)))
        return out.reshape(32, 3)

    def init(self, seed: int):
        _check(_load().aleator_gpu_mlp_init(self._h, C.c_uint64(seed)))

    def use_graph(self, enable: bool):
        """Replay training iterations from one captured hipGraph (default) or launch each kernel."""
        _check(_load().aleator_gpu_mlp_use_graph(self._h, C.c_int32(1 if enable else 0)))

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        _check(_load().aleator_gpu_mlp_get_params(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("wrong parameter count")
        _check(_load().aleator_gpu_mlp_set_params(self._h, p.ctypes.data_as(C.c_void_p)))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        _check(_load().aleator_gpu_mlp_get_grads(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def predict_device(self, frac_ptr: int, n: int, out_ptr: int):
        _check(_load().aleator_gpu_mlp_predict(self._h, C.c_void_p(int(frac_ptr)), C.c_int64(n),
                                               C.c_void_p(int(out_ptr))))

    def predict(self, frac) -> np.ndarray:
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
        d_in = DeviceBuffer.from_array(frac)
        d_out = DeviceBuffer((len(frac) + 1) // 2)
        self.predict_device(int(d_in), len(frac), int(d_out))
        return d_out.to_host(np.float32)[:len(frac)].copy()

    def loss_grad_device(self, frac_ptr: int, target_ptr: int, n: int) -> float:
        loss = C.c_float()
        _check(_load().aleator_gpu_mlp_loss_grad(self._h, C.c_void_p(int(frac_ptr)), C.c_void_p(int(target_ptr)),
                                                 C.c_int32(n), C.byref(loss)))
        return loss.value

    def loss_grad(self, frac, target) -> float:
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
        d_f = DeviceBuffer.from_array(frac)
        d_t = DeviceBuffer.from_array(np.ascontiguousarray(target, dtype=np.float64))
        return self.loss_grad_device(int(d_f), int(d_t), len(frac))

    def train_device(self, frac_ptr: int, target_ptr: int, n_samples: int, batch: int, n_iters: int, lr: float,
                     seed: int = 0, stream: int = 0, trace: bool = True):
        losses = np.zeros(max(n_iters, 1), dtype=np.float32)
        _check(_load().aleator_gpu_mlp_train(self._h, C.c_void_p(int(frac_ptr)), C.c_void_p(int(target_ptr)),
                                             C.c_int64(n_samples), C.c_int32(batch), C.c_int32(n_iters),
                                             C.c_float(lr), C.c_uint64(seed), C.c_uint64(stream),
                                             losses.ctypes.data_as(C.c_void_p) if trace else None))
        return losses[:n_iters] if trace else None

    def close(self):
        if self._h:
            _load().aleator_gpu_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _charged_sites(sites):
    sites = np.asarray(sites, dtype=np.float64).reshape(-1, 4)
    arr = (_ChargedSite * max(len(sites), 1))()
    for i, r in enumerate(sites):
        arr[i] = _ChargedSite(*r)
    return arr, len(sites)


class EwaldState:
    """forcefield::EwaldIncrementalState on the device, initialised from a charged framework."""

    def __init__(self, fw: Framework):
        self.fw = fw
        self._h = C.c_void_p()
        _check(_load().aleator_gpu_ewald_state_create(fw.handle, C.byref(self._h)))
        fw.ctx._adopt(self)

    def close(self):
        if self._h:
            _load().aleator_gpu_ewald_state_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def propose(self, removed, added) -> float:
        r, nr = _charged_sites(removed)
        a, na = _charged_sites(added)
        out = C.c_double()
        _check(_load().aleator_gpu_ewald_state_propose(self._h, r, C.c_int32(nr), a, C.c_int32(na), C.byref(out)))
        return out.value

    def propose_insertion(self, sites):
        return self.propose([], sites)

    def propose_deletion(self, sites):
        return self.propose(sites, [])

    def commit(self, removed, added):
        r, nr = _charged_sites(removed)
        a, na = _charged_sites(added)
        _check(_load().aleator_gpu_ewald_state_commit(self._h, r, C.c_int32(nr), a, C.c_int32(na)))

    def propose_batch(self, removed, added):
        """removed: [P][nr][4], added: [P][na][4] -> dE[P]."""
        added = np.asarray(added, dtype=np.float64)
        removed = np.asarray(removed, dtype=np.float64)
        p = added.shape[0] if added.size else removed.shape[0]
        na = added.shape[1] if added.size else 0
        nr = removed.shape[1] if removed.size else 0
        a, _ = _charged_sites(added.reshape(-1, 4) if added.size else [])
        r, _ = _charged_sites(removed.reshape(-1, 4) if removed.size else [])
        out = np.zeros(p)
        _check(_load().aleator_gpu_ewald_state_propose_batch(self._h, C.c_uint64(p), r, C.c_int32(nr), a,
                                                             C.c_int32(na), _p(out)))
        return out

    def committed_energy(self) -> float:
        out = C.c_double()
        _check(_load().aleator_gpu_ewald_state_committed_energy(self._h, C.byref(out)))
        return out.value

    @staticmethod
    def _particles(x, y, z, q):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q)]
        return _Particles(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), None, len(keep[0])), keep

    @staticmethod
    def _molecules(molecules):
        off = np.zeros(len(molecules) + 1, dtype=np.uint64)
        flat = []
        for i, m in enumerate(molecules):
            flat.extend(m)
            off[i + 1] = len(flat)
        return off, np.ascontiguousarray(flat if flat else [0], dtype=np.uint64)

    def resync(self, x, y, z, q, molecules):
        parts, keep = self._particles(x, y, z, q)
        off, flat = self._molecules(molecules)
        _check(_load().aleator_gpu_ewald_state_resync(self._h, C.byref(parts), _p(off, _u64p), _p(flat, _u64p),
                                                      C.c_uint64(len(molecules))))

    def recompute_from_scratch(self, x, y, z, q, molecules) -> float:
        parts, keep = self._particles(x, y, z, q)
        off, flat = self._molecules(molecules)
        out = C.c_double()
        _check(_load().aleator_gpu_ewald_state_recompute(self._h, C.byref(parts), _p(off, _u64p), _p(flat, _u64p),
                                                         C.c_uint64(len(molecules)), C.byref(out)))
        return out.value

    def raw(self) -> np.ndarray:
        n = self.fw.num_k
        out = np.zeros(2 * n + 2)
        w = C.c_uint64()
        _check(_load().aleator_gpu_ewald_state_raw(self._h, _p(out), C.c_uint64(len(out)), C.byref(w)))
        return out

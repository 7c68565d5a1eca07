"""The reference's real-structure GCMC protocols on IRMOF-1 (tests only).

Inputs are the reference's own io::readCif output for its IRMOF-1 CIF files
(tests/golden/cif/expected_structures.npz), so these systems do not depend on the reader under
test.  Protocols:
  methane: test_gcmc_ch4_irmof1_isotherm.cc:86-170 (and the CLI run of gcmc_ch4_irmof1.toml,
           cli/main.cc:481-588): UFF "GenericMOFs" framework, CH4 158.5 K / 3.72 A, rc 12 A
           shifted, 298 K, Peng-Robinson methane fugacity, seed 2026, stream = point index,
           30000 equilibration + 90000 production steps, <N> over every production step.
  co2:     test_gcmc_co2_irmof1_isotherm.cc:160-334: DDEC charges shifted to neutral, 2x2x2
           supercell, CO2 sites O(-1.16) C O(+1.16), rc 12.8 A shifted, Ewald alpha 0.25 kmax 6,
           seed 2026, stream = point index, 2000 + 40000 steps in 8 blocks."""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

import oracle_lib as ol

SEED = 2026
AMU_KG = 1.66053906660e-27  # CODATA 2018 (test_gcmc_co2_irmof1_isotherm.cc:155)
AVOGADRO = 6.02214076e23
UFF = {"Zn": (62.3992, 2.46155), "O": (48.1581, 3.03315), "C": (47.8562, 3.47299), "H": (7.64893, 2.84642)}
METHANE_PR = (190.6, 4.599e6, 0.011)  # engines::kMethane
CO2_PR = (304.1282, 7377300.0, 0.22394)
METHANE_BAR = (0.1, 1.0, 5.0, 10.0)
CO2_PA = (1.0e4, 1.0e5)

# Numbers the reference records for these protocols
CLI_MEAN_N = "0.231456"            # test_cli_end_to_end.cc:170 (methane, 0.1 bar)
CLI_MMOL_G = "0.0375804"           # test_cli_end_to_end.cc:171
METHANE_MMOL_G = (0.0376, 0.3860, 1.9061, 3.7551)  # data/irmof1/known_deviation_baseline.md:19-24

HERE = os.path.dirname(os.path.abspath(__file__))
TRACES = os.path.join(HERE, "golden", "irmof1_gcmc_traces.npz")


def _seq_sum(a) -> float:
    # left-to-right accumulation, as the reference's range-for sums
    s = 0.0
    for v in np.asarray(a, dtype=np.float64):
        s += float(v)
    return s


def structure(stem: str) -> dict:
    z = np.load(os.path.join(HERE, "golden", "cif", "expected_structures.npz"))
    keys = ("lattice", "x", "y", "z", "mass", "charge", "species", "species_symbols")
    return {k: z[f"{stem}.{k}"] for k in keys}


@dataclass
class System:
    fw: ol.Framework
    lj: ol.LJ
    sites: list
    fugacity: float
    temperature: float
    stream: int
    ewald: tuple | None
    equilibration: int
    production: int
    framework_mass_amu: float

    @property
    def steps(self) -> int:
        return self.equilibration + self.production


def _lj(symbols, guests, cutoff):
    eps = [UFF[str(e)][0] for e in symbols] + [g[0] for g in guests]
    sig = [UFF[str(e)][1] for e in symbols] + [g[1] for g in guests]
    return ol.LJ(np.array(eps), np.array(sig), cutoff, ol.SHIFTED)


def methane(point: int) -> System:
    s = structure("IRMOF_1")
    fw = ol.Framework(m=s["lattice"], x=s["x"], y=s["y"], z=s["z"], q=s["charge"], species=s["species"])
    ch4 = len(s["species_symbols"])
    f = ol.pr_fugacity(*METHANE_PR, 298.0, METHANE_BAR[point] * 1.0e5)
    return System(fw, _lj(s["species_symbols"], [(158.5, 3.72)], 12.0), [(0.0, 0.0, 0.0, ch4, 0.0)], f, 298.0,
                  point, None, 30000, 90000, _seq_sum(s["mass"]))


def co2(point: int) -> System:
    s = structure("IRMOF_1_primitive_DDEC")
    q = np.asarray(s["charge"], dtype=np.float64)
    q = q - _seq_sum(q) / len(q)
    reps = (2, 2, 2)
    m = np.asarray(s["lattice"])
    # buildSupercell of the test (cells in i, j, k order, shift (a0 i + a1 j) + a2 k)
    xs, ys, zs = [], [], []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                sh = [(m[0][d] * i + m[1][d] * j) + m[2][d] * k for d in range(3)]
                xs.append(s["x"] + sh[0])
                ys.append(s["y"] + sh[1])
                zs.append(s["z"] + sh[2])
    n_cells = int(np.prod(reps))
    fw = ol.Framework(m=m * np.array(reps, dtype=np.float64)[:, None], x=np.concatenate(xs), y=np.concatenate(ys),
                      z=np.concatenate(zs), q=np.tile(q, n_cells), species=np.tile(s["species"], n_cells))
    nsp = len(s["species_symbols"])
    c_i, o_i = nsp, nsp + 1
    sites = [(0.0, 0.0, -1.16, o_i, -0.35), (0.0, 0.0, 0.0, c_i, 0.70), (0.0, 0.0, 1.16, o_i, -0.35)]
    f = ol.pr_fugacity(*CO2_PR, 298.0, CO2_PA[point])
    return System(fw, _lj(s["species_symbols"], [(27.0, 2.8), (79.0, 3.05)], 12.8), sites, f, 298.0, point,
                  (0.25, 12.8, 6), 2000, 40000, _seq_sum(np.tile(s["mass"], n_cells)))


def all_points():
    for p in range(len(METHANE_BAR)):
        yield f"methane_{p}", methane(p), 120000
    for p in range(len(CO2_PA)):
        yield f"co2_{p}", co2(p), 42000


def encode_trace(trace):
    trace = np.asarray(trace, dtype=np.int64)
    prev = np.concatenate([[0], trace[:-1]])
    at = np.nonzero(trace != prev)[0]
    return at.astype(np.int32), trace[at].astype(np.int16)


def decode_trace(steps, counts, n):
    out = np.zeros(n, dtype=np.int64)
    for s, c in zip(steps, counts):
        out[s:] = c
    return out


def golden_trace(name: str, n: int):
    z = np.load(TRACES)
    return decode_trace(z[f"{name}.steps"], z[f"{name}.counts"], n)


def production_mean(trace, system: System) -> float:
    # sum of counts after each production step / steps (cli/main.cc:548-563)
    return _seq_sum(np.asarray(trace[system.equilibration:system.steps], dtype=np.float64)) / system.production


def loading_mmol_per_g(mean_n: float, system: System) -> float:
    return (mean_n / AVOGADRO) / (system.framework_mass_amu * AMU_KG)

"""Synthetic benchmark workloads (no reference counterpart), built from the library's own
generator (aleator_synth_framework) and host utilities only.

  c0  cubic 25.83 A, 424 atoms, CH4 probe, rc 12.8 A, 0.2 A grid (129^3 by the reference rule)
  c1  rhombohedral 24 A 60/60/60, 600 charged atoms -> 2x2x2 supercell (4800), CO2 probe,
      Ewald alpha 0.265 kmax 8 rc 12 A, 0.15 A grid over the primitive cell (131^3)
  c2  256 frameworks: per framework f, Philox4x32Rng(seed 2, stream 2f) draws the cell (edges
      U(10, 40) A; orthorhombic when the fourth draw < 0.5, else angles U(60, 120) deg; the whole
      cell is redrawn until its metric is valid), then a number density U(0.03, 0.10) A^-3 sets
      N = clamp(round(rho V), 50, 2000); atoms from stream 2f+1; the minimum supercell for
      rc 12.8 A; CH4 probe; 0.2 A grid over the primitive cell
  c3  64 CO2 chains on the c1 framework at 298 K: 8 pressures 0.1 * 300^(i/7) bar x 8 seeds,
      fugacity from Peng-Robinson (CO2: Tc 304.1282 K, Pc 7377300 Pa, omega 0.22394)
"""
from __future__ import annotations

import numpy as np

from . import grid_point_counts, peng_robinson_fugacity, philox_uniforms, replicate_supercell, \
    supercell_replication, synth_framework

C0_TYPES = ((52.8, 3.43), (30.2, 3.12), (22.1, 2.57), (62.4, 2.46))
METHANE = (148.0, 3.73)
CO2_C = (27.0, 2.80)
CO2_O = (79.0, 3.05)
CO2_SITES = [(0.0, 0.0, -1.16, 5, -0.35), (0.0, 0.0, 0.0, 4, 0.70), (0.0, 0.0, 1.16, 5, -0.35)]
CH4_SITES = [(0.0, 0.0, 0.0, 4, 0.0)]
CO2_PR = (304.1282, 7377300.0, 0.22394)


def cell_from_params(a, b, c, alpha, beta, gamma):
    """Row-vector lattice from edges (A) and angles (deg); None if the metric is invalid."""
    al, be, ga = np.radians([alpha, beta, gamma])
    ca, cb, cg, sg = np.cos(al), np.cos(be), np.cos(ga), np.sin(ga)
    cy = (ca - cb * cg) / sg
    cz2 = 1.0 - cb * cb - cy * cy
    if not cz2 > 1e-6:
        return None
    return np.array([[a, 0, 0], [b * cg, b * sg, 0], [c * cb, c * cy, c * np.sqrt(cz2)]], dtype=np.float64)


def c0():
    m = np.diag([25.83] * 3)
    x, y, z, q, sp = synth_framework(m, 424, seed=1, stream=0)
    eps = np.array([t[0] for t in C0_TYPES] + [METHANE[0]])
    sig = np.array([t[1] for t in C0_TYPES] + [METHANE[1]])
    return dict(m=m, x=x, y=y, z=z, q=q, sp=sp, eps=eps, sig=sig, rc=12.8, counts=grid_point_counts(m, 0.2))


def c1():
    m = cell_from_params(24.0, 24.0, 24.0, 60.0, 60.0, 60.0)
    x, y, z, q, sp = synth_framework(m, 600, seed=1, stream=0, charge_sigma=0.5, charge_stream=2)
    eps = np.array([t[0] for t in C0_TYPES] + [CO2_C[0], CO2_O[0]])
    sig = np.array([t[1] for t in C0_TYPES] + [CO2_C[1], CO2_O[1]])
    reps = supercell_replication(m, 12.0)
    sm, sx, sy, sz, sq, ss = replicate_supercell(m, x, y, z, q, sp, reps)
    return dict(prim=m, sup=sm, x=sx, y=sy, z=sz, q=sq, sp=ss, eps=eps, sig=sig, rc=12.0, alpha=0.265, kmax=8,
                counts=grid_point_counts(m, 0.15), reps=reps)


def c2_framework(f: int, rc=12.8, spacing=0.2):
    seed, stream = 2, 2 * f
    call = 0

    def u():
        nonlocal call
        v = philox_uniforms(seed, stream, call, 1)[0]
        call += 1
        return v

    while True:
        a, b, c = (10.0 + 30.0 * u() for _ in range(3))
        if u() < 0.5:
            ang = (90.0, 90.0, 90.0)
        else:
            ang = tuple(60.0 + 60.0 * u() for _ in range(3))
        m = cell_from_params(a, b, c, *ang)
        if m is not None:
            break
    rho = 0.03 + 0.07 * u()
    vol = abs(np.linalg.det(m))
    n = int(min(max(round(rho * vol), 50), 2000))
    x, y, z, q, sp = synth_framework(m, n, seed=seed, stream=stream + 1)
    reps = supercell_replication(m, rc)
    sm, sx, sy, sz, sq, ss = replicate_supercell(m, x, y, z, q, sp, reps)
    return dict(prim=m, sup=sm, x=sx, y=sy, z=sz, q=sq, sp=ss, n_atoms=n, reps=reps,
                counts=grid_point_counts(m, spacing), rc=rc)


def c2(n_frameworks=256):
    eps = np.array([t[0] for t in C0_TYPES] + [METHANE[0]])
    sig = np.array([t[1] for t in C0_TYPES] + [METHANE[1]])
    return [c2_framework(f) for f in range(n_frameworks)], eps, sig


def c3_chains(temperature=298.0):
    chains = []
    for i in range(8):
        p_pa = 0.1 * 300.0 ** (i / 7.0) * 1e5
        f = peng_robinson_fugacity(*CO2_PR, temperature, p_pa)
        for s in range(8):
            chains.append((f, temperature, 1000 + s, 8 * i + s))
    return chains

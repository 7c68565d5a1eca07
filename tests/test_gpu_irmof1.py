"""Device GCMC chains on the reference's real IRMOF-1 protocols, built end to end through the
product path (CIF reader -> supercell -> device framework -> batched chains), follow the
reference's trajectories step for step: the molecule count after every step equals the oracle
fixture (tests/golden/irmof1_gcmc_traces.npz, itself pinned to the reference's recorded numbers
by test_oracle_irmof1.py), so the device reproduces the reference CLI's <N> = 0.231456 and
0.0375804 mmol/g and its methane isotherm baseline exactly."""
import os

import numpy as np
import pytest

import aleator_amd as aa
import irmof1_systems as sy

pytestmark = pytest.mark.gpu
CIF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cif")


def product_framework(ctx, system, cif, reps=None, neutralize=False):
    s = aa.read_cif(os.path.join(CIF, cif))
    q = s.charge
    if neutralize:
        tot = 0.0
        for v in q:
            tot += float(v)
        q = q - tot / len(q)
    lat, x, y, z, species = s.lattice, s.x, s.y, s.z, s.species
    if reps is not None:
        lat, x, y, z, q, species = aa.replicate_supercell(lat, x, y, z, q, species, reps)
    for a, b in ((lat, system.fw.m), (x, system.fw.x), (y, system.fw.y), (z, system.fw.z), (q, system.fw.q),
                 (species, system.fw.species)):
        assert np.array_equal(a, b)
    lj = system.lj
    return aa.Framework(ctx, lat, x, y, z, species, lj.eps, lj.sigma, lj.cutoff, truncation=lj.truncation,
                        mixing=lj.mixing, charge=q, ewald=system.ewald)


def test_methane_isotherm_chains_follow_reference(ctx):
    print("irmof1 methane: 4 chains x 120000 steps", flush=True)
    systems = [sy.methane(p) for p in range(4)]
    for s in systems:
        assert aa.peng_robinson_fugacity(*sy.METHANE_PR, 298.0, sy.METHANE_BAR[s.stream] * 1e5) == s.fugacity
    fw = product_framework(ctx, systems[0], "IRMOF-1.cif")
    gc = aa.Gcmc(fw, systems[0].sites, [(s.fugacity, s.temperature, sy.SEED, s.stream) for s in systems],
                 max_molecules=256)
    trace = gc.run(120000, trace=True)
    for c, s in enumerate(systems):
        gold = sy.golden_trace(f"methane_{c}", s.steps)
        diverge = np.nonzero(trace[c] != gold)[0]
        assert len(diverge) == 0, f"chain {c} diverged at step {diverge[0]}"
        loading = sy.loading_mmol_per_g(sy.production_mean(trace[c], s), s)
        print(f"  {sy.METHANE_BAR[c]} bar: <N> {sy.production_mean(trace[c], s):.6g}, {loading:.6g} mmol/g")
        assert round(loading, 4) == sy.METHANE_MMOL_G[c]
    mean0 = sy.production_mean(trace[0], systems[0])
    assert f"{mean0:.6g}" == sy.CLI_MEAN_N
    assert f"{sy.loading_mmol_per_g(mean0, systems[0]):.6g}" == sy.CLI_MMOL_G


def test_co2_ddec_chains_follow_reference(ctx):
    print("irmof1 co2 (DDEC, 2x2x2, Ewald): 2 chains x 42000 steps", flush=True)
    systems = [sy.co2(p) for p in range(2)]
    fw = product_framework(ctx, systems[0], "IRMOF-1_primitive_DDEC.cif", reps=(2, 2, 2), neutralize=True)
    gc = aa.Gcmc(fw, systems[0].sites, [(s.fugacity, s.temperature, sy.SEED, s.stream) for s in systems],
                 max_molecules=256)
    trace = gc.run(42000, trace=True)
    for c, s in enumerate(systems):
        gold = sy.golden_trace(f"co2_{c}", s.steps)
        diverge = np.nonzero(trace[c] != gold)[0]
        assert len(diverge) == 0, f"chain {c} diverged at step {diverge[0]}"
        blocks = trace[c][s.equilibration:s.steps].reshape(8, -1).mean(axis=1)
        print(f"  {sy.CO2_PA[c]:.0e} Pa: {sy.loading_mmol_per_g(blocks.mean(), s):.4f} mol/kg")

"""The oracle's GCMC on the reference's real IRMOF-1 protocols reproduces the numbers the
reference itself records, and the trajectory fixtures the device tests compare against.

  methane, 0.1 bar: <N> = 0.231456 and 0.0375804 mmol/g, printed by the reference CLI
                    (test_cli_end_to_end.cc:170-171);
  methane, 0.1 / 1 / 5 / 10 bar: 0.0376 / 0.3860 / 1.9061 / 3.7551 mmol/g
                    (data/irmof1/known_deviation_baseline.md:19-24).
The CO2/DDEC fixture is checked here over a prefix (the whole run takes minutes on the CPU) and
against the compiled reference engine where oracle/_ref is built."""
import numpy as np
import pytest

import irmof1_systems as sy
import oracle_lib as ol

need_ref = pytest.mark.skipif(not ol.have_ref(), reason="oracle/_ref not built (needs the reference sources)")


def run(system, steps):
    g = ol.Gcmc(system.lj, system.fw, system.sites, system.fugacity, system.temperature, sy.SEED, system.stream,
                ewald=system.ewald)
    return g, g.run(steps, trace=True).astype(np.int64)


def test_methane_cli_point_reproduces_reference_output():
    system = sy.methane(0)
    _, trace = run(system, system.steps)
    mean_n = sy.production_mean(trace, system)
    assert f"{mean_n:.6g}" == sy.CLI_MEAN_N
    assert f"{sy.loading_mmol_per_g(mean_n, system):.6g}" == sy.CLI_MMOL_G
    assert np.array_equal(trace, sy.golden_trace("methane_0", system.steps))


@pytest.mark.parametrize("point", [1, 2, 3])
def test_methane_isotherm_points_reproduce_reference_baseline(point):
    system = sy.methane(point)
    _, trace = run(system, system.steps)
    loading = sy.loading_mmol_per_g(sy.production_mean(trace, system), system)
    assert round(loading, 4) == sy.METHANE_MMOL_G[point]
    assert np.array_equal(trace, sy.golden_trace(f"methane_{point}", system.steps))


@pytest.mark.parametrize("point", [0, 1])
def test_co2_ddec_fixture_prefix(point):
    system = sy.co2(point)
    assert system.fw.n == 848 and abs(float(np.sum(system.fw.q))) < 1e-9
    _, trace = run(system, 1500)
    assert np.array_equal(trace, sy.golden_trace(f"co2_{point}", 1500))


@need_ref
def test_methane_cli_point_matches_reference_engine():
    """The compiled reference engine ends the 120000-step 0.1 bar run in the oracle's state."""
    system = sy.methane(0)
    g, _ = run(system, system.steps)
    n_ref, parts, _ = ol.ref_gcmc_run(system.lj, system.fw, system.sites, system.fugacity, system.temperature,
                                      sy.SEED, system.stream, system.steps)
    assert n_ref == g.n_molecules
    for a, b in zip(g.particles(), parts):
        assert np.array_equal(a, b)

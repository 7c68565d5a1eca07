"""Writes tests/golden/irmof1_gcmc_traces.npz: molecule-count trajectories of the reference's GCMC
protocols on the real IRMOF-1 structures, computed by the oracle's FP64 restatement of
MonteCarloEngine::run (oracle/gcmc.c), whose trajectories are bitwise equal to the compiled
reference engine's.

Systems (inputs are the reference's own readCif output, tests/golden/cif/expected_structures.npz):
  methane  IRMOF-1.cif, UFF framework + CH4 (158.5 K, 3.72 A), rc 12 A shifted, 298 K, Peng-Robinson
           fugacity at 0.1, 1, 5, 10 bar; seed 2026, stream = point index; 30000 + 90000 steps
           (test_gcmc_ch4_irmof1_isotherm.cc:86-170 and the CLI example gcmc_ch4_irmof1.toml)
  co2      IRMOF-1_primitive_DDEC.cif, charges shifted to neutral, 2x2x2 supercell, UFF + CO2
           (C 27.0/2.8, O 79.0/3.05, q +0.70/-0.35, O-C-O 1.16 A), rc 12.8 A shifted, Ewald
           alpha 0.25 kmax 6, 298 K, 1e4 and 1e5 Pa; seed 2026, stream = point index; 2000 + 40000
           steps (test_gcmc_co2_irmof1_isotherm.cc:160-334)
Each trace is stored as the steps at which the count changes and the count after that step
(<name>.steps int32, <name>.counts int16), from an empty start.

    python tests/golden/make_irmof1_gcmc_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import irmof1_systems as sy  # noqa: E402
import oracle_lib as ol  # noqa: E402

OUT = os.path.join(HERE, "irmof1_gcmc_traces.npz")


def main():
    arrays = {}
    for name, system, steps in sy.all_points():
        g = ol.Gcmc(system.lj, system.fw, system.sites, system.fugacity, system.temperature, sy.SEED, system.stream,
                    ewald=system.ewald)
        trace = g.run(steps, trace=True).astype(np.int64)
        chg, cnt = sy.encode_trace(trace)
        arrays[f"{name}.steps"], arrays[f"{name}.counts"] = chg, cnt
        print(f"{name}: {len(chg)} changes, production mean {sy.production_mean(trace, system):.6g}", flush=True)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

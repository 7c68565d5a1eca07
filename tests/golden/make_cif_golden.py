"""Writes tests/golden/cif/expected_structures.npz: the reference's io::readCif output (compiled
from the reference sources into oracle/_ref by `make -C oracle ref`) for every CIF file under
tests/golden/cif/.  Per file <stem>: lattice (3x3), x, y, z, mass, charge, occupancy (float64),
species (uint32), species_symbols and source_labels (unicode arrays).

    python tests/golden/make_cif_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

CIF_DIR = os.path.join(HERE, "cif")
OUT = os.path.join(CIF_DIR, "expected_structures.npz")


def main():
    arrays = {}
    for name in sorted(os.listdir(CIF_DIR)):
        if not name.endswith(".cif"):
            continue
        stem = name[:-4].replace("-", "_")
        r = ol.ref_read_cif(os.path.join(CIF_DIR, name))
        assert r is not None, name
        for k, v in r.items():
            arrays[f"{stem}.{k}"] = np.asarray(v) if not isinstance(v, list) else np.array(v, dtype=str)
        print(f"{name}: {len(r['x'])} atoms, species {r['species_symbols']}")
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

"""The oracle's Coulomb grid potential (real space as ewald.cc:223-249, reciprocal as the linear
term of ewald_incremental_state.cc:117-121) against an independent extended-precision evaluation,
at c1 nodes where the potential is a near-total cancellation of its two Ewald parts.  This pins
the FP64 rounding floor that the device's FP64 Coulomb bar is written against: the oracle is
within 1e-9 K of the exact value there, and the two parts are each right to their own rounding."""
import numpy as np

import oracle_lib as ol
from precise_coulomb import PreciseCoulomb

# c1 nodes whose potential is 1-3 K while each Ewald part is 2e3-7e4 K (found on the device grid)
HARD_NODES = [777860, 334564, 13225, 1341924, 189603, 992227]


def test_oracle_coulomb_at_cancellation_nodes():
    prim, _ = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    counts = ol.grid_counts(prim.m, 0.15)
    pc = PreciseCoulomb(sc, 0.265, 8, 12.0)
    st = ol.EwaldState(sc, 0.265, 8)
    nodes = np.array(HARD_NODES, dtype=np.int64)
    real, rec = ol.coulomb_nodes(st, 12.0, sc, prim.m, counts, nodes)
    for i, node in enumerate(HARD_NODES):
        er, ek = pc.at(pc.node_position(prim.m, counts, node))
        exact = float(er + ek)
        assert 0.5 < abs(exact) < 5.0 and abs(float(er)) > 1e3  # a cancellation node
        print(f"node {node}: phi {exact:.6f} K, oracle - exact {real[i] + rec[i] - exact:+.2e} K")
        assert abs(real[i] - float(er)) <= 1e-9 * max(1.0, abs(float(er))) * 10
        assert abs(rec[i] - float(ek)) <= 1e-9 * max(1.0, abs(float(ek))) * 10
        assert abs(real[i] + rec[i] - exact) <= 1e-9


def test_precise_coulomb_planted_mistake():
    """A wrong reciprocal prefactor (2 pi/V instead of 4 pi/V) moves the potential by thousands of
    kelvin at these nodes: the check above would catch it."""
    prim, _ = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    counts = ol.grid_counts(prim.m, 0.15)
    pc = PreciseCoulomb(sc, 0.265, 8, 12.0)
    er, ek = pc.at(pc.node_position(prim.m, counts, HARD_NODES[0]))
    assert abs(float(er + ek / 2)) > 100.0

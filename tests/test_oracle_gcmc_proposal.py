"""Pins of the oracle's biased-insertion GCMC (insertion anchors drawn from a piecewise-constant
density, acceptance corrected by V p(r)), which the device chains are checked against.

Detailed balance is what the correction must preserve, so the pins are statistical laws of an
ideal gas (all LJ well depths zero, no charges): the molecule count is Poisson with mean
f V / T, and molecules fill the cell uniformly, however strongly the proposal favours one
region.  Dropping or inverting either correction factor moves the occupancy of the favoured
bin from 1/8 towards 50/57."""
import numpy as np
import pytest

import oracle_lib as ol

K_B = 1.380649e-23


def ideal_gas(mean_n=4.0, a=20.0):
    m = np.diag([a, a, a])
    fw = ol.Framework(m, np.array([1.0]), np.array([1.0]), np.array([1.0]), np.zeros(1),
                      np.zeros(1, dtype=np.uint32))
    lj = ol.LJ(np.array([0.0, 0.0]), np.array([1.0, 1.0]), 5.0)
    t = 300.0
    f_pa = mean_n * t / a ** 3 * (K_B * 1e30)
    return fw, lj, f_pa, t


def occupancy(g, steps, every, nb=2):
    """Per-bin fraction of molecules (anchor bins of the 2x2x2 grid), sampled every `every` steps,
    and the mean molecule count over all steps."""
    counts = np.zeros(nb ** 3)
    ns = []
    for _ in range(steps // every):
        tr = g.run(every, trace=True)
        ns.append(tr.mean())
        x, y, z, _ = g.particles()
        sites = g.molecule_sites()
        if len(sites):
            f = np.stack([x[sites[:, 0]], y[sites[:, 0]], z[sites[:, 0]]], axis=1) / 20.0
            f -= np.floor(f)
            b = np.minimum((f * nb).astype(int), nb - 1)
            np.add.at(counts, (b[:, 0] * nb + b[:, 1]) * nb + b[:, 2], 1)
    return counts / counts.sum(), float(np.mean(ns))


def test_proposal_tables():
    fw, lj, f_pa, t = ideal_gas()
    g = ol.Gcmc(lj, fw, [(0.0, 0.0, 0.0, 1, 0.0)], f_pa, t, 1, 0)
    g.set_proposal(np.array([3.0, 1.0]), (2, 1, 1), (1, 1, 1), 0.5)
    # p = 0.5 * (3/4, 1/4) + 0.5 / 2 = (0.625, 0.375); factor n p = (1.25, 0.75)
    assert g.g.prop_cdf[0] == 0.625 and g.g.prop_cdf[1] == 1.0
    assert g.g.prop_fac[0] == 1.25 and g.g.prop_fac[1] == 0.75
    with pytest.raises(ValueError):
        g.set_proposal(np.array([1.0, 0.0]), (2, 1, 1), (1, 1, 1), 0.0)  # a bin that can never be drawn
    with pytest.raises(ValueError):
        g.set_proposal(np.array([1.0, -1.0]), (2, 1, 1), (1, 1, 1), 0.5)
    g.set_proposal(None, (1, 1, 1), (1, 1, 1), 0.0)
    assert not g.g.prop_cdf


@pytest.mark.parametrize("reps", [(1, 1, 1), (2, 1, 2)])
def test_biased_insertion_keeps_the_ideal_gas_uniform(reps):
    fw, lj, f_pa, t = ideal_gas()
    g = ol.Gcmc(lj, fw, [(0.0, 0.0, 0.0, 1, 0.0)], f_pa, t, 3, 1)
    nb = 2
    # over the whole cell the favoured bin is bin 0 of every image; the check bins the cell
    # itself 2x2x2, which coincides with the density's bins only for reps = 1
    w = np.ones(nb ** 3)
    w[0] = 50.0
    g.set_proposal(w, (nb, nb, nb), reps, 0.0)
    occ, mean_n = occupancy(g, 200000, 50)
    assert mean_n == pytest.approx(4.0, abs=0.15)
    if reps == (1, 1, 1):
        assert np.all(np.abs(occ - 1 / 8) < 0.02), occ
    # the biased sampler really is biased: most insertion attempts land in the favoured bins
    att, acc = g.counts
    assert att[2] > 40000 and acc[2] > 0


def test_uniform_insertion_ideal_gas_reference_point():
    """The same statistics without a proposal (the reference's uniform insertion)."""
    fw, lj, f_pa, t = ideal_gas()
    g = ol.Gcmc(lj, fw, [(0.0, 0.0, 0.0, 1, 0.0)], f_pa, t, 3, 1)
    occ, mean_n = occupancy(g, 100000, 50)
    assert mean_n == pytest.approx(4.0, abs=0.2)
    assert np.all(np.abs(occ - 1 / 8) < 0.03), occ


def test_uncorrected_biased_insertion_is_caught():
    """A planted mistake (the oracle's correction factors overwritten with 1: biased proposals,
    plain acceptance) skews the occupancy far outside the tolerance the tests above use."""
    fw, lj, f_pa, t = ideal_gas()
    g = ol.Gcmc(lj, fw, [(0.0, 0.0, 0.0, 1, 0.0)], f_pa, t, 3, 1)
    w = np.ones(8)
    w[0] = 50.0
    g.set_proposal(w, (2, 2, 2), (1, 1, 1), 0.0)
    for b in range(8):
        g.g.prop_fac[b] = 1.0
    occ, _ = occupancy(g, 50000, 50)
    assert occ[0] > 0.5, occ

"""The analytic LJ tail correction on device outputs (c0's "truncated LJ at 12.8 A with analytic
tail correction"): with aleator_gpu_framework_set_lj_tail on, every c0 CH4 grid node and every
Widom trial carries exactly the oracle's U_LRC(framework + CH4) - U_LRC(framework)
(tailEnergyCorrection, lennard_jones.cc:218-257, restated in oracle.c), and K_H scales by
exp(-delta/T).  Charged CO2 on a small cell: the molecule's three sites enter together."""
import ctypes as C

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework

pytestmark = pytest.mark.gpu


def oracle_tail_delta(lj, fw, site_species):
    l, keep = ol.lj_struct(lj)
    lat = ol.lattice(fw.m)
    p0, k0 = ol.particles(fw)
    n = len(site_species)
    grown = ol.Framework(fw.m, np.append(fw.x, np.zeros(n)), np.append(fw.y, np.zeros(n)),
                         np.append(fw.z, np.zeros(n)), np.append(fw.q, np.zeros(n)),
                         np.append(fw.species, np.asarray(site_species, dtype=np.uint32)))
    p1, k1 = ol.particles(grown)
    return (ol.oracle().orc_lj_tail_energy(C.byref(l), C.byref(p1), C.byref(lat)) -
            ol.oracle().orc_lj_tail_energy(C.byref(l), C.byref(p0), C.byref(lat)))


def test_tail_on_c0_grid_and_widom(ctx):
    fw, lj = ol.c0_system()
    delta = oracle_tail_delta(lj, fw, [4])
    assert delta < 0.0
    dev = device_framework(ctx, fw, lj)
    counts = ol.grid_counts(fw.m, 0.2)
    for prec in (aa.FP64, aa.MIXED):
        plain = dev.grid([4], counts, precision=prec)[0].ravel()
        dev.set_lj_tail(True)
        tail = dev.grid([4], counts, precision=prec)[0].ravel()
        dev.set_lj_tail(False)
        free = plain < 1e5  # uncapped without the tail, so (delta < 0) uncapped with it
        d = tail[free] - plain[free]
        print(f"c0 grid precision {prec}: delta {delta:.9f} K, tail - plain in [{d.min():.9f}, {d.max():.9f}]")
        assert np.allclose(d, delta, rtol=0, atol=1e-9 * max(1.0, np.abs(plain[free]).max()))
        assert np.all((tail[~free] <= 1e5) & (tail[~free] >= 1e5 + delta))
    s0, du0 = dev.widom([(0.0, 0.0, 0.0, 4, 0.0)], 5000, 298.0, seed=1, stream=1, precision=aa.FP64)
    dev.set_lj_tail(True)
    s1, du1 = dev.widom([(0.0, 0.0, 0.0, 4, 0.0)], 5000, 298.0, seed=1, stream=1, precision=aa.FP64)
    dev.set_lj_tail(False)
    assert np.allclose(du1 - du0, delta, rtol=0, atol=1e-9 * np.maximum(1.0, np.abs(du0)).max())
    assert abs(s1 / s0 - np.exp(-delta / 298.0)) <= 1e-10


def test_tail_of_a_charged_co2_insertion(ctx):
    m = ol.cell_from_params(20.0, 21.0, 22.0, 70.0, 80.0, 95.0)
    fw = ol.synth_framework(m, 250, seed=7, stream=0, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 8.0)
    delta = oracle_tail_delta(lj, fw, [5, 4, 5])
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 8.0, 6))
    _, du0 = dev.widom(ol.co2_sites(), 300, 298.0, seed=2, stream=3, precision=aa.FP64)
    dev.set_lj_tail(True)
    _, du1 = dev.widom(ol.co2_sites(), 300, 298.0, seed=2, stream=3, precision=aa.FP64)
    print(f"CO2 tail {delta:.9f} K")
    assert np.allclose(du1 - du0, delta, rtol=0, atol=1e-9 * np.maximum(1.0, np.abs(du0)))

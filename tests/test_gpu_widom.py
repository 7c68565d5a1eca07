"""Parity of batched Widom insertion on the GPU with the FP64 oracle, trial by trial
(MonteCarloEngine::widomInsertionHenryCoefficient, monte_carlo_engine.cc:850-917).

The per-trial dU comparison pins the device Philox draw indices (a wrong index moves the
trial and changes dU by orders of magnitude), the rotation and every energy term.
Tolerances: FP64 1e-10 relative (|dU| floored at 1 K); MIXED 1e-5 relative on every trial with
|dU| > 1 K, against the oracle (and against the device FP64 run on large batches)."""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework, rel_err

pytestmark = pytest.mark.gpu

CH4 = [(0.0, 0.0, 0.0, 4, 0.0)]


def test_widom_c0_methane_fp64(ctx):
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    n = 3000
    kh_ref, du_ref = ol.widom(lj, fw, CH4, seed=1, stream=1, n_trials=n, temperature=298.0)
    s, du = dev.widom(CH4, n, 298.0, seed=1, stream=1, precision=aa.FP64)
    err = rel_err(du, du_ref)
    kh = s / n / 298.0
    print(f"c0 widom: max rel dU err {err.max():.3e}; K_H gpu {kh:.12e} oracle {kh_ref:.12e}")
    assert err.max() <= 1e-10
    assert abs(kh - kh_ref) <= 1e-10 * abs(kh_ref)


def test_widom_first_call_index_offsets_the_stream(ctx):
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    _, du_all = ol.widom(lj, fw, CH4, seed=9, stream=4, n_trials=40, temperature=298.0)
    _, du = dev.widom(CH4, 25, 298.0, seed=9, stream=4, first_call_index=15 * 3)
    assert rel_err(du, du_all[15:]).max() <= 1e-10


def test_widom_charged_co2_triclinic_fp64(ctx):
    m = ol.cell_from_params(20.0, 21.0, 22.0, 70.0, 80.0, 95.0)
    fw = ol.synth_framework(m, 250, seed=7, stream=0, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 8.0)
    st = ol.EwaldState(fw, 0.35, 6)
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 8.0, 6))
    n = 150
    kh_ref, du_ref = ol.widom(lj, fw, ol.co2_sites(), seed=2, stream=3, n_trials=n, temperature=298.0,
                              ewald=st, rc_real=8.0)
    s, du = dev.widom(ol.co2_sites(), n, 298.0, seed=2, stream=3, precision=aa.FP64)
    err = rel_err(du, du_ref)
    print(f"charged CO2 widom: max rel dU err {err.max():.3e}; dU range {du_ref.min():.3g}..{du_ref.max():.3g}")
    assert err.max() <= 1e-10
    kh = s / n / 298.0
    assert abs(kh - kh_ref) <= 1e-9 * abs(kh_ref)


def mixed_bar(name, du, du_ref):
    sel = np.abs(du_ref) > 1.0
    err = np.abs(du - du_ref)[sel] / np.abs(du_ref[sel])
    print(f"{name} MIXED: {sel.sum()} trials with |dU| > 1 K, max rel {err.max():.3e}, median {np.median(err):.3e}")
    assert err.max() <= 1e-5
    return float(err.max())


def test_widom_mixed_precision_c0(ctx):
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    n = 3000
    kh_ref, du_ref = ol.widom(lj, fw, CH4, seed=1, stream=1, n_trials=n, temperature=298.0)
    s, du = dev.widom(CH4, n, 298.0, seed=1, stream=1, precision=aa.MIXED)
    mixed_bar("c0 CH4 widom", du, du_ref)
    assert abs(s / n / 298.0 - kh_ref) <= 1e-5 * abs(kh_ref)


def test_widom_mixed_precision_charged_c1(ctx):
    """Charged CO2 on the c1 supercell (4800 atoms, alpha 0.265, kmax 8 = 2204 k) in MIXED against
    the oracle: 300 trials on the direct (split) kernel and the first 60 of a 1e5-trial batch,
    which takes the cube kernel; K_H of the 300 within 1e-5."""
    print("widom c1 CO2 MIXED: 300 trials and 1e5 trials against the oracle", flush=True)
    fw, lj = ol.c1_system()
    sc = ol.replicate(fw, ol.min_supercell(fw.m, 12.0))
    st = ol.EwaldState(sc, 0.265, 8)
    dev = device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))
    n = 300
    kh_ref, du_ref = ol.widom(lj, sc, ol.co2_sites(), seed=6, stream=2, n_trials=n, temperature=298.0, ewald=st,
                              rc_real=12.0)
    s, du = dev.widom(ol.co2_sites(), n, 298.0, seed=6, stream=2, precision=aa.MIXED)
    mixed_bar("c1 CO2 widom (direct)", du, du_ref)
    assert abs(s / n / 298.0 - kh_ref) <= 1e-5 * abs(kh_ref)
    _, big = dev.widom(ol.co2_sites(), 100_000, 298.0, seed=6, stream=2, precision=aa.MIXED)
    mixed_bar("c1 CO2 widom (cube)", big[:60], du_ref[:60])
    _, big64 = dev.widom(ol.co2_sites(), 100_000, 298.0, seed=6, stream=2, precision=aa.FP64)
    assert rel_err(big64[:60], du_ref[:60]).max() <= 1e-10
    mixed_bar("c1 CO2 widom (cube, all 1e5 vs device FP64)", big, big64)


def test_widom_rejects_bad_arguments(ctx):
    fw, lj = ol.c0_system(60, a=12.0)
    lj.cutoff = 5.9
    dev = device_framework(ctx, fw, lj)
    with pytest.raises(ValueError):
        dev.widom(CH4, 10, -1.0)
    with pytest.raises(IndexError):
        dev.widom([(0, 0, 0, 99, 0.0)], 10, 298.0)


def test_widom_cell_path_matches_the_oracle(ctx):
    """Large batches group trials by cube and sum against per-cube LDS lists; small ones scan
    every atom.  On c0 a 1e7-trial batch takes the first path and a 200-trial batch the second:
    both give the oracle's dU for the first 200 trials, and MIXED meets its bar against FP64 over
    all 1e7."""
    print("widom c0: 200 and 1e7 trials, FP64 and MIXED", flush=True)
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    _, du_ref = ol.widom(lj, fw, CH4, seed=3, stream=2, n_trials=200, temperature=298.0)
    _, small = dev.widom(CH4, 200, 298.0, seed=3, stream=2, precision=aa.FP64)
    _, big = dev.widom(CH4, 10_000_000, 298.0, seed=3, stream=2, precision=aa.FP64)
    assert rel_err(small, du_ref).max() <= 1e-10
    assert rel_err(big[:200], du_ref).max() <= 1e-10
    _, big32 = dev.widom(CH4, 10_000_000, 298.0, seed=3, stream=2, precision=aa.MIXED)
    mixed_bar("c0 CH4 widom 1e7 (cube, vs device FP64)", big32, big)
    mixed_bar("c0 CH4 widom 1e7 (cube, first 200 vs oracle)", big32[:200], du_ref)


def test_widom_cube_sort_paths_agree_bitwise(ctx):
    """The trials are grouped by cube by per-block LDS histograms when the cube grid has at most
    12288 cubes (1e7 trials on c0: 22^3) and by global atomics above that (2e7 trials: 27^3).
    A trial's dU is its atom-order sum over the images within the cutoff whatever cube holds it;
    only the rounding of the cube-relative coordinates differs (amplified near the r^-12 wall),
    so the first 1e7 trials of both batches agree within the FP64 parity bar (a misplaced trial
    would differ by orders of magnitude), and both equal the oracle's first trials."""
    print("widom c0: 1e7 and 2e7 trials, FP64", flush=True)
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    _, a = dev.widom(CH4, 10_000_000, 298.0, seed=5, stream=1, precision=aa.FP64)
    _, b = dev.widom(CH4, 20_000_000, 298.0, seed=5, stream=1, precision=aa.FP64)
    err = rel_err(b[:10_000_000], a)
    print(f"  1e7 vs 2e7 batches, first 1e7 trials: max rel {err.max():.3e}, bitwise equal {np.mean(err == 0):.4f}")
    assert err.max() <= 1e-10
    _, du_ref = ol.widom(lj, fw, CH4, seed=5, stream=1, n_trials=100, temperature=298.0)
    assert rel_err(a[:100], du_ref).max() <= 1e-10
    assert rel_err(b[:100], du_ref).max() <= 1e-10


def test_widom_cell_path_charged_c1(ctx):
    """Charged CO2 on the c1 supercell (4800 atoms, Ewald alpha 0.265, kmax 8): a 1e5-trial batch
    runs the cube path; its first 40 trials equal the oracle's dU."""
    print("widom c1 CO2: 1e5 trials against the oracle's first 40", flush=True)
    fw, lj = ol.c1_system()
    sc = ol.replicate(fw, ol.min_supercell(fw.m, 12.0))
    st = ol.EwaldState(sc, 0.265, 8)
    dev = device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))
    _, du_ref = ol.widom(lj, sc, ol.co2_sites(), seed=4, stream=0, n_trials=40, temperature=298.0, ewald=st,
                         rc_real=12.0)
    _, du = dev.widom(ol.co2_sites(), 100_000, 298.0, seed=4, stream=0, precision=aa.FP64)
    assert rel_err(du[:40], du_ref).max() <= 1e-10


def test_widom_all_atom_path_for_large_molecules(ctx):
    """A molecule wider than the cell leaves no cell grid whose list radius fits inside the cell,
    so the all-atom kernel runs; it matches the oracle too."""
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    wide = [(0.0, 0.0, -13.5, 4, 0.0), (0.0, 0.0, 13.5, 4, 0.0)]  # radius 13.5 > L - rc = 13.03
    _, du_ref = ol.widom(lj, fw, wide, seed=5, stream=1, n_trials=100, temperature=298.0)
    _, du = dev.widom(wide, 100, 298.0, seed=5, stream=1, precision=aa.FP64)
    assert rel_err(du, du_ref).max() <= 1e-10


def test_widom_split_kernel_c0_1e5(ctx):
    """c0's configured batch, 1e5 CH4 trials: the cube grid is rejected (the list sphere would
    cover more than 70 % of the cell), so the split kernel runs (two trials per lane, each trial's
    atoms split over 8 waves, fractional units in the cubic cell); its first 300 trials equal the
    oracle's in FP64 and meet the MIXED bar, and K_H equals the oracle's over those 300."""
    print("widom c0 1e5 trials, split kernel", flush=True)
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    kh_ref, du_ref = ol.widom(lj, fw, CH4, seed=1, stream=1, n_trials=300, temperature=298.0)
    s300, _ = dev.widom(CH4, 300, 298.0, seed=1, stream=1, precision=aa.FP64)
    assert abs(s300 / 300 / 298.0 - kh_ref) <= 1e-10 * abs(kh_ref)
    s, du = dev.widom(CH4, 100_000, 298.0, seed=1, stream=1, precision=aa.FP64)
    assert rel_err(du[:300], du_ref).max() <= 1e-10
    # the batch sum, against the per-trial energies (K_H =
    # <exp(-dU/T)> / T, monte_carlo_engine.cc:880-916), and again on a second call
    s_host = np.sum(np.exp(-du / 298.0))
    assert abs(s - s_host) <= 1e-12 * abs(s_host)
    s_again, _ = dev.widom(CH4, 100_000, 298.0, seed=1, stream=1, precision=aa.FP64)
    assert s_again == s
    _, du32 = dev.widom(CH4, 100_000, 298.0, seed=1, stream=1, precision=aa.MIXED)
    mixed_bar("c0 CH4 widom 1e5 (split)", du32[:300], du_ref)
    mixed_bar("c0 CH4 widom 1e5 (split, all vs device FP64)", du32, du)

"""Run-time consistency checks on the device, and the chain checkpoint they are tested through.

- Ewald drift gate at every resync: the reference tests |fresh - committed| / max(1, |fresh|)
  <= 1e-10 of the Ewald energy (test_ewald_incremental_state.cc:137-222); the GCMC chains
  record that drift before each resync (every 100 accepted moves, monte_carlo_engine.cc:380-403)
  replaces the committed S(k), self and exclusion terms by the from-scratch ones.  A drift is
  forced from the host by loading a chain checkpoint whose S(k) was perturbed: the next resync
  must report it, and heal it.
- Chain checkpoints: a chain saved, loaded into a fresh chain set and continued follows the
  original step for step (same draws, same molecule counts, same positions bitwise).
- NaN / Inf sentinel of the grid tile kernel: a NaN charge spoils the framework's S(k) and so
  every Coulomb value; the framework's count rises by exactly the non-finite values written,
  and stays at zero for a clean framework."""
import struct

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework

pytestmark = pytest.mark.gpu

EWALD = (0.35, 7.5, 5)
HEADER = 144  # bytes before S(k) in a chain checkpoint record


def charged_system(ctx):
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    return fw, lj, device_framework(ctx, fw, lj, ewald=EWALD)


CHAINS = [(5e6, 298.0, 4, 2), (1e7, 298.0, 9, 3)]


def run_until_resyncs(gc, want, max_steps=40000, chunk=2000):
    done = 0
    while done < max_steps:
        gc.run(chunk)
        done += chunk
        drift, count = gc.resync_check()
        if count.min() >= want:
            return drift, count
    return gc.resync_check()


def test_resync_drift_within_the_reference_gate(ctx):
    _, _, dev = charged_system(ctx)
    gc = aa.Gcmc(dev, ol.co2_sites(), CHAINS, max_molecules=256)
    drift, count = run_until_resyncs(gc, 3)
    print(f"resyncs {count.tolist()}, max relative drift {drift.tolist()}")
    assert count.min() >= 3
    assert np.all(np.isfinite(drift)) and drift.max() <= 1e-10
    gc.reset_stats()
    drift, count = gc.resync_check()
    assert drift.max() == 0.0 and count.max() == 0


def test_uncharged_chains_report_no_resyncs(ctx):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    gc = aa.Gcmc(device_framework(ctx, fw, lj), [(0.0, 0.0, 0.0, 4, 0.0)], [(1e8, 300.0, 5, 0)], max_molecules=256)
    gc.run(3000)
    drift, count = gc.resync_check()
    assert drift[0] == 0.0 and count[0] == 0


def test_checkpoint_resumes_the_chain_step_for_step(ctx):
    _, _, dev = charged_system(ctx)
    a = aa.Gcmc(dev, ol.co2_sites(), CHAINS, max_molecules=256)
    a.run(1500)
    rec = a.save_chain(1)
    magic, version, ns, n_mol, nkh = struct.unpack_from("<IIiii", rec, 0)
    assert (magic, version, ns) == (0x4B434741, 1, 3) and n_mol == a.stats()[1]["n_molecules"] > 0
    assert len(rec) == HEADER + 16 * nkh + 24 * n_mol * ns
    # a fresh chain set whose chain 0 continues chain 1 of the first
    b = aa.Gcmc(dev, ol.co2_sites(), CHAINS[1:], max_molecules=256)
    b.load_chain(0, rec)
    assert b.save_chain(0) == rec
    ta = a.run(2000, trace=True)[1]
    tb = b.run(2000, trace=True)[0]
    assert np.array_equal(ta, tb)
    assert np.array_equal(a.positions(1), b.positions(0))
    assert a.electrostatic_energy(1) == b.electrostatic_energy(0)
    sa, sb = a.stats()[1], b.stats()[0]
    assert sa["attempted"] == sb["attempted"] and sa["accepted"] == sb["accepted"] and sa["steps"] == sb["steps"]


def test_forced_drift_is_reported_and_healed_by_the_resync(ctx):
    _, _, dev = charged_system(ctx)
    gc = aa.Gcmc(dev, ol.co2_sites(), CHAINS, max_molecules=256)
    gc.run(1500)
    rec = bytearray(gc.save_chain(0))
    nkh = struct.unpack_from("<i", rec, 16)[0]
    sre = np.frombuffer(bytes(rec[HEADER:HEADER + 8 * nkh]), dtype=np.float64).copy()
    sre *= 1.0 + 1e-4  # the cached structure factor drifts by 1e-4 relative
    rec[HEADER:HEADER + 8 * nkh] = sre.tobytes()
    gc.reset_stats()
    gc.load_chain(0, bytes(rec))
    drift, count = run_until_resyncs(gc, 1)
    print(f"forced drift: resyncs {count.tolist()}, max relative drift {drift.tolist()}")
    assert count[0] >= 1 and drift[0] > 1e-8
    assert count[1] >= 1 and drift[1] <= 1e-10
    # the resync replaced the perturbed S(k): later resyncs see no drift again
    gc.reset_stats()
    drift, count = run_until_resyncs(gc, 2)
    assert count.min() >= 2 and drift.max() <= 1e-10


def test_checkpoint_errors(ctx):
    _, _, dev = charged_system(ctx)
    gc = aa.Gcmc(dev, ol.co2_sites(), CHAINS, max_molecules=256)
    gc.run(500)
    rec = gc.save_chain(0)
    with pytest.raises(IndexError):
        gc.save_chain(2)
    with pytest.raises(RuntimeError):
        gc.load_chain(0, b"XXXX" + rec[4:])
    with pytest.raises(ValueError):
        gc.load_chain(0, rec + b"\0" * 8)
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    other = aa.Gcmc(device_framework(ctx, fw, lj), [(0.0, 0.0, 0.0, 4, 0.0)], [(1e8, 300.0, 5, 0)])
    with pytest.raises(ValueError):
        other.load_chain(0, rec)  # 3 charged sites and a k-set against 1 site, uncharged
    small = aa.Gcmc(dev, ol.co2_sites(), CHAINS, max_molecules=1)
    n_mol = struct.unpack_from("<i", rec, 12)[0]
    if n_mol > 1:
        with pytest.raises(ValueError):
            small.load_chain(0, rec)


@pytest.mark.parametrize("precision", [aa.FP64, aa.MIXED])
def test_nonfinite_sentinel_counts_what_a_nan_charge_spoils(ctx, precision):
    m = ol.cell_from_params(12.0, 13.0, 12.5, 60.0, 70.0, 80.0)
    fw = ol.synth_framework(m, 90, seed=5, stream=0, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1]])
    lj = ol.LJ(eps, sig, 4.5)
    counts = ol.grid_counts(fw.m, 0.7)
    clean = device_framework(ctx, fw, lj, ewald=(0.3, 4.5, 4))
    got = clean.grid([4], counts, coulomb=True, precision=precision)
    assert np.all(np.isfinite(got)) and clean.nonfinite_count() == 0
    # a NaN charge passes the neutrality test (|NaN| > tol is false, as in ewald.cc:181-195) and
    # spoils S(k), so every Coulomb value; the LJ table stays finite
    fw.q[0] = np.nan
    bad_fw = device_framework(ctx, fw, lj, ewald=(0.3, 4.5, 4))
    assert bad_fw.nonfinite_count() == 0
    got = bad_fw.grid([4], counts, coulomb=True, precision=precision)
    n_bad = int((~np.isfinite(got)).sum())
    print(f"precision {precision}: {n_bad} non-finite values of {got.size}")
    assert np.all(np.isfinite(got[0])) and not np.any(np.isfinite(got[1]))
    assert bad_fw.nonfinite_count() == n_bad == got[1].size
    bad_fw.grid([4], counts, coulomb=True, precision=precision)
    assert bad_fw.nonfinite_count() == 2 * n_bad
    assert clean.nonfinite_count() == 0

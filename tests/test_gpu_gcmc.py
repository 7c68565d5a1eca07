"""Parity of the batched device GCMC chains with the oracle's restatement of the reference
engine (MonteCarloEngine::run, monte_carlo_engine.cc:405-671), itself bitwise-equal to the
compiled reference (test_oracle_vs_reference.py).

With the same (seed, stream) a device chain takes the same draws in the same order, so it follows
the same trajectory: the molecule count after every step must match exactly, and final
positions and Ewald energy to 1e-9 relative.  (A flip would need two energies that differ only
by rounding to straddle an acceptance threshold; these seeds are checked not to.)"""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework

pytestmark = pytest.mark.gpu


def oracle_positions(g: ol.Gcmc):
    x, y, z, _ = g.particles()
    sites = g.molecule_sites()
    return np.stack([np.stack([x[sites[:, k]], y[sites[:, k]], z[sites[:, k]]], axis=1)
                     for k in range(sites.shape[1])], axis=1) if len(sites) else np.zeros((0, sites.shape[1], 3))


def test_gcmc_methane_chains_follow_oracle(ctx):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0), (3e7, 300.0, 6, 2), (1e8, 250.0, 7, 1)]
    gc = aa.Gcmc(dev, sites, chains, max_molecules=256)
    trace = gc.run(3000, trace=True)
    stats = gc.stats()
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, sites, f, t, seed, stream)
        ref_trace = g.run(3000, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        assert stats[c]["n_molecules"] == g.n_molecules > 0
        att, acc = g.counts
        assert stats[c]["attempted"] == att and stats[c]["accepted"] == acc
        pos = gc.positions(c)
        ref = oracle_positions(g)
        assert np.allclose(pos, ref, rtol=0, atol=1e-9)


def test_gcmc_charged_co2_chain_follows_oracle(ctx):
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 7.5, 5))
    chains = [(1e6, 298.0, 3, 1), (5e6, 298.0, 4, 2)]
    gc = aa.Gcmc(dev, ol.co2_sites(), chains, max_molecules=256)
    trace = gc.run(1500, trace=True)
    stats = gc.stats()
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, ol.co2_sites(), f, t, seed, stream, ewald=(0.35, 7.5, 5))
        ref_trace = g.run(1500, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        assert stats[c]["n_molecules"] == g.n_molecules > 0
        assert np.allclose(gc.positions(c), oracle_positions(g), rtol=0, atol=1e-9)
        e_dev, e_ref = gc.electrostatic_energy(c), g.electrostatic_energy()
        assert abs(e_dev - e_ref) <= 1e-9 * abs(e_ref)


def test_gcmc_errors(ctx):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    dev = device_framework(ctx, fw, lj)
    with pytest.raises(ValueError):
        aa.Gcmc(dev, [(0, 0, 0, 4, 0.0)], [(1e5, -1.0, 1, 0)])
    with pytest.raises(ValueError):
        aa.Gcmc(dev, [(0, 0, 0, 4, 0.0)], [(0.0, 300.0, 1, 0)])
    gc = aa.Gcmc(dev, [(0, 0, 0, 4, 0.0)], [(1e9, 300.0, 1, 0)], max_molecules=2)
    with pytest.raises(RuntimeError):
        gc.run(2000)
    assert gc.stats()[0]["error"] == 1  # capacity exceeded is reported, not silently dropped


def _proposal_weights(nb, seed):
    rng = np.random.default_rng(seed)
    return rng.random(int(np.prod(nb))) ** 4 + 1e-3  # strongly non-uniform, every bin > 0


def test_biased_methane_chains_follow_oracle(ctx):
    """Biased insertion (density over a 2x1x2 sub-cell, repeated): same draws, same trajectory."""
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0), (3e7, 300.0, 6, 2)]
    nb, reps, mix = (6, 7, 5), (2, 1, 2), 0.1
    w = _proposal_weights(nb, 1)
    gc = aa.Gcmc(dev, sites, chains, max_molecules=256)
    gc.set_proposal(w, nb, reps, mix)
    trace = gc.run(3000, trace=True)
    stats = gc.stats()
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, sites, f, t, seed, stream)
        g.set_proposal(w, nb, reps, mix)
        ref_trace = g.run(3000, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        att, acc = g.counts
        assert stats[c]["attempted"] == att and stats[c]["accepted"] == acc
        assert np.allclose(gc.positions(c), oracle_positions(g), rtol=0, atol=1e-9)


def test_biased_charged_co2_chain_follows_oracle(ctx):
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 7.5, 5))
    chains = [(5e6, 298.0, 4, 2)]
    nb, reps, mix = (5, 5, 5), (1, 1, 1), 0.05
    w = _proposal_weights(nb, 2)
    gc = aa.Gcmc(dev, ol.co2_sites(), chains, max_molecules=256)
    gc.set_proposal(w, nb, reps, mix)
    trace = gc.run(1500, trace=True)
    g = ol.Gcmc(lj, fw, ol.co2_sites(), 5e6, 298.0, 4, 2, ewald=(0.35, 7.5, 5))
    g.set_proposal(w, nb, reps, mix)
    ref_trace = g.run(1500, trace=True)
    assert np.array_equal(trace[0], ref_trace.astype(np.int32))
    assert gc.stats()[0]["n_molecules"] == g.n_molecules > 0
    assert np.allclose(gc.positions(0), oracle_positions(g), rtol=0, atol=1e-9)
    assert abs(gc.electrostatic_energy(0) - g.electrostatic_energy()) <= 1e-9 * abs(g.electrostatic_energy())


def test_biased_insertion_ideal_gas_on_device(ctx):
    """Detailed balance on the device: an ideal gas stays Poisson(f V / T) and uniform under a
    proposal that favours one eighth of the cell 50:1."""
    a, t, mean_n = 20.0, 300.0, 4.0
    m = np.diag([a, a, a])
    dev = aa.Framework(ctx, m, [1.0], [1.0], [1.0], np.zeros(1, dtype=np.uint32), [0.0, 0.0], [1.0, 1.0], 5.0)
    f_pa = mean_n * t / a ** 3 * (1.380649e-23 * 1e30)
    chains = [(f_pa, t, 100 + c, c) for c in range(32)]
    gc = aa.Gcmc(dev, [(0.0, 0.0, 0.0, 1, 0.0)], chains, max_molecules=64)
    w = np.ones(8)
    w[0] = 50.0
    gc.set_proposal(w, (2, 2, 2), (1, 1, 1), 0.0)
    counts = np.zeros(8)
    for _ in range(40):
        gc.run(250)
        for c in range(len(chains)):
            f = gc.positions(c)[:, 0, :] / a
            f -= np.floor(f)
            b = np.minimum((f * 2).astype(int), 1)
            np.add.at(counts, (b[:, 0] * 2 + b[:, 1]) * 2 + b[:, 2], 1)
    st = gc.stats()
    mean = sum(s["sum_molecules"] for s in st) / sum(s["steps"] for s in st)
    occ = counts / counts.sum()
    print(f"ideal gas, biased insertion: <N> {mean:.3f} (expect {mean_n}), occupancy {np.round(occ, 3)}")
    assert mean == pytest.approx(mean_n, abs=0.15)
    assert np.all(np.abs(occ - 1 / 8) < 0.02)
    assert all(s["error"] == 0 for s in st)


def test_proposal_errors(ctx):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    dev = device_framework(ctx, fw, lj)
    gc = aa.Gcmc(dev, [(0, 0, 0, 4, 0.0)], [(1e5, 300.0, 1, 0)])
    with pytest.raises(ValueError):
        gc.set_proposal(np.array([1.0, 0.0]), (2, 1, 1), (1, 1, 1), 0.0)
    with pytest.raises(ValueError):
        gc.set_proposal(np.ones(2), (2, 1, 1), (1, 1, 1), 1.5)
    off_centre = aa.Gcmc(dev, [(0.5, 0, 0, 4, 0.0), (1.5, 0, 0, 4, 0.0)], [(1e5, 300.0, 1, 0)])
    with pytest.raises(ValueError):
        off_centre.set_proposal(np.ones(2), (2, 1, 1), (1, 1, 1), 0.1)
    gc.set_proposal(None, (1, 1, 1), (1, 1, 1))


def test_gcmc_methane_with_energy_grid_follows_oracle(ctx):
    """Grid mode (MonteCarloEngine given a FrameworkEnergyGrid, monte_carlo_engine.cc:341-374):
    the oracle's grid mode is bitwise the compiled reference's (test_oracle_vs_reference.py).
    Chains 0-2 use the oracle's tables; a fourth chain object uses tables built on the device
    (FP64, within 1e-10 of them) and must follow the same trajectory.  Switching the grid off
    mid-run goes back to the direct scan on both sides."""
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    counts = ol.grid_counts(fw.m, 0.5)
    tables = ol.grid_build(lj, fw, [4], counts)
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0), (3e7, 300.0, 6, 2), (1e8, 250.0, 7, 1)]
    gc = aa.Gcmc(dev, sites, chains, max_molecules=256)
    gc.set_energy_grid(counts, [4], tables)
    buf = aa.DeviceBuffer(tables.size)
    dev.grid_device([4], counts, int(buf))
    gd = aa.Gcmc(dev, sites, chains[:1], max_molecules=256)
    gd.set_energy_grid(counts, [4], buf)
    buf.free()  # the chains keep their own copy
    trace = gc.run(3000, trace=True)
    trace_d = gd.run(3000, trace=True)
    gc.set_energy_grid(None, None, None)
    trace2 = gc.run(1000, trace=True)
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, sites, f, t, seed, stream)
        g.set_grid(counts, [4], tables)
        ref_trace = g.run(3000, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        if c == 0:
            assert np.array_equal(trace_d[0], ref_trace.astype(np.int32)), "device-built tables diverged"
            assert np.allclose(gd.positions(0), oracle_positions(g), rtol=0, atol=1e-9)
        g.set_grid(None, None, None)
        ref_trace2 = g.run(1000, trace=True)
        assert np.array_equal(trace2[c], ref_trace2.astype(np.int32)), f"chain {c} diverged after the grid was removed"
        assert np.allclose(gc.positions(c), oracle_positions(g), rtol=0, atol=1e-9)
        att, acc = g.counts
        assert gc.stats()[c]["attempted"] == att and gc.stats()[c]["accepted"] == acc


def test_gcmc_charged_co2_with_energy_grid_follows_oracle(ctx):
    """Grid mode with Ewald: LJ of C and O from their tables, real-space Coulomb with the
    framework summed directly, S(k) as without a grid."""
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 7.5)
    counts = ol.grid_counts(fw.m, 0.6)
    tables = ol.grid_build(lj, fw, [4, 5], counts)
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 7.5, 5))
    chains = [(1e6, 298.0, 3, 1), (5e6, 298.0, 4, 2)]
    gc = aa.Gcmc(dev, ol.co2_sites(), chains, max_molecules=256)
    gc.set_energy_grid(counts, [4, 5], tables)
    trace = gc.run(1500, trace=True)
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, ol.co2_sites(), f, t, seed, stream, ewald=(0.35, 7.5, 5))
        g.set_grid(counts, [4, 5], tables)
        ref_trace = g.run(1500, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        assert gc.stats()[c]["n_molecules"] == g.n_molecules > 0
        assert np.allclose(gc.positions(c), oracle_positions(g), rtol=0, atol=1e-9)
        e_dev, e_ref = gc.electrostatic_energy(c), g.electrostatic_energy()
        assert abs(e_dev - e_ref) <= 1e-9 * abs(e_ref)


def test_gcmc_energy_grid_errors(ctx):
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    dev = device_framework(ctx, fw, lj)
    gc = aa.Gcmc(dev, [(0, 0, 0, 4, 0.0)], [(1e5, 300.0, 1, 0)])
    with pytest.raises(IndexError):  # framework_energy_grid.cc:254-257: no table for species 4
        gc.set_energy_grid((4, 4, 4), [3], np.zeros((1, 4, 4, 4)))
    with pytest.raises(ValueError):
        gc.set_energy_grid((4, 4, 4), [4], np.zeros((1, 4, 4, 3)))
    with pytest.raises(ValueError):
        gc.set_energy_grid((0, 4, 4), [4], np.zeros(0))
    gc.run(10)
    assert gc.stats()[0]["steps"] == 10


def test_guest_cells_follow_the_full_scan_and_the_oracle(ctx):
    """Guest-guest candidates through guest cells (the engine's guest CellList): in a 28 A box with
    a 6.5 A cutoff there are 4 x 4 x 4 cells, and with a few hundred methanes the chains
    still follow the oracle's full-scan trajectory step for step, as the reference's cell list
    follows its full scan (test_gcmc_cell_list_bit_identical.cc:111-175); forcing the full scan
    gives the same trajectory."""
    print("gcmc guest cells: CH4, 28 A box, 3000 steps", flush=True)
    fw, lj = ol.c0_system(40, a=28.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0), (3e8, 300.0, 6, 2)]
    gc = aa.Gcmc(dev, sites, chains, max_molecules=1024)
    assert gc.guest_cells() == (0, 0, 0)  # the full scan is the default
    gc.set_guest_cells(1)
    assert gc.guest_cells() == (4, 4, 4)
    gc.set_guest_cells(2)  # every step, not only from 1024 guest sites on
    assert gc.guest_cells() == (4, 4, 4)
    trace = gc.run(3000, trace=True)
    full = aa.Gcmc(dev, sites, chains, max_molecules=1024)
    assert full.guest_cells() == (0, 0, 0)
    trace_full = full.run(3000, trace=True)
    assert np.array_equal(trace, trace_full)
    for c, (f, t, seed, stream) in enumerate(chains):
        g = ol.Gcmc(lj, fw, sites, f, t, seed, stream)
        ref_trace = g.run(3000, trace=True)
        assert np.array_equal(trace[c], ref_trace.astype(np.int32)), f"chain {c} trajectory diverged"
        assert g.n_molecules > 100
        assert np.allclose(gc.positions(c), oracle_positions(g), rtol=0, atol=1e-9)


def test_guest_cells_forced_on_a_small_charged_cell(ctx):
    """Mode 2 uses cells even with 3 or fewer per axis, where each cell's periodic neighbours
    cover the box: the charged CO2 chains still follow the oracle."""
    print("gcmc guest cells forced: charged CO2, 1500 steps", flush=True)
    m = ol.cell_from_params(16.0, 17.0, 18.0, 80.0, 95.0, 100.0)
    fw = ol.synth_framework(m, 120, seed=4, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    lj = ol.LJ(eps, sig, 5.0)
    dev = device_framework(ctx, fw, lj, ewald=(0.35, 5.0, 5))
    chains = [(5e6, 298.0, 4, 2)]
    gc = aa.Gcmc(dev, ol.co2_sites(), chains, max_molecules=256)
    gc.set_guest_cells(1)
    assert gc.guest_cells() == (0, 0, 0)  # fewer than 4 cells per axis
    gc.set_guest_cells(2)
    cells = gc.guest_cells()
    assert min(cells) >= 2 and max(cells) <= 3
    trace = gc.run(1500, trace=True)
    g = ol.Gcmc(lj, fw, ol.co2_sites(), *chains[0], ewald=(0.35, 5.0, 5))
    ref_trace = g.run(1500, trace=True)
    assert np.array_equal(trace[0], ref_trace.astype(np.int32))
    assert g.n_molecules > 0
    assert np.allclose(gc.positions(0), oracle_positions(g), rtol=0, atol=1e-9)
    e_dev, e_ref = gc.electrostatic_energy(0), g.electrostatic_energy()
    assert abs(e_dev - e_ref) <= 1e-9 * abs(e_ref)
    with pytest.raises(IndexError):
        aa.Gcmc(dev, ol.co2_sites(), chains, max_molecules=2000).set_guest_cells(2)
    with pytest.raises(ValueError):
        gc.set_guest_cells(3)


def test_closing_a_context_closes_what_was_created_on_it():
    """Objects still open when their context closes are closed first, newest first, so none of
    their native handles is left pointing into the freed context."""
    c = aa.Context(0)
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff = 6.5
    dev = device_framework(c, fw, lj)
    chains = aa.Gcmc(dev, [(0.0, 0.0, 0.0, 4, 0.0)], [(1e6, 300.0, 1, 0)], max_molecules=16)
    chains.run(10)
    c.close()
    assert not chains._h and not dev._h and not c._h
    chains.close()


def test_gcmc_overflow_stops_the_chain_at_the_step_that_does_not_fit(ctx):
    """max_molecules 3 at a high fugacity: the fourth accepted insertion does not fit.  run raises,
    the trace reads -1 from that step on, and the chain's state and counters are those of the
    oracle after exactly the steps that ran; a second run raises without advancing."""
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0)]
    gc = aa.Gcmc(dev, sites, chains, max_molecules=3)
    with pytest.raises(RuntimeError) as exc:
        gc.run(3000, trace=True)
    trace = exc.value.trace[0]
    st = gc.stats()[0]
    ran = int(st["steps"])
    assert 0 < ran < 3000 and st["n_molecules"] == 3
    assert np.all(trace[ran:] == -1) and np.all(trace[:ran] >= 0)
    g = ol.Gcmc(lj, fw, sites, *chains[0])
    ref = g.run(ran, trace=True)
    assert np.array_equal(trace[:ran], ref.astype(np.int32))
    att, acc = g.counts
    assert st["attempted"] == att and st["accepted"] == acc
    assert np.allclose(gc.positions(0), oracle_positions(g), rtol=0, atol=1e-9)
    with pytest.raises(RuntimeError):
        gc.run(100)
    assert gc.stats()[0]["steps"] == ran
    gc.close()


def test_gcmc_graph_replay_equals_launches(ctx):
    """run() replayed from a captured hipGraph continues the chains exactly as direct launches."""
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chains = [(1e8, 300.0, 5, 0), (3e7, 300.0, 6, 2)]
    a = aa.Gcmc(dev, sites, chains, max_molecules=256)
    b = aa.Gcmc(dev, sites, chains, max_molecules=256)
    b.use_graph(True)
    for _ in range(5):
        a.run(300)
        b.run(300)
    b.run(100)  # a different step count recaptures
    a.run(100)
    for c in range(2):
        assert a.stats()[c] == b.stats()[c]
        assert np.array_equal(a.positions(c), b.positions(c))
    a.close()
    b.close()

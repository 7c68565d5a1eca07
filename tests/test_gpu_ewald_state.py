"""Parity of the device EwaldIncrementalState with the FP64 oracle
(ewald_incremental_state.cc:55-257) and the invariants the reference tests assert
(test_ewald_incremental_state.cc:137-430): drift against a from-scratch recompute,
insert/delete antisymmetry, rollback leaves the state bit-identical.
Tolerance: FP64 1e-10 relative (energies floored at 1 K; S(k) against max|S|)."""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework, rel_err

pytestmark = pytest.mark.gpu

ALPHA, KMAX, RC = 0.3, 6, 9.0


def system():
    m = ol.cell_from_params(19.0, 20.0, 21.0, 75.0, 85.0, 100.0)
    fw = ol.synth_framework(m, 300, seed=13, stream=0, charge_sigma=0.5)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.CO2_C[0], ol.CO2_O[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.CO2_C[1], ol.CO2_O[1]])
    return fw, ol.LJ(eps, sig, 9.0)


def random_co2(rng, m):
    c = rng.uniform(0, 1, 3) @ m
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    return np.array([[*(c - 1.16 * u), -0.35], [*c, 0.70], [*(c + 1.16 * u), -0.35]])


def check_raw(gpu_raw, oracle_raw, nk):
    s_scale = np.abs(oracle_raw[: 2 * nk]).max()
    assert np.abs(gpu_raw[: 2 * nk] - oracle_raw[: 2 * nk]).max() <= 1e-10 * s_scale
    assert rel_err(gpu_raw[2 * nk:], oracle_raw[2 * nk:]).max() <= 1e-10


def test_state_init_matches_oracle(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    ost = ol.EwaldState(fw, ALPHA, KMAX)
    assert dev.num_k == ost.nk == 948
    check_raw(st.raw(), ost.raw(), ost.nk)
    assert rel_err(st.committed_energy(), ost.committed()) <= 1e-10


def test_propose_commit_sequence_tracks_oracle(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    ost = ol.EwaldState(fw, ALPHA, KMAX)
    rng = np.random.default_rng(5)
    molecules = []
    worst = 0.0
    for step in range(60):
        kind = rng.integers(3) if molecules else 0
        if kind == 0:
            new = random_co2(rng, fw.m)
            d_gpu, d_ref = st.propose([], new), ost.propose([], new)
            st.commit([], new)
            ost.commit([], new)
            molecules.append(new)
        elif kind == 1:
            i = rng.integers(len(molecules))
            old = molecules[i]
            new = old.copy()
            new[:, :3] += rng.uniform(-1.5, 1.5, 3)
            d_gpu, d_ref = st.propose(old, new), ost.propose(old, new)
            st.commit(old, new)
            ost.commit(old, new)
            molecules[i] = new
        else:
            i = rng.integers(len(molecules))
            old = molecules.pop(i)
            d_gpu, d_ref = st.propose(old, []), ost.propose(old, [])
            st.commit(old, [])
            ost.commit(old, [])
        worst = max(worst, float(rel_err(d_gpu, d_ref)))
        assert rel_err(d_gpu, d_ref) <= 1e-10, (step, d_gpu, d_ref)
    check_raw(st.raw(), ost.raw(), ost.nk)
    assert rel_err(st.committed_energy(), ost.committed()) <= 1e-10
    print(f"60 propose/commit steps: worst rel err {worst:.3e}")


def test_rollback_is_bit_exact_and_batch_equals_single(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    rng = np.random.default_rng(6)
    st.commit([], random_co2(rng, fw.m))
    before = st.raw()
    trials = np.stack([random_co2(rng, fw.m) for _ in range(37)])
    singles = np.array([st.propose_insertion(t) for t in trials])
    after = st.raw()
    assert np.array_equal(before, after)  # rejected trials never touch the committed cache
    batch = st.propose_batch(np.zeros((37, 0, 4)), trials)
    assert np.array_equal(batch, singles)


def test_large_batches_take_the_run_kernel_and_match_oracle(ctx):
    """Batches of 64 or more proposals without commit run the k-set run kernel (several
    proposals per workgroup, per-axis factors by recurrence): within 1e-10 of the oracle's
    proposals and 1e-12 of the one-at-a-time path."""
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    ost = ol.EwaldState(fw, ALPHA, KMAX)
    rng = np.random.default_rng(8)
    for _ in range(5):
        m = random_co2(rng, fw.m)
        st.commit([], m)
        ost.commit([], m)
    n = 200
    ins = np.stack([random_co2(rng, fw.m) for _ in range(n)])
    old = np.stack([random_co2(rng, fw.m) for _ in range(n)])
    new = old.copy()
    new[:, :, :3] += rng.uniform(-1.5, 1.5, (n, 1, 3))
    b_ins = st.propose_batch(np.zeros((n, 0, 4)), ins)
    b_mov = st.propose_batch(old, new)
    worst = 0.0
    for i in range(n):
        for got, ref in ((b_ins[i], ost.propose([], ins[i])), (b_mov[i], ost.propose(old[i], new[i]))):
            worst = max(worst, float(rel_err(got, ref)))
    print(f"run kernel, {n} insertions and {n} moves: worst rel err vs oracle {worst:.3e}")
    assert worst <= 1e-10
    singles = np.array([st.propose_insertion(t) for t in ins[:20]])
    assert rel_err(b_ins[:20], singles).max() <= 1e-12


def test_insert_then_delete_is_antisymmetric(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    rng = np.random.default_rng(7)
    mol = random_co2(rng, fw.m)
    d_ins = st.propose_insertion(mol)
    st.commit([], mol)
    d_del = st.propose_deletion(mol)
    assert abs(d_ins + d_del) <= 1e-9 * abs(d_ins)


def test_resync_and_recompute_match_oracle(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj, ewald=(ALPHA, RC, KMAX))
    st = aa.EwaldState(dev)
    ost = ol.EwaldState(fw, ALPHA, KMAX)
    rng = np.random.default_rng(8)
    mols = [random_co2(rng, fw.m) for _ in range(5)]
    for mdl in mols:
        st.commit([], mdl)
        ost.commit([], mdl)
    allp = np.concatenate([mdl for mdl in mols])
    x = np.concatenate([fw.x, allp[:, 0]])
    y = np.concatenate([fw.y, allp[:, 1]])
    z = np.concatenate([fw.z, allp[:, 2]])
    q = np.concatenate([fw.q, allp[:, 3]])
    molecules = [list(range(fw.n + 3 * i, fw.n + 3 * i + 3)) for i in range(5)]
    full = ol.Framework(fw.m, x, y, z, q, np.zeros(len(x), dtype=np.uint32))
    drift = abs(st.recompute_from_scratch(x, y, z, q, molecules) - st.committed_energy())
    ref = ost.recompute(full, molecules)
    assert rel_err(st.recompute_from_scratch(x, y, z, q, molecules), ref) <= 1e-10
    assert drift <= 1e-10 * abs(ref)
    st.resync(x, y, z, q, molecules)
    ost.resync(full, molecules)
    check_raw(st.raw(), ost.raw(), ost.nk)


def test_state_errors(ctx):
    fw, lj = system()
    dev = device_framework(ctx, fw, lj)  # uncharged: no Ewald parameters
    with pytest.raises(ValueError):
        aa.EwaldState(dev)
    bad = ol.Framework(fw.m, fw.x, fw.y, fw.z, fw.q + 0.01, fw.species)
    with pytest.raises(ValueError):
        device_framework(ctx, bad, lj, ewald=(ALPHA, RC, KMAX))  # not charge-neutral
    with pytest.raises(ValueError):
        device_framework(ctx, fw, lj, ewald=(-1.0, RC, KMAX))
    with pytest.raises(ValueError):
        device_framework(ctx, fw, ol.LJ(lj.eps, lj.sigma, 50.0))  # cutoff above L_perp/2

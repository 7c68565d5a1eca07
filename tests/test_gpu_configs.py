"""The benchmark configurations beyond the headline, checked on the device against the oracle at
reduced size:

- c2: eight of the 256 synthetic frameworks (cells 10-40 A, angles 60-120 deg, 738-10688 atoms
  after the minimum supercell, rc 12.8 A, CH4, 0.2 A) built in one batched, graph-replayed
  GridPlan, MIXED and FP64, against FrameworkEnergyGrid's direct sum
  (framework_energy_grid.cc:193-249) on sampled and hard nodes of each;
- c3: CO2 chains of the c3 set on the c1 supercell (4800 atoms, Ewald alpha 0.265, kmax 8 =
  2204 k, Peng-Robinson fugacities) step for step against the oracle's restatement of the
  engine's move loop (monte_carlo_engine.cc:405-671) and of EwaldIncrementalState
  (ewald_incremental_state.cc:95-136);
- the S(k) state at kmax 8 on the c1 supercell: a propose/commit sequence and a batch of
  proposals against the oracle.
Tolerances as in the headline tests: FP64 1e-10 relative (|E| floored at 1 K), MIXED 1e-5 on
every checked node with |E| > 1 K (capped nodes left out)."""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from aleator_amd import workloads as wl
from gpu_helpers import device_framework, rel_err
from test_gpu_ewald_state import check_raw, random_co2
from test_gpu_grid import hard_nodes, mixed_vs_oracle

pytestmark = pytest.mark.gpu

C2_PICK = [0, 37, 73, 110, 146, 183, 219, 255]


def test_c2_batched_frameworks_against_oracle(ctx):
    print(f"c2 frameworks {C2_PICK}: one batched plan, MIXED and FP64", flush=True)
    fws, eps, sig = wl.c2(256)
    sel = [fws[f] for f in C2_PICK]
    devs = [aa.Framework(ctx, f["sup"], f["x"], f["y"], f["z"], f["sp"], eps, sig, f["rc"]) for f in sel]
    counts = [f["counts"] for f in sel]
    nodes = [int(np.prod(c)) for c in counts]
    out = {}
    for prec in (aa.MIXED, aa.FP64):
        bufs = [aa.DeviceBuffer(n) for n in nodes]
        plan = aa.GridPlan(ctx, devs, [4], counts, [int(b.ptr.value) for b in bufs], precision=prec,
                           grid_lattices=[f["prim"] for f in sel], tile=-1, use_graph=True)
        plan.execute()
        plan.execute()  # a replay
        out[prec] = [b.to_host() for b in bufs]
        plan.close()
        for b in bufs:
            b.free()
    lj = ol.LJ(eps, sig, 12.8)
    for k, f in enumerate(sel):
        ofw = ol.Framework(f["sup"], f["x"], f["y"], f["z"], f["q"], np.ascontiguousarray(f["sp"], dtype=np.uint32))
        pick = hard_nodes(out[aa.FP64][k], 300, 100, seed=20 + k)
        ref = ol.grid_nodes_lj(lj, ofw, f["prim"], counts[k], 4, pick)
        e64 = rel_err(out[aa.FP64][k][pick], ref)
        print(f"c2 framework {C2_PICK[k]} ({len(f['x'])} atoms, {counts[k]}): FP64 max rel {e64.max():.3e}")
        assert e64.max() <= 1e-10
        mixed_vs_oracle(f"c2 framework {C2_PICK[k]}", out[aa.MIXED][k][pick], ref)


def c1_device(ctx):
    prim, lj = ol.c1_system()
    sc = ol.replicate(prim, (2, 2, 2))
    return prim, sc, lj, device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))


def test_c3_chains_follow_oracle_at_kmax8(ctx):
    """Three chains of the c3 set for 1000 steps each: 19 and 36 insert molecules that then
    translate and rotate under Ewald, 42 inserts and deletes (the c1 framework is dense: most c3
    chains accept no insertion within 1000 steps)."""
    print("c3 chains 19, 36, 42 on c1 at kmax 8, 1000 steps", flush=True)
    prim, sc, lj, dev = c1_device(ctx)
    chains = wl.c3_chains()
    pick = [19, 36, 42]
    sel = [chains[c] for c in pick]
    gc = aa.Gcmc(dev, wl.CO2_SITES, sel, max_molecules=256)
    trace = gc.run(1000, trace=True)
    stats = gc.stats()
    for c, (f, t, seed, stream) in enumerate(sel):
        g = ol.Gcmc(lj, sc, ol.co2_sites(), f, t, seed, stream, ewald=(0.265, 12.0, 8))
        ref = g.run(1000, trace=True)
        assert np.array_equal(trace[c], ref.astype(np.int32)), f"c3 chain {pick[c]} diverged"
        att, acc = g.counts
        assert stats[c]["attempted"] == att and stats[c]["accepted"] == acc
        e_dev, e_ref = gc.electrostatic_energy(c), g.electrostatic_energy()
        assert abs(e_dev - e_ref) <= 1e-9 * max(abs(e_ref), 1.0)
        print(f"  chain {pick[c]}: N after 1000 steps {ref[-1]}, accepted {acc}")
    gc.close()


def test_sk_state_at_kmax8_on_c1(ctx):
    """EwaldIncrementalState on the c1 supercell (2204 k, half set 1102): 40 propose/commit steps
    against the oracle, then a batch of 128 insertions and moves (the run kernel).  Bar: 1e-10
    relative with |dE| floored at 10 K.  Here |S(k)| ~ 35 (4800 charges), so single k-terms of
    the reciprocal delta reach ~4e4 K and a delta of a few K is a cancellation: FP64 sums of
    those terms in different orders (the oracle's included) differ by ~1e-10 K, as for the
    Coulomb grid (test_oracle_coulomb_precision.py)."""
    print("S(k) state on c1 at kmax 8", flush=True)
    prim, sc, lj, dev = c1_device(ctx)
    assert dev.num_k == 2204
    st = aa.EwaldState(dev)
    ost = ol.EwaldState(sc, 0.265, 8)
    check_raw(st.raw(), ost.raw(), ost.nk)
    rng = np.random.default_rng(21)
    mols = []
    worst = 0.0
    for step in range(40):
        if not mols or rng.random() < 0.5:
            new = random_co2(rng, sc.m)
            d, r = st.propose([], new), ost.propose([], new)
            st.commit([], new)
            ost.commit([], new)
            mols.append(new)
        else:
            i = rng.integers(len(mols))
            old = mols[i]
            new = old.copy()
            new[:, :3] += rng.uniform(-1.5, 1.5, 3)
            d, r = st.propose(old, new), ost.propose(old, new)
            st.commit(old, new)
            ost.commit(old, new)
            mols[i] = new
        worst = max(worst, float(rel_err(d, r, floor=10.0)))
        assert rel_err(d, r, floor=10.0) <= 1e-10, (step, d, r)
    check_raw(st.raw(), ost.raw(), ost.nk)
    n = 128
    ins = np.stack([random_co2(rng, sc.m) for _ in range(n)])
    old = np.stack([random_co2(rng, sc.m) for _ in range(n)])
    new = old.copy()
    new[:, :, :3] += rng.uniform(-1.5, 1.5, (n, 1, 3))
    b_ins = st.propose_batch(np.zeros((n, 0, 4)), ins)
    b_mov = st.propose_batch(old, new)
    for i in range(0, n, 8):
        worst = max(worst, float(rel_err(b_ins[i], ost.propose([], ins[i]), floor=10.0)),
                    float(rel_err(b_mov[i], ost.propose(old[i], new[i]), floor=10.0)))
    print(f"  worst rel err vs oracle {worst:.3e}")
    assert worst <= 1e-10
    st.close()

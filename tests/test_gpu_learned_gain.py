"""What the trained insertion density gains over plain uniform insertion, at equilibrium
(gcmc_acceptance.hpp:28-51 with the proposal correction; the move loop of
monte_carlo_engine.cc:405-439).

System: the c0 framework (cubic 25.83 A, 424 atoms), CH4 at 298 K, rc 12.8 A truncated, at
5.9 and 30 bar (Peng-Robinson), 32 seeds each, both methods from empty cells; a framework where
plain GCMC equilibrates within the budget (the c1 framework's pockets need millions of plain
steps).  The density: the c0 CH4 grid built on the device, 2e5 samples, an MLP trained on the
device (1000 Adam steps), 64^3 bins.  Asserted:
- the learned chains reach 90 % of plain's equilibrium loading in fewer steps;
- after a burn-in, the two methods' loadings agree within 2 sigma at both pressures (the
  proposal correction keeps the sampled distribution exact);
- insertion acceptance rises;
- the 32-seed standard error of the mean loading after equal steps is smaller with the density
  (faster decorrelation of the molecule count)."""
import numpy as np
import pytest

import aleator_amd as aa
from aleator_amd import learned, workloads as wl

pytestmark = pytest.mark.gpu

PRESSURES_BAR = (5.8799, 30.0)
SEEDS = 32
NB = (64, 64, 64)


def c0_methane(ctx):
    w = wl.c0()
    fw = aa.Framework(ctx, w["m"], w["x"], w["y"], w["z"], w["sp"], w["eps"], w["sig"], w["rc"])
    chains = []
    for i, p in enumerate(PRESSURES_BAR):
        f = aa.peng_robinson_fugacity(190.6, 4.599e6, 0.011, 298.0, p * 1e5)
        chains += [(f, 298.0, 500 + s, 64 * i + s) for s in range(SEEDS)]
    return w, fw, chains


def run(fw, chains, weights, fill_steps, seg, prod_steps):
    """Per-segment mean loadings [n_seg][pressures] over the filling phase, then per-chain time
    averages [pressures][seeds] over a production block, and the insertion acceptance."""
    g = aa.Gcmc(fw, wl.CH4_SITES, chains, max_molecules=512)
    if weights is not None:
        g.set_proposal(weights, NB, (1, 1, 1), 0.05)
    fill = []
    for _ in range(fill_steps // seg):
        g.reset_stats()
        g.run(seg)
        st = g.stats()
        n = np.array([s["sum_molecules"] / s["steps"] for s in st]).reshape(len(PRESSURES_BAR), SEEDS)
        fill.append(n.mean(axis=1))
    g.reset_stats()
    g.run(prod_steps)
    st = g.stats()
    assert sum(s["error"] for s in st) == 0
    prod = np.array([s["sum_molecules"] / s["steps"] for s in st]).reshape(len(PRESSURES_BAR), SEEDS)
    acc = sum(s["accepted"][2] for s in st) / sum(s["attempted"][2] for s in st)
    g.close()
    return np.array(fill), prod, acc


def test_learned_density_gain_at_equilibrium(ctx):
    print("learned vs plain GCMC, c0 CH4 at 5.9 and 30 bar, 32 seeds", flush=True)
    w, fw, chains = c0_methane(ctx)
    tab = fw.grid([4], w["counts"], precision=aa.MIXED)
    frac, e = learned.sample_grid(tab, [1.0], w["counts"], 200_000, low_fraction=0.5, low_cut=1000.0)
    ld = learned.LearnedDensity(ctx, max_rows=8192)
    ld.fit(frac, e, e_clip=1000.0, n_iters=1000, batch=4096, lr=1e-3)
    weights = ld.bin_weights(NB, 298.0)
    seg = 250
    fill_p, prod_p, acc_p = run(fw, chains, None, 10_000, seg, 60_000)
    fill_l, prod_l, acc_l = run(fw, chains, weights, 10_000, seg, 60_000)
    n_p, n_l = prod_p.mean(axis=1), prod_l.mean(axis=1)
    s_p = prod_p.std(axis=1, ddof=1) / np.sqrt(SEEDS)
    s_l = prod_l.std(axis=1, ddof=1) / np.sqrt(SEEDS)

    def steps_to(fill, target):
        hit = np.flatnonzero(fill >= target)
        return (hit[0] + 1) * seg if hit.size else np.inf

    for i, p in enumerate(PRESSURES_BAR):
        t_p = steps_to(fill_p[:, i], 0.9 * n_p[i])
        t_l = steps_to(fill_l[:, i], 0.9 * n_p[i])
        print(f"  {p:g} bar: plain {n_p[i]:.4f} +- {s_p[i]:.4f}, learned {n_l[i]:.4f} +- {s_l[i]:.4f}; "
              f"steps to 90 % of plain's loading: plain {t_p}, learned {t_l}")
        assert abs(n_l[i] - n_p[i]) <= 2.0 * np.hypot(s_p[i], s_l[i])
        assert t_l < t_p
        assert s_l[i] < s_p[i]
    print(f"  insertion acceptance: plain {acc_p:.4f}, learned {acc_l:.4f}")
    assert acc_l > 2.0 * acc_p

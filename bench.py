#!/usr/bin/env python3
"""Benchmark of the MI355X adsorption hot path.

Headline workload (one step = one complete build, device-resident input to device-resident
output): the charged triclinic framework -- rhombohedral 24 A cell (60/60/60 deg), 600 atoms,
charges N(0, 0.5 e) shifted to zero sum, automatic 2x2x2 supercell (4800 atoms) -- with a rigid
CO2 probe: three tables (C LJ, O LJ, Ewald Coulomb with alpha 0.265, kmax 8, rc 12 A) at
0.15 A over the primitive cell (131^3 nodes), in MIXED precision (FP64 pair arithmetic with the
hardware reciprocal square root refined by one Newton step and the erf kernel polynomial cut at
1e-12; FP64 accumulation and reciprocal space).  value = grid table points per second (nodes x
tables).

Also reported (extra): the FP64 build of the same grid, the device's error against the FP64 oracle
on sampled and hard nodes, the cubic 424-atom CH4 LJ grid (129^3), Widom insertion rates and the
other legs.  Every host-side timing (the cpu_baseline leg) runs the oracle restatement
(oracle/liboracle.so); nothing else in this file touches oracle/.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published dense FP32 vector
PEAK_FP64_TFLOPS = 78.6
# tile shapes of the grid kernel by index: nodes per tile along x, y, z and nodes per thread along z
TILE_SHAPES = {0: (8, 8, 4, 1), 1: (8, 8, 4, 2), 2: (8, 8, 4, 4), 3: (4, 8, 8, 2)}
# algorithmic FLOP per (node, atom within the cutoff) pair
F_LJ = 20.0
F_REAL = 30.0

C0_TYPES = ((52.8, 3.43), (30.2, 3.12), (22.1, 2.57), (62.4, 2.46))
PROPORTIONS = (0.45, 0.30, 0.17, 0.08)


def cell_from_params(a, b, c, alpha, beta, gamma):
    al, be, ga = np.radians([alpha, beta, gamma])
    ca, cb, cg, sg = np.cos(al), np.cos(be), np.cos(ga), np.sin(ga)
    cy = (ca - cb * cg) / sg
    return np.array([[a, 0, 0], [b * cg, b * sg, 0], [c * cb, c * cy, c * np.sqrt(1 - cb * cb - cy * cy)]])


def replicate(m, x, y, z, q, sp, reps):
    xs, ys, zs, qs, ss = [], [], [], [], []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                o = i * m[0] + j * m[1] + k * m[2]
                xs.append(x + o[0]); ys.append(y + o[1]); zs.append(z + o[2]); qs.append(q); ss.append(sp)
    mm = m * np.array(reps)[:, None]
    return mm, np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(qs), np.concatenate(ss)


def c1_workload(aa):
    m = cell_from_params(24.0, 24.0, 24.0, 60.0, 60.0, 60.0)
    x, y, z, q, sp = aa.synth_framework(m, 600, seed=1, stream=0, charge_sigma=0.5, charge_stream=2)
    eps = np.array([t[0] for t in C0_TYPES] + [27.0, 79.0])
    sig = np.array([t[1] for t in C0_TYPES] + [2.80, 3.05])
    reps = aa.supercell_replication(m, 12.0)  # (2, 2, 2)
    sm, sx, sy, sz, sq, ss = replicate(m, x, y, z, q, sp, reps)
    return dict(prim=m, sup=sm, x=sx, y=sy, z=sz, q=sq, sp=ss, eps=eps, sig=sig, rc=12.0,
                alpha=0.265, kmax=8, counts=aa.grid_point_counts(m, 0.15), reps=reps)


def c0_workload(aa):
    m = np.diag([25.83] * 3)
    x, y, z, q, sp = aa.synth_framework(m, 424, seed=1, stream=0)
    eps = np.array([t[0] for t in C0_TYPES] + [148.0])
    sig = np.array([t[1] for t in C0_TYPES] + [3.73])
    return dict(m=m, x=x, y=y, z=z, q=q, sp=sp, eps=eps, sig=sig, rc=12.8, counts=aa.grid_point_counts(m, 0.2))


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def c4_learned_vs_plain(aa, ctx, fw, w, sync, plain_steps=1_000_000, learned_max=600_000, seg=20_000):
    """Learned-density GCMC against plain uniform insertion on the 64 c3 chains (8 pressures x 8
    seeds), both from empty cells.  The plain method judges the loading: its 8-seed mean loading per
    pressure over the last segment of its run (plain_steps) with that mean's standard error sigma.
    A method's time is the GPU-seconds until its segment-mean loading first reaches
    (plain final - sigma) at every pressure; the learned time includes the density's construction
    (two LJ grid tables, the host-side draw of the 2e6-sample training set, MLP training, bin
    table)."""
    from aleator_amd import learned, workloads as wl

    chains = wl.c3_chains()
    nb = (64, 64, 64)

    def once(fn):
        sync()
        t0 = time.perf_counter()
        r = fn()
        sync()
        return r, time.perf_counter() - t0

    tabs, t_grid = once(lambda: fw.grid([4, 5], w["counts"], precision=aa.MIXED, grid_lattice=w["prim"]))
    (frac, e), t_sample = once(lambda: learned.sample_grid(tabs, [1.0, 2.0], w["counts"], 2_000_000,
                                                           low_fraction=0.5, low_cut=1000.0))
    ld = learned.LearnedDensity(ctx, max_rows=8192)
    losses, t_train = once(lambda: ld.fit(frac, e, e_clip=1000.0, n_iters=3000, batch=4096, lr=1e-3))
    wts, t_table = once(lambda: ld.bin_weights(nb, 298.0))
    t_density = t_grid + t_sample + t_train + t_table

    def trajectory(weights, max_steps, target=None, sigma=None, keep=False):
        g = aa.Gcmc(fw, wl.CO2_SITES, chains, max_molecules=1024)
        if weights is not None:
            g.set_proposal(weights, nb, w["reps"], 0.05)
        segs, total, att, acc = [], 0.0, 0, 0
        for _ in range(max_steps // seg):
            g.reset_stats()
            _, dt = once(lambda: g.run(seg))
            total += dt
            st = g.stats()
            n = np.array([s["sum_molecules"] / s["steps"] for s in st]).reshape(8, 8)
            att += sum(s["attempted"][2] for s in st)
            acc += sum(s["accepted"][2] for s in st)
            segs.append((total, n.mean(axis=1), n.std(axis=1, ddof=1) / np.sqrt(8)))
            if target is not None and np.all(segs[-1][1] >= target - sigma):
                break
        errors = sum(s["error"] for s in g.stats())
        if keep:
            return segs, acc / max(att, 1), errors, g
        g.close()
        return segs, acc / max(att, 1), errors

    plain, plain_acc, e1 = trajectory(None, plain_steps)
    target, sigma = plain[-1][1], plain[-1][2]
    t_plain = next(t for t, m, _ in plain if np.all(m >= target - sigma))
    lrn, lrn_acc, e2, g_filled = trajectory(wts, learned_max, target, sigma, keep=True)
    reached = bool(np.all(lrn[-1][1] >= target - sigma))
    # the c3 chains at that loading (the filled cells), continued with uniform insertion: the
    # per-step cost of the c3 workload at a representative occupancy
    g_filled.set_proposal(None, nb, w["reps"])
    g_filled.reset_stats()
    _, dt_filled = once(lambda: g_filled.run(5000))
    st_f = g_filled.stats()
    c3_filled = {"c3_filled_chain_steps_per_s": len(chains) * 5000 / dt_filled,
                 "c3_filled_mean_N": float(np.mean([s_["sum_molecules"] / s_["steps"] for s_ in st_f])),
                 "c3_filled_us_per_step": dt_filled / 5000 * 1e6}
    g_filled.close()
    t_learned = lrn[-1][0] + t_density if reached else None
    return {
        "c4_density": {"train_samples": 2_000_000, "adam_steps": 3000, "batch": 4096,
                       "final_loss_std_units": float(losses[-100:].mean()), "bins": list(nb),
                       "s_grid_sample_train_table": [t_grid, t_sample, t_train, t_table]},
        "c4_plain_steps": len(plain) * seg, "c4_plain_gpu_s": plain[-1][0],
        "c4_plain_final_loading": target.round(3).tolist(), "c4_plain_final_sigma": sigma.round(3).tolist(),
        "c4_plain_gpu_s_to_loading": t_plain,
        "c4_learned_steps": len(lrn) * seg, "c4_learned_loading": lrn[-1][1].round(3).tolist(),
        "c4_learned_gpu_s_to_loading_incl_training": t_learned,
        "c4_speedup_gpu_s": (t_plain / t_learned) if t_learned else None,
        "c4_insertion_acceptance": {"plain": plain_acc, "learned": lrn_acc},
        "c4_errors": e1 + e2,
        **c3_filled,
    }


def c4_equilibrium(aa, ctx, sync, seeds=32, fill_steps=10_000, seg=250, prod_steps=1_600_000):
    """Learned density against plain uniform insertion at equilibrium, on a framework where plain
    GCMC equilibrates within the budget: the c0 framework (cubic 25.83 A, 424 atoms), CH4 at 298 K
    at the 8 c3 pressures (0.1-30 bar, Peng-Robinson), `seeds` chains each, both from empty.
    Per method: the filling phase in segments, then a production block.  Reported per pressure:
    plain's loading N_p +- sigma_p (standard error over seeds), the learned N_l +- sigma_l, steps to
    90 % of N_p, and the GPU-seconds the learned method needs for plain's precision:
    t_density + t_learned (sigma_l/sigma_p)^2 (the standard error falls as 1/sqrt(steps)).  The
    density's cost (grid, 2e6-sample training set drawn on the host, MLP training, bin table) is
    counted in full."""
    from aleator_amd import learned, workloads as wl
    w = wl.c0()
    fw = aa.Framework(ctx, w["m"], w["x"], w["y"], w["z"], w["sp"], w["eps"], w["sig"], w["rc"])
    pres = [0.1 * 300.0 ** (i / 7.0) for i in range(8)]
    chains = []
    for i, p in enumerate(pres):
        f = aa.peng_robinson_fugacity(190.6, 4.599e6, 0.011, 298.0, p * 1e5)
        chains += [(f, 298.0, 500 + s, 64 * i + s) for s in range(seeds)]
    nb = (64, 64, 64)
    sync()
    t0 = time.perf_counter()
    tab = fw.grid([4], w["counts"], precision=aa.MIXED)
    frac, e = learned.sample_grid(tab, [1.0], w["counts"], 2_000_000, low_fraction=0.5, low_cut=1000.0)
    ld = learned.LearnedDensity(ctx, max_rows=8192)
    ld.fit(frac, e, e_clip=1000.0, n_iters=3000, batch=4096, lr=1e-3)
    weights = ld.bin_weights(nb, 298.0)
    sync()
    t_density = time.perf_counter() - t0

    def run(wts):
        g = aa.Gcmc(fw, wl.CH4_SITES, chains, max_molecules=512)
        if wts is not None:
            g.set_proposal(wts, nb, (1, 1, 1), 0.05)
        fill = []
        sync()
        t0 = time.perf_counter()
        for _ in range(fill_steps // seg):
            g.reset_stats()
            g.run(seg)
            st = g.stats()
            fill.append(np.array([s["sum_molecules"] / s["steps"] for s in st]).reshape(8, seeds).mean(axis=1))
        g.reset_stats()
        g.run(prod_steps)
        sync()
        dt = time.perf_counter() - t0
        st = g.stats()
        prod = np.array([s["sum_molecules"] / s["steps"] for s in st]).reshape(8, seeds)
        acc = sum(s["accepted"][2] for s in st) / max(1, sum(s["attempted"][2] for s in st))
        err = sum(s["error"] for s in st)
        g.close()
        return np.array(fill), prod, acc, dt, err

    fill_p, prod_p, acc_p, t_p, e1 = run(None)
    fill_l, prod_l, acc_l, t_l, e2 = run(weights)
    n_p, n_l = prod_p.mean(axis=1), prod_l.mean(axis=1)
    s_p = prod_p.std(axis=1, ddof=1) / np.sqrt(seeds)
    s_l = prod_l.std(axis=1, ddof=1) / np.sqrt(seeds)

    def steps_to(fill, target):
        hit = np.flatnonzero(fill >= target)
        return int((hit[0] + 1) * seg) if hit.size else None

    big = [i for i in range(8) if n_p[i] >= 0.5]
    t_equal = [t_density + t_l * (s_l[i] / s_p[i]) ** 2 for i in big]
    return {
        "c4eq_system": "c0 framework, CH4, 298 K, 8 pressures 0.1-30 bar x 32 seeds, from empty",
        "c4eq_pressures_bar": [round(p, 4) for p in pres],
        "c4eq_plain_loading": n_p.round(4).tolist(), "c4eq_plain_sigma": s_p.round(4).tolist(),
        "c4eq_learned_loading": n_l.round(4).tolist(), "c4eq_learned_sigma": s_l.round(4).tolist(),
        "c4eq_agree_within_2sigma": bool(np.all(np.abs(n_l - n_p) <= 2 * np.hypot(s_p, s_l))),
        "c4eq_steps_to_90pct_plain": [steps_to(fill_p[:, i], 0.9 * n_p[i]) for i in big],
        "c4eq_steps_to_90pct_learned": [steps_to(fill_l[:, i], 0.9 * n_p[i]) for i in big],
        "c4eq_insertion_acceptance": {"plain": acc_p, "learned": acc_l},
        "c4eq_gpu_s_plain": t_p, "c4eq_gpu_s_learned_run": t_l, "c4eq_gpu_s_density": t_density,
        "c4eq_gpu_s_learned_for_plain_precision": [round(v, 4) for v in t_equal],
        "c4eq_gain_gpu_s_min_over_pressures": min(t_p / v for v in t_equal),
        "c4eq_variance_ratio_equal_steps": [round(float((s_p[i] / s_l[i]) ** 2), 2) for i in big],
        "c4eq_steps_per_method": fill_steps + prod_steps, "c4eq_errors": e1 + e2,
    }


def measured_traffic(tile_shape):
    """HBM bytes per launch of the headline tile kernel with the tile shape this run uses, from the
    committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE runs scaled by a device copy of
    known size); None when the summary has no figure for that shape."""
    import re
    try:
        with open(os.path.join(ROOT, "profiles", "grid_tile_kernel_c1_co2_3tables.md")) as f:
            lines = re.findall(r"traffic_bytes_per_launch: (\d+) \(tile shape (\d+)\)", f.read())
    except OSError:
        return None
    by_shape = {int(k): float(v) for v, k in lines}
    return by_shape.get(tile_shape)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline_leg(w, c0, tables_mixed, tables_fp64, threads, seed=3):
    """The CPU baseline and the device's error against it, on the oracle restatement of the
    reference's FP64 path (oracle/liboracle.so; the reference's all-pairs loops, 125-candidate
    triclinic minimum image, per-pair mixing).  Only this leg imports the oracle.

    - c1 grid: 1200 random nodes (seed 3) of the 131^3 grid, all 3 tables, on one thread, and 1000
      nodes per thread over disjoint node ranges on `threads` host threads (the calls release the
      interpreter lock).
    - The device's MIXED and FP64 tables against the oracle on those nodes plus, per table, up to
      100 nodes of |E| in (1, 30] K (the hard ones: large terms cancelling), picked from the device
      FP64 table: max relative error over |E| > 1 K, capped nodes left out.
    - Widom: c0 CH4 (2e4 trials per thread) and c1 CO2 with Ewald (100 trials per thread), one
      thread and `threads` threads, each thread on its own Philox stream."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    counts = np.array(w["counts"], dtype=np.int32)
    total = int(np.prod(counts))
    rng = np.random.default_rng(seed)
    fw = ol.Framework(np.ascontiguousarray(w["sup"]), w["x"], w["y"], w["z"], w["q"],
                      np.ascontiguousarray(w["sp"], dtype=np.uint32))
    lj = ol.LJ(w["eps"], w["sig"], w["rc"])
    st = ol.EwaldState(fw, w["alpha"], w["kmax"])
    prim = np.ascontiguousarray(w["prim"])

    def grid_nodes(nodes):
        lj4 = ol.grid_nodes_lj(lj, fw, prim, counts, 4, nodes)
        lj5 = ol.grid_nodes_lj(lj, fw, prim, counts, 5, nodes)
        real, rec = ol.coulomb_nodes(st, w["rc"], fw, prim, counts, nodes)
        return lj4, lj5, real, rec

    def timed_nodes(per_thread, nt):
        nodes = np.sort(rng.choice(total, size=per_thread * nt, replace=False)).astype(np.int64)
        res = [None] * nt
        pool = [threading.Thread(target=lambda k=k: res.__setitem__(
            k, grid_nodes(nodes[k * per_thread:(k + 1) * per_thread]))) for k in range(nt)]
        t0 = time.perf_counter()
        for t in pool:
            t.start()
        for t in pool:
            t.join()
        dt = time.perf_counter() - t0
        vals = [np.concatenate([r[i] for r in res]) for i in range(4)]
        return nodes, vals, dt

    out = {}
    n1, v1, dt1 = timed_nodes(1200, 1)
    one = {"value": 3 * len(n1) / dt1, "unit": "grid pts/s", "cores": 1, "kind": "port",
           "sample": f"{len(n1)} random nodes (seed {seed}) of the 131^3 grid, all 3 tables, 1 host thread; "
                     f"{dt1:.1f} s; CPU {cpu_model()}"}
    nm, vm, dtm = timed_nodes(1000, threads)
    cpu = {"value": 3 * len(nm) / dtm, "unit": "grid pts/s", "cores": threads, "kind": "port",
           "sample": f"{len(nm)} random nodes (seed {seed}) of the 131^3 grid, all 3 tables, {threads} host "
                     f"threads over disjoint node ranges; {dtm:.1f} s; CPU {cpu_model()}, "
                     f"{os.cpu_count()} logical cores on host"}
    out["cpu_baseline_1_thread"] = one

    # the device against the oracle: the timed nodes plus hard nodes of each table
    hard = []
    for t in range(3):
        flat = tables_fp64[t]
        cand = np.flatnonzero((np.abs(flat) > 1.0) & (np.abs(flat) <= 30.0))
        if cand.size:
            hard.append(rng.choice(cand, size=min(100, cand.size), replace=False))
    extra_nodes = (np.setdiff1d(np.unique(np.concatenate(hard)), np.concatenate([n1, nm])).astype(np.int64)
                   if hard else np.zeros(0, dtype=np.int64))
    ve = grid_nodes(extra_nodes)
    nodes = np.concatenate([n1, nm, extra_nodes])
    refs = [np.concatenate([v1[i], vm[i], ve[i]]) for i in range(4)]
    phi = refs[2] + refs[3]
    errs = {}
    for name, tabs in (("mixed", tables_mixed), ("fp64", tables_fp64)):
        worst = []
        for t, ref in enumerate((refs[0], refs[1], phi)):
            got = tabs[t][nodes]
            sel = (np.abs(ref) > 1.0) & (ref < 1e5 if t < 2 else np.isfinite(ref))
            worst.append(float(np.max(np.abs(got[sel] - ref[sel]) / np.abs(ref[sel]))))
        errs[name] = worst
    out["c1_max_rel_err_vs_oracle"] = {
        "mixed": errs["mixed"], "fp64": errs["fp64"], "tables": ["C LJ", "O LJ", "Coulomb"],
        "nodes": int(len(nodes)), "hard_nodes": int(len(extra_nodes)),
        "rule": "max |dev - oracle| / |oracle| over nodes with |E| > 1 K, LJ nodes at the 1e5 K cap left out"}

    # Widom on the host: the oracle's widomInsertionHenryCoefficient restatement
    cases = {
        "c0_ch4": (ol.Framework(c0["m"], c0["x"], c0["y"], c0["z"], np.zeros(len(c0["x"])),
                                np.ascontiguousarray(c0["sp"], dtype=np.uint32)),
                   ol.LJ(c0["eps"], c0["sig"], c0["rc"]), [(0.0, 0.0, 0.0, 4, 0.0)], None, 20000),
        "c1_co2": (fw, lj, [(0.0, 0.0, -1.16, 5, -0.35), (0.0, 0.0, 0.0, 4, 0.70), (0.0, 0.0, 1.16, 5, -0.35)],
                   st, 100),
    }
    for name, (f_, l_, sites, ew, n) in cases.items():
        for nt in (1, threads):
            pool = [threading.Thread(target=lambda k=k: ol.widom(l_, f_, sites, 1, k + 1, n, 298.0, ewald=ew,
                                                                 rc_real=w["rc"] if ew else 0.0))
                    for k in range(nt)]
            t0 = time.perf_counter()
            for t in pool:
                t.start()
            for t in pool:
                t.join()
            out[f"cpu_oracle_widom_{name}_{nt}_thread_inserts_per_s"] = nt * n / (time.perf_counter() - t0)
    return cpu, out


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


LEGS = ("c1", "c0", "c2", "c3", "ewald", "mlp", "irmof1", "gcmc_grid", "guest_cells", "c4", "c4eq")


def guest_cell_rates(aa, ctx, sync, n_chains=64, fugacity=1e7):
    """Guest cells against the full guest scan at high occupancy: 3-site CO2-shaped LJ molecules
    (no charges) in a sparse synthetic framework (424 atoms in a 51.66 A cube, rc 12.8 A: 4 cells
    per axis), 64 chains at 1e7 Pa and 298 K filled by 3e4 steps to several hundred molecules,
    then 5000 steps timed with cells (mode 1: from 1024 guest sites on) and 5000 more with the
    full scan (the default)."""
    from aleator_amd import workloads as wl
    m = np.diag([51.66] * 3)
    x, y, z, _, sp = wl.synth_framework(m, 424, seed=1, stream=0)
    eps = np.array([t[0] for t in wl.C0_TYPES] + [wl.CO2_C[0], wl.CO2_O[0]])
    sig = np.array([t[1] for t in wl.C0_TYPES] + [wl.CO2_C[1], wl.CO2_O[1]])
    fw = aa.Framework(ctx, m, x, y, z, sp, eps, sig, 12.8)
    sites = [(x, y, z, sp, 0.0) for x, y, z, sp, _ in wl.CO2_SITES]
    gc = aa.Gcmc(fw, sites, [(fugacity, 298.0, 11, c) for c in range(n_chains)], max_molecules=1365)
    gc.set_guest_cells(1)
    out = {"guest_cells_per_axis": list(gc.guest_cells())}
    gc.run(30000)
    sync()
    for mode, name in ((1, "cells"), (0, "full_scan")):
        gc.set_guest_cells(mode)
        gc.reset_stats()
        t0 = time.perf_counter()
        gc.run(5000)
        sync()
        dt = time.perf_counter() - t0
        st = gc.stats()
        out[f"guest_cells_{name}_chain_steps_per_s"] = n_chains * 5000 / dt
        out[f"guest_cells_{name}_mean_N"] = float(np.mean([x["sum_molecules"] / x["steps"] for x in st]))
        out[f"guest_cells_{name}_errors"] = sum(x["error"] for x in st)
    gc.close()
    fw.close()
    return out


def ewald_proposal_rate(aa, ctx, fw, sync, n_prop=100_000):
    """EwaldIncrementalState proposals against the c1 framework's S(k) (half set 1102 k), batched
    (16 insertions or 8 moves per workgroup over the 183 runs of consecutive n2 in the k-set):
    CO2 insertions (3 sites) and moves (3 removed + 3 added).  Algorithmic work per proposal:
    sites x (Nk/2) x 16 FLOP (FP64); S(k) state read per proposal: (Nk/2) x 24 B (weight 8 B,
    S 16 B), through L1/L2."""
    rng = np.random.default_rng(11)
    nk = fw.num_k // 2
    st = aa.EwaldState(fw)
    co2 = np.array([(0.0, 0.0, -1.16, -0.35), (0.0, 0.0, 0.0, 0.70), (0.0, 0.0, 1.16, -0.35)])

    def molecules(n):
        c = rng.random((n, 1, 3)) @ np.asarray(fw.lattice)
        out = np.repeat(co2[None], n, axis=0)
        out[:, :, :3] += c
        return out

    res = {}
    for name, removed, added in (("insert", np.zeros((n_prop, 0, 4)), molecules(n_prop)),
                                 ("move", molecules(n_prop), molecules(n_prop))):
        st.propose_batch(removed[:1000], added[:1000])  # warm
        ctx.timing(True)
        for _ in range(3):
            st.propose_batch(removed, added)
        ms, n = ctx.timing(False)
        t = ms / max(n, 1) / 1e3
        sites = removed.shape[1] + added.shape[1]
        res[f"ewald_{name}_proposals_per_s"] = n_prop / t
        res[f"ewald_{name}_fp64_tflops"] = n_prop * sites * nk * 16 / t / 1e12
        res[f"ewald_{name}_state_read_gb_per_s"] = n_prop * nk * 24 / t / 1e9
    st.close()
    return res


def irmof1_gcmc_rates(aa, ctx, sync):
    """GCMC on the reference's real IRMOF-1 systems (CIFs under tests/golden/cif/), device chains:
      methane, UFF, rc 12 A shifted, 298 K, 0.1 and 10 bar (Peng-Robinson), 424 atoms;
      CO2, DDEC charges, 2x2x2 (848 atoms), Ewald alpha 0.25 kmax 6, rc 12.8 A, 1e6 Pa, the
      reference's high-occupancy benchmark point (bench_gcmc_co2_irmof1.cc:116-160).
    Per-step latency of one chain, aggregate chain-steps/s of many, and the reference CLI's
    <N> (seed 2026, 30000 + 90000 steps at 0.1 bar) computed on the device."""
    cif = os.path.join(ROOT, "tests", "golden", "cif")
    uff = {"Zn": (62.3992, 2.46155), "O": (48.1581, 3.03315), "C": (47.8562, 3.47299), "H": (7.64893, 2.84642)}
    out = {}

    def timed_run(gc, steps):
        sync()
        t0 = time.perf_counter()
        gc.run(steps)
        sync()
        return time.perf_counter() - t0

    s = aa.read_cif(os.path.join(cif, "IRMOF-1.cif"))
    eps = [uff[e][0] for e in s.species_symbols] + [158.5]
    sig = [uff[e][1] for e in s.species_symbols] + [3.72]
    fw = aa.Framework(ctx, s.lattice, s.x, s.y, s.z, s.species, eps, sig, 12.0, truncation=aa.SHIFTED)
    ch4 = [(0.0, 0.0, 0.0, len(s.species_symbols), 0.0)]
    for bar in (0.1, 10.0):
        f = aa.peng_robinson_fugacity(190.6, 4.599e6, 0.011, 298.0, bar * 1e5)
        for n in (1, 256):
            gc = aa.Gcmc(fw, ch4, [(f, 298.0, 2026, c) for c in range(n)], max_molecules=1024)
            gc.run(2000)
            dt = timed_run(gc, 20000)
            key = f"irmof1_ch4_{bar:g}bar_{n}chains"
            out[key + "_us_per_step"] = dt / 20000 * 1e6
            out[key + "_chain_steps_per_s"] = n * 20000 / dt
            gc.close()
    f = aa.peng_robinson_fugacity(190.6, 4.599e6, 0.011, 298.0, 0.1e5)
    gc = aa.Gcmc(fw, ch4, [(f, 298.0, 2026, 0)], max_molecules=1024)
    gc.run(30000)
    gc.reset_stats()
    gc.run(90000)
    st = gc.stats()[0]
    out["irmof1_ch4_cli_run_mean_N"] = st["sum_molecules"] / st["steps"]  # the reference CLI prints 0.231456
    gc.close()

    d = aa.read_cif(os.path.join(cif, "IRMOF-1_primitive_DDEC.cif"))
    tot = 0.0
    for v in d.charge:
        tot += float(v)
    q = d.charge - tot / len(d.charge)
    lat, x, y, z, qq, sp = aa.replicate_supercell(d.lattice, d.x, d.y, d.z, q, d.species, (2, 2, 2))
    ns = len(d.species_symbols)
    eps = [uff[e][0] for e in d.species_symbols] + [27.0, 79.0]
    sig = [uff[e][1] for e in d.species_symbols] + [2.8, 3.05]
    fw2 = aa.Framework(ctx, lat, x, y, z, sp, eps, sig, 12.8, truncation=aa.SHIFTED, charge=qq, ewald=(0.25, 12.8, 6))
    co2 = [(0.0, 0.0, -1.16, ns + 1, -0.35), (0.0, 0.0, 0.0, ns, 0.70), (0.0, 0.0, 1.16, ns + 1, -0.35)]
    f = aa.peng_robinson_fugacity(304.1282, 7377300.0, 0.22394, 298.0, 1.0e6)
    for n in (1, 64):
        gc = aa.Gcmc(fw2, co2, [(f, 298.0, 2026, c) for c in range(n)], max_molecules=1024)
        gc.run(2000)
        dt = timed_run(gc, 2000)
        key = f"irmof1_co2_1e6Pa_{n}chains"
        out[key + "_us_per_step"] = dt / 2000 * 1e6
        out[key + "_chain_steps_per_s"] = n * 2000 / dt
        out[key + "_molecules_after_4000_steps"] = float(np.mean([c["n_molecules"] for c in gc.stats()]))
        gc.close()

    return out


def gcmc_grid_rates(aa, ctx, sync):
    """GCMC with the framework LJ taken from an energy grid, in the reference's own measurement
    (CH4 in IRMOF-1, UFF, rc 12 A shifted, 298 K, 0.1 bar Peng-Robinson, seed 7, 0.2 A grid; the
    single cell, 129^3 nodes, and the 2x2x2 supercell, 258^3 nodes): the device build of the
    FP64 table and the device per-step time of 1 and 256 chains with and without the grid."""
    cif = os.path.join(ROOT, "tests", "golden", "cif")
    uff = {"Zn": (62.3992, 2.46155), "O": (48.1581, 3.03315), "C": (47.8562, 3.47299), "H": (7.64893, 2.84642)}
    s = aa.read_cif(os.path.join(cif, "IRMOF-1.cif"))
    eps = [uff[e][0] for e in s.species_symbols] + [158.5]
    sig = [uff[e][1] for e in s.species_symbols] + [3.72]
    ch4 = len(s.species_symbols)
    sites = [(0.0, 0.0, 0.0, ch4, 0.0)]
    f = aa.peng_robinson_fugacity(190.6, 4.599e6, 0.011, 298.0, 0.1e5)
    out = {}

    def timed_run(gc, steps):
        sync()
        t0 = time.perf_counter()
        gc.run(steps)
        sync()
        return time.perf_counter() - t0

    for label, reps in (("1x1x1", (1, 1, 1)), ("2x2x2", (2, 2, 2))):
        lat, x, y, z, q, sp = aa.replicate_supercell(s.lattice, s.x, s.y, s.z, s.charge, s.species, reps)
        fw = aa.Framework(ctx, lat, x, y, z, sp, eps, sig, 12.0, truncation=aa.SHIFTED)
        n = aa.grid_point_counts(lat, 0.2)
        nodes = n[0] * n[1] * n[2]
        buf = aa.DeviceBuffer(nodes)
        fw.grid_device([ch4], n, int(buf))  # warm (tile choice, code load)
        sync()
        t0 = time.perf_counter()
        fw.grid_device([ch4], n, int(buf))
        sync()
        key = f"gcmc_grid_irmof1_{label}"
        out[key + "_grid_nodes"] = nodes
        out[key + "_device_grid_build_s"] = time.perf_counter() - t0
        for nch in (1, 256):
            for mode in ("direct", "grid"):
                gc = aa.Gcmc(fw, sites, [(f, 298.0, 7, c) for c in range(nch)], max_molecules=1024)
                if mode == "grid":
                    gc.set_energy_grid(n, [ch4], buf)
                gc.run(2000)
                dt = timed_run(gc, 20000)
                out[f"{key}_{mode}_{nch}chains_us_per_step"] = dt / 20000 * 1e6
                out[f"{key}_{mode}_{nch}chains_chain_steps_per_s"] = nch * 20000 / dt
                gc.close()
        buf.free()
        fw.close()
    return out


def mlp_training_rate(aa, ctx, sync, batch=4096, iters=500):
    """Adam steps of the insertion-density MLP (64 Fourier features, 4 x 256 SiLU, 1 output) on a
    smooth periodic target over 2e5 resident rows: forward + backward + Adam per step, replayed
    from its hipGraph.  FLOP per step = 3 x 2 x (weights of the 5 layers) x batch."""
    import torch

    rng = np.random.default_rng(7)
    frac = rng.random((200_000, 3))
    target = np.sin(2 * np.pi * frac[:, 0]) * np.cos(2 * np.pi * frac[:, 1]) + 0.5 * np.cos(4 * np.pi * frac[:, 2])
    d_f = torch.from_numpy(frac).cuda()
    d_t = torch.from_numpy(target).cuda()
    net = aa.Mlp(ctx, max_rows=batch, seed=1)
    net.train_device(d_f.data_ptr(), d_t.data_ptr(), len(frac), batch, 20, 1e-3, seed=1, stream=0)  # capture + warm
    sync()
    t0 = time.perf_counter()
    losses = net.train_device(d_f.data_ptr(), d_t.data_ptr(), len(frac), batch, iters, 1e-3, seed=2, stream=0)
    sync()
    dt = (time.perf_counter() - t0) / iters
    weights = 64 * 256 + 3 * 256 * 256 + 256
    flop = 3 * 2 * weights * batch
    return {"mlp_ms_per_adam_step_batch4096": dt * 1e3, "mlp_tflops": flop / dt / 1e12,
            "mlp_final_loss": float(np.mean(losses[-20:]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="only the headline build (profiling runs)")
    ap.add_argument("--no-graph", action="store_true", help="launch the headline plan without its hipGraph")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--c4-plain-steps", type=int, default=1_000_000)
    ap.add_argument("--legs", default="all",
                    help="comma-separated extra legs to run after the headline: " + ",".join(LEGS) + " (or all)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    args = ap.parse_args()
    legs = set() if args.headline_only else (set(LEGS) if args.legs == "all" else set(args.legs.split(",")))
    if legs - set(LEGS):
        ap.error(f"unknown legs {sorted(legs - set(LEGS))}")
    if args.no_c4:
        legs.discard("c4")
        legs.discard("c4eq")
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the hot path would shard by grid slab, trial range or chain, with no exchange "
              "beyond a final sum)", file=sys.stderr)
        sys.exit(2)

    import torch
    import aleator_amd as aa

    prec = aa.MIXED if args.precision == "mixed" else aa.FP64
    ctx = aa.Context(0)
    sync = torch.cuda.synchronize

    # ---------------- headline: three c1 tables
    w = c1_workload(aa)
    fw = aa.Framework(ctx, w["sup"], w["x"], w["y"], w["z"], w["sp"], w["eps"], w["sig"], w["rc"], charge=w["q"],
                      ewald=(w["alpha"], w["rc"], w["kmax"]))
    nx, ny, nz = w["counts"]
    nodes = nx * ny * nz
    out = torch.empty(3 * nodes, dtype=torch.float64, device="cuda")
    # the build is planned once (constants uploaded, tile shape autotuned, launches captured in a
    # hipGraph); a step replays it: device-resident input to device-resident tables
    plan = aa.GridPlan(ctx, [fw], [4, 5], [w["counts"]], [out.data_ptr()], coulomb=True, precision=prec,
                       grid_lattices=[w["prim"]], tile=-1, use_graph=not args.no_graph)
    tile, tuning = plan.tile()
    step = plan.execute
    for _ in range(args.warmup):
        step()
    sync()
    ctx.timing(True)
    t_step = timed(step, args.steps, 0, sync)
    k_ms, k_n = ctx.timing(False)
    k_avg = k_ms / max(k_n, 1) / 1e3
    rho = len(w["x"]) / abs(np.linalg.det(w["sup"]))
    nbar = rho * 4.0 / 3.0 * np.pi * w["rc"] ** 3
    flop = nodes * nbar * (2 * F_LJ + F_REAL)
    achieved = flop / k_avg / 1e12
    value = 3 * nodes / t_step

    tables_main = out.cpu().numpy().reshape(3, nodes)
    # the erf kernel polynomial degree the build used (the tile kernel's template argument)
    erf_deg = plan.erf_degree()[0]

    # ---------------- extras
    extra = {"nodes_per_s": nodes / t_step, "tile_kernel_ms": k_avg * 1e3, "pairs_per_node": nbar,
             "tile_shape": tile, "tile_atoms_per_pass": plan.atoms_per_pass(), "autotune_ms_per_shape": tuning,
             "erf_degree_inner_shell": list(plan.erf_degree())}
    # the same grid in the other precision (FP64 beside the MIXED headline), same plan machinery;
    # profiling runs (--headline-only) leave it out so that their counters cover the headline alone
    tables_mixed = tables_fp64 = None
    if not args.headline_only:
        other = aa.FP64 if prec == aa.MIXED else aa.MIXED
        other_name = "fp64" if prec == aa.MIXED else "mixed"
        out2 = torch.empty(3 * nodes, dtype=torch.float64, device="cuda")
        plan_o = aa.GridPlan(ctx, [fw], [4, 5], [w["counts"]], [out2.data_ptr()], coulomb=True, precision=other,
                             grid_lattices=[w["prim"]], tile=-1, use_graph=not args.no_graph)
        ctx.timing(True)
        t_o = timed(plan_o.execute, max(3, args.steps // 2), 1, sync)
        ko_ms, ko_n = ctx.timing(False)
        extra[f"c1_grid_{other_name}_pts_per_s"] = 3 * nodes / t_o
        extra[f"c1_grid_{other_name}_tile_kernel_ms"] = ko_ms / max(ko_n, 1)
        extra[f"c1_grid_{other_name}_tile_shape"] = plan_o.tile()[0]
        tables_other = out2.cpu().numpy().reshape(3, nodes)
        plan_o.close()
        del out2
        tables_mixed, tables_fp64 = (tables_main, tables_other) if prec == aa.MIXED else (tables_other, tables_main)
    if args.copy_calibration:
        # a plain device copy of known size, for scaling HBM byte counters (FETCH_SIZE, WRITE_SIZE)
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        extra["copy_calibration_bytes_each_way"] = src.numel() * 4
    if "c1" in legs:
        t = timed(lambda: fw.grid_device([4, 5], w["counts"], out.data_ptr(), coulomb=True, precision=prec,
                                         grid_lattice=w["prim"]), 5, 1, sync)
        extra["c1_grid_per_call_entry_point_pts_per_s"] = 3 * nodes / t
    if "c0" in legs:
        c0 = c0_workload(aa)
        fw0 = aa.Framework(ctx, c0["m"], c0["x"], c0["y"], c0["z"], c0["sp"], c0["eps"], c0["sig"], c0["rc"])
        n0 = int(np.prod(c0["counts"]))
        out0 = torch.empty(n0, dtype=torch.float64, device="cuda")
        for p, name in ((aa.MIXED, "mixed"), (aa.FP64, "fp64")):
            t = timed(lambda: fw0.grid_device([4], c0["counts"], out0.data_ptr(), precision=p), 10, 2, sync)
            extra[f"c0_ch4_grid_129^3_{name}_pts_per_s"] = n0 / t
        du = torch.empty(10_000_000, dtype=torch.float64, device="cuda")
        sm = torch.empty(1, dtype=torch.float64, device="cuda")
        ch4 = [(0.0, 0.0, 0.0, 4, 0.0)]
        for n_trials in (100_000, 1_000_000, 10_000_000):
            for p, name in ((aa.MIXED, "mixed"), (aa.FP64, "fp64")):
                t = timed(lambda: fw0.widom_device(ch4, n_trials, 298.0, du.data_ptr(), sm.data_ptr(), seed=1, stream=1,
                                                   precision=p), 5, 1, sync)
                extra[f"c0_widom_ch4_{n_trials}_{name}_inserts_per_s"] = n_trials / t
        # algorithmic work per insertion: CH4 on c0, n = 216.1 atoms within rc x 20 FLOP (LJ)
        rate = extra["c0_widom_ch4_10000000_mixed_inserts_per_s"]
        extra["c0_widom_ch4_1e7_mixed_tflops"] = rate * 216.1 * F_LJ / 1e12
        co2 = [(0.0, 0.0, -1.16, 5, -0.35), (0.0, 0.0, 0.0, 4, 0.70), (0.0, 0.0, 1.16, 5, -0.35)]
        # CO2 on c1: 3 sites x (444.3 x (20 + 30) FLOP direct + 1102 half-set k x 16 FLOP FP64)
        f_ins = 3 * (nbar * (F_LJ + F_REAL) + (fw.num_k // 2) * 16)
        for n_trials in (100_000, 1_000_000):
            t = timed(lambda: fw.widom_device(co2, n_trials, 298.0, du.data_ptr(), sm.data_ptr(), seed=1, stream=1,
                                              precision=prec), 3, 1, sync)
            extra[f"c1_widom_co2_{n_trials}_{args.precision}_inserts_per_s"] = n_trials / t
            extra[f"c1_widom_co2_{n_trials}_{args.precision}_tflops"] = n_trials / t * f_ins / 1e12
        del du

    from aleator_amd import workloads as wl

    if "c2" in legs:
        # c2: 256 frameworks, one batched (graph-captured, autotuned) launch per step, MIXED pair terms
        try:
            fws2, eps2, sig2 = wl.c2(256)
            devs = [aa.Framework(ctx, f["sup"], f["x"], f["y"], f["z"], f["sp"], eps2, sig2, f["rc"]) for f in fws2]
            counts2 = [f["counts"] for f in fws2]
            nodes2 = [int(np.prod(c)) for c in counts2]
            buf2 = torch.empty(sum(nodes2), dtype=torch.float64, device="cuda")
            offs = np.concatenate([[0], np.cumsum(nodes2)[:-1]])
            ptrs = [buf2.data_ptr() + 8 * int(o) for o in offs]
            plan2 = aa.GridPlan(ctx, devs, [4], counts2, ptrs, precision=aa.MIXED,
                                grid_lattices=[f["prim"] for f in fws2], tile=-1, use_graph=True)
            t = timed(plan2.execute, 3, 1, sync)
            extra["c2_batched_256_frameworks_pts_per_s"] = sum(nodes2) / t
            extra["c2_nodes"] = sum(nodes2)
            # algorithmic work: sum over frameworks of nodes x atoms within rc (density x sphere) x 20 FLOP
            work2 = sum(n_f * len(f["x"]) / abs(np.linalg.det(f["sup"])) * (4.0 / 3.0) * np.pi * f["rc"] ** 3 * F_LJ
                        for n_f, f in zip(nodes2, fws2))
            extra["c2_tflops"] = work2 / t / 1e12
            extra["c2_ms_per_batch"] = t * 1e3
            extra["c2_tile_shape"] = plan2.tile()[0]
            extra["c2_tile_atoms_per_pass"] = plan2.atoms_per_pass()
            plan2.close()
            del buf2, devs
        except Exception as e:  # report, do not hide the headline
            extra["c2_error"] = repr(e)

    if "c3" in legs:
        # c3: 64 CO2 GCMC chains on the charged framework, one persistent launch
        try:
            chains = wl.c3_chains()
            gc = aa.Gcmc(fw, wl.CO2_SITES, chains, max_molecules=1024)
            # 2e4 cycles per chain, a cycle taken as one move attempt (the reference's run() counts
            # steps); one persistent launch advances all 64 chains
            n_steps = 20000
            t0 = time.perf_counter()
            gc.run(n_steps)
            sync()
            dt = time.perf_counter() - t0
            st = gc.stats()
            extra["c3_chain_steps_per_s"] = len(chains) * n_steps / dt
            extra["c3_gpu_s_2e4_steps_all_chains"] = dt
            extra["c3_insertion_acceptance"] = (sum(s["accepted"][2] for s in st) /
                                                max(1, sum(s["attempted"][2] for s in st)))
            extra["c3_mean_N_by_pressure"] = [float(np.mean([st[8 * i + s]["sum_molecules"] / st[8 * i + s]["steps"]
                                                             for s in range(8)])) for i in range(8)]
            extra["c3_errors"] = sum(s["error"] for s in st)
            gc.close()
        except Exception as e:
            extra["c3_error"] = repr(e)

    if "ewald" in legs:
        try:
            extra.update(ewald_proposal_rate(aa, ctx, fw, sync))
        except Exception as e:
            extra["ewald_error"] = repr(e)

    if "mlp" in legs:
        try:
            extra.update(mlp_training_rate(aa, ctx, sync))
        except Exception as e:
            extra["mlp_error"] = repr(e)

    if "irmof1" in legs:
        try:
            extra.update(irmof1_gcmc_rates(aa, ctx, sync))
        except Exception as e:
            extra["irmof1_error"] = repr(e)

    if "gcmc_grid" in legs:
        try:
            extra.update(gcmc_grid_rates(aa, ctx, sync))
        except Exception as e:
            extra["gcmc_grid_error"] = repr(e)

    if "guest_cells" in legs:
        try:
            extra.update(guest_cell_rates(aa, ctx, sync))
        except Exception as e:
            extra["guest_cells_error"] = repr(e)

    if "c4eq" in legs:
        try:
            extra.update(c4_equilibrium(aa, ctx, sync))
        except Exception as e:
            extra["c4eq_error"] = repr(e)

    if "c4" in legs:
        # c4: learned insertion density vs plain uniform insertion on the c3 chains
        try:
            extra.update(c4_learned_vs_plain(aa, ctx, fw, w, sync, plain_steps=args.c4_plain_steps))
        except Exception as e:
            extra["c4_error"] = repr(e)

    cpu = None
    if not args.no_cpu_baseline and not args.headline_only:
        th = host_threads()
        cpu, cpu_extra = cpu_baseline_leg(w, c0_workload(aa), tables_mixed, tables_fp64, th)
        extra.update(cpu_extra)
        extra["speedup_vs_cpu_baseline_1_thread"] = value / extra["cpu_baseline_1_thread"]["value"]
        extra["speedup_vs_cpu_baseline"] = value / cpu["value"]
        if "c0" in legs:
            for n_trials in (100_000, 1_000_000, 10_000_000):
                extra[f"widom_c0_{n_trials}_speedup_vs_cpu_all_threads"] = (
                    extra[f"c0_widom_ch4_{n_trials}_mixed_inserts_per_s"] /
                    extra[f"cpu_oracle_widom_c0_ch4_{th}_thread_inserts_per_s"])
            extra["widom_c1_1000000_speedup_vs_cpu_all_threads"] = (
                extra[f"c1_widom_co2_1000000_{args.precision}_inserts_per_s"] /
                extra[f"cpu_oracle_widom_c1_co2_{th}_thread_inserts_per_s"])

    line = {
        "metric": "grid pts/s (3-table CO2 grid: C LJ, O LJ, Ewald Coulomb)",
        "value": value,
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (MIXED: hardware rsq + 1 Newton step, erf polynomial to 1e-12)" if prec == aa.MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": "charged triclinic rhombohedral 24 A 60/60/60, 600 atoms -> 2x2x2 supercell "
                               "(4800 atoms), CO2 probe, 3 tables at 0.15 A over the primitive cell (131^3), "
                               "Ewald alpha 0.265 kmax 8 (2204 k), rc 12 A",
                   "nodes": nodes, "tables": 3, "precision": args.precision},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(tile),
                     "kernel": "grid_tile_kernel<{}, 2, {}, 0, {}, {}, {}, {}>".format(
                         "true" if prec == aa.FP64 else "false", erf_deg, *TILE_SHAPES[tile]),
                     "flop_per_launch": flop,
                     "peak_note": "FP64 vector peak: the pair arithmetic is FP64 (FP32 peak would be 157.3)"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

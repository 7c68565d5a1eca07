"""Grid plans: one batched launch over many frameworks, hipGraph replay and tile autotuning must
give the tables of the per-call build (FrameworkEnergyGrid ctor, framework_energy_grid.cc:193-249).
A given tile shape is bitwise deterministic; different shapes agree to rounding (1e-12)."""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework, rel_err

pytestmark = pytest.mark.gpu


def frameworks(ctx, n=3):
    out = []
    rng = np.random.default_rng(3)
    eps = np.array([t[0] for t in ol.C0_TYPES] + [ol.METHANE[0]])
    sig = np.array([t[1] for t in ol.C0_TYPES] + [ol.METHANE[1]])
    for i in range(n):
        cell = (rng.uniform(10, 16), rng.uniform(10, 16), rng.uniform(10, 16),
                *(rng.uniform(70, 110, 3) if i % 2 else (90.0, 90.0, 90.0)))
        m = ol.cell_from_params(*cell)
        fw = ol.synth_framework(m, int(rng.integers(50, 200)), seed=10 + i, stream=0)
        reps = ol.min_supercell(m, 6.0)
        sc = ol.replicate(fw, reps)
        out.append((fw, sc, ol.LJ(eps, sig, 6.0), ol.grid_counts(m, 0.5)))
    return out


def test_batched_plan_equals_per_call_builds(ctx):
    systems = frameworks(ctx)
    devs = [device_framework(ctx, sc, lj) for fw, sc, lj, counts in systems]
    bufs = [aa.DeviceBuffer(int(np.prod(c))) for _, _, _, c in systems]
    for shape in range(4):
        plan = aa.GridPlan(ctx, devs, [4], [c for *_, c in systems], [int(b) for b in bufs], precision=aa.FP64,
                           grid_lattices=[fw.m for fw, *_ in systems], tile=shape)
        plan.execute()
        for (fw, sc, lj, counts), dev, buf in zip(systems, devs, bufs):
            got = buf.to_host().reshape(counts[::-1])
            ref = dev.grid([4], counts, precision=aa.FP64, grid_lattice=fw.m)[0]
            assert rel_err(got, ref).max() <= 1e-12
        plan.close()
    # and the oracle on a sample of one framework
    fw, sc, lj, counts = systems[1]
    nodes = np.arange(0, int(np.prod(counts)), 97, dtype=np.int64)
    ref = ol.grid_nodes_lj(lj, sc, fw.m, counts, 4, nodes)
    assert rel_err(bufs[1].to_host()[nodes], ref).max() <= 1e-10


def test_atoms_per_pass_on_a_single_image_framework(ctx):
    """The headline framework (c1 2x2x2 supercell, search radius below half its perpendicular
    width) lets the list build stage up to 1024 atoms per pass; every shape's tables agree with
    the per-call build and the oracle."""
    fw, lj = ol.c1_system()
    sc = ol.replicate(fw, ol.min_supercell(fw.m, 12.0))
    dev = device_framework(ctx, sc, lj)
    counts = ol.grid_counts(fw.m, 0.3)  # 65^3 nodes: every tile shape keeps 2 (rc + tile radius) < L_perp
    buf = aa.DeviceBuffer(int(np.prod(counts)))
    ref = dev.grid([4], counts, precision=aa.FP64, grid_lattice=fw.m)[0].ravel()
    for shape in range(4):
        plan = aa.GridPlan(ctx, [dev], [4], [counts], [int(buf)], precision=aa.FP64, grid_lattices=[fw.m], tile=shape)
        assert plan.atoms_per_pass() == 1024
        plan.execute()
        assert rel_err(buf.to_host(), ref).max() <= 1e-10  # FP64 bar; tile-relative rounding differs by shape
        plan.close()
    nodes = np.arange(0, int(np.prod(counts)), 997, dtype=np.int64)
    assert rel_err(ref[nodes], ol.grid_nodes_lj(lj, sc, fw.m, counts, 4, nodes)).max() <= 1e-10


def test_graph_replay_is_bitwise_and_repeatable(ctx):
    systems = frameworks(ctx, 2)
    devs = [device_framework(ctx, sc, lj) for fw, sc, lj, counts in systems]
    bufs = [aa.DeviceBuffer(int(np.prod(c))) for _, _, _, c in systems]
    eager = aa.GridPlan(ctx, devs, [4], [c for *_, c in systems], [int(b) for b in bufs], precision=aa.MIXED,
                        grid_lattices=[fw.m for fw, *_ in systems], tile=1)
    eager.execute()
    ref = [b.to_host() for b in bufs]
    graph = aa.GridPlan(ctx, devs, [4], [c for *_, c in systems], [int(b) for b in bufs], precision=aa.MIXED,
                        grid_lattices=[fw.m for fw, *_ in systems], tile=1, use_graph=True)
    for replay in range(3):
        for b in bufs:  # eager work between replays
            ctx_zero = np.zeros(b.nbytes // 8)
            aa._hip_rt().hipMemcpy(b.ptr, ctx_zero.ctypes.data_as(aa.C.c_void_p), aa.C.c_size_t(b.nbytes), 1)
        graph.execute()
        for b, r in zip(bufs, ref):
            assert np.array_equal(b.to_host(), r), f"replay {replay}"


def test_autotune_picks_a_fitting_shape(ctx):
    fw, lj = ol.c0_system()
    dev = device_framework(ctx, fw, lj)
    counts = ol.grid_counts(fw.m, 0.2)
    buf = aa.DeviceBuffer(int(np.prod(counts)))
    plan = aa.GridPlan(ctx, [dev], [4], [counts], [int(buf)], precision=aa.MIXED, tile=-1)
    shape, ms = plan.tile()
    per_pass = plan.atoms_per_pass()
    print(f"autotune c0 MIXED: chose {shape} with {per_pass} atoms per list pass, medians (ms) {ms}")
    assert 0 <= shape < 4 and ms[shape] > 0
    assert per_pass == 128  # c0's search radius admits two images of an atom: 8 images per atom, 128 per pass
    plan.execute()
    got = buf.to_host()
    ref = dev.grid([4], counts, precision=aa.MIXED)[0].ravel()
    assert rel_err(got, ref).max() <= 1e-6


def test_erf_degrees_of_the_shell_split_on_c1(ctx):
    """c1 (alpha 0.265, rc 12 A): in MIXED (tolerance 1e-12 on the dropped Chebyshev terms) the
    entries inside both cutoffs take the degree-18 erf polynomial and the shell entries one of
    degree 11 fitted over their own distances; FP64 (1e-17) takes 24 and 17.  Both plans agree with each other to the MIXED bar
    (1e-5 relative where |E| > 1 K) on every node, the Coulomb table included."""
    fw, lj = ol.c1_system()
    sc = ol.replicate(fw, ol.min_supercell(fw.m, 12.0))
    counts = ol.grid_counts(fw.m, 0.15)
    dev = device_framework(ctx, sc, lj, ewald=(0.265, 12.0, 8))
    n = int(np.prod(counts))
    tabs = {}
    for prec, want in ((aa.MIXED, (18, 11)), (aa.FP64, (24, 17))):
        buf = aa.DeviceBuffer(3 * n)
        plan = aa.GridPlan(ctx, [dev], [4, 5], [counts], [int(buf)], coulomb=True, precision=prec,
                           grid_lattices=[fw.m], tile=1)
        assert plan.erf_degree() == want
        plan.execute()
        tabs[prec] = buf.to_host().reshape(3, n)
        plan.close()
    for t in range(3):
        ref = tabs[aa.FP64][t]
        sel = (np.abs(ref) > 1.0) & (ref < 1e5)
        err = np.abs(tabs[aa.MIXED][t][sel] - ref[sel]) / np.abs(ref[sel])
        assert err.max() <= 1e-5, (t, err.max())

"""Native object lifetimes.  A garbage collector finalises the objects of an unreachable cycle in
no set order, so destroy calls can reach the library parent first.  A context and a framework
are reference counted: every object created on them keeps them alive, so a child used or
destroyed after its parent's handle was dropped still works."""
import ctypes as C

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol
from gpu_helpers import device_framework

pytestmark = pytest.mark.gpu


def _drop_handle(obj, destroy):
    """Destroy obj's native handle directly, bypassing the binding's children-first close."""
    h = obj._h
    obj._h = C.c_void_p()
    if hasattr(obj, "_owned"):
        obj._owned = []
    assert destroy(h) == 0


def test_children_outlive_dropped_parent_handles():
    print("lifetimes: context and framework handles destroyed before their chains, state and network",
          flush=True)
    lib = aa._load()
    ctx = aa.Context(0)
    fw, lj = ol.c0_system(40, a=14.0)
    lj.cutoff, lj.truncation = 6.5, ol.SHIFTED
    sites = [(0.0, 0.0, 0.0, 4, 0.0)]
    dev = device_framework(ctx, fw, lj)
    chain = (1e8, 300.0, 5, 0)
    gc = aa.Gcmc(dev, sites, [chain], max_molecules=128)
    net = aa.Mlp(ctx, max_rows=64, seed=3)
    frac = np.random.default_rng(1).random((50, 3))
    y_before = net.predict(frac)

    _drop_handle(ctx, lib.aleator_gpu_destroy)
    _drop_handle(dev, lib.aleator_gpu_framework_destroy)

    # the chains still read the framework's atoms and run on the context's stream
    trace = gc.run(800, trace=True)
    ref = ol.Gcmc(lj, fw, sites, *chain).run(800, trace=True)
    assert np.array_equal(trace[0], ref.astype(np.int32))
    assert np.array_equal(net.predict(frac), y_before)
    gc.close()  # the framework's last reference: frees it, then drops one on the context
    net.close()  # the context's last reference


def test_destroy_order_does_not_matter_for_plans_and_states():
    print("lifetimes: grid plan and S(k) state destroyed after their framework and context", flush=True)
    lib = aa._load()
    ctx = aa.Context(0)
    fw = ol.synth_framework(ol.cell_from_params(14.0, 15.0, 16.0, 90.0, 90.0, 90.0), 60, seed=2,
                            charge_sigma=0.5)
    lj = ol.LJ(np.array([t[0] for t in ol.C0_TYPES]), np.array([t[1] for t in ol.C0_TYPES]), 6.5)
    dev = device_framework(ctx, fw, lj, ewald=(0.3, 6.5, 4))
    counts = ol.grid_counts(fw.m, 0.9)
    out = aa.DeviceBuffer(int(np.prod(counts)))
    plan = aa.GridPlan(ctx, [dev], [0], [counts], [out], precision=aa.FP64)
    st = aa.EwaldState(dev)
    _drop_handle(ctx, lib.aleator_gpu_destroy)
    _drop_handle(dev, lib.aleator_gpu_framework_destroy)
    plan.execute()
    got = out.to_host().reshape(counts[::-1])
    ref = ol.grid_build(lj, fw, [0], counts)
    assert np.max(np.abs(got.ravel() - np.asarray(ref).ravel()) / np.maximum(np.abs(np.asarray(ref).ravel()), 1.0)) < 1e-10
    st.close()
    plan.close()
    out.free()

"""The C-ABI library (CPU-side checks): it loads, exports every function include/aleator_gpu.h
declares, and its host-only entry points agree with the oracle.  No GPU call is made."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol

HEADER = os.path.join(ol.ROOT, "include", "aleator_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aleator_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(aa.library_path())
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_grid_point_counts_match_reference_rule():
    for cell, spacing in (((25.83, 25.83, 25.83, 90, 90, 90), 0.2), ((24, 24, 24, 60, 60, 60), 0.15),
                          ((10.0, 37.0, 22.0, 70, 110, 95), 0.2), ((3.0, 3.0, 3.0, 90, 90, 90), 10.0)):
        m = ol.cell_from_params(*cell)
        assert aa.grid_point_counts(m, spacing) == ol.grid_counts(m, spacing)
    assert aa.grid_point_counts(ol.cubic(25.83), 0.2) == (129, 129, 129)


def test_grid_point_counts_errors():
    with pytest.raises(ValueError):
        aa.grid_point_counts(ol.cubic(10.0), 0.0)
    with pytest.raises(ValueError):
        aa.grid_point_counts(np.zeros((3, 3)), 0.2)


@pytest.mark.parametrize("cell,n,sigma", [((24, 24, 24, 60, 60, 60), 600, 0.5), ((25.83,) * 3 + (90,) * 3, 424, 0.0),
                                          ((12.0, 30.0, 18.0, 65, 115, 80), 900, 0.5)])
def test_synthetic_framework_equals_oracle_generator(cell, n, sigma):
    m = ol.cell_from_params(*cell)
    got = aa.synth_framework(m, n, seed=1, stream=0, charge_sigma=sigma)
    ref = ol.synth_framework(m, n, seed=1, stream=0, charge_sigma=sigma)
    for a, b in zip(got, (ref.x, ref.y, ref.z, ref.q, ref.species)):
        assert np.array_equal(a, b)


def test_supercell_utilities_equal_oracle():
    prim, _ = ol.c1_system()
    reps = aa.supercell_replication(prim.m, 12.0)
    assert reps == ol.min_supercell(prim.m, 12.0) == (2, 2, 2)
    m, x, y, z, q, sp = aa.replicate_supercell(prim.m, prim.x, prim.y, prim.z, prim.q, prim.species, reps)
    ref = ol.replicate(prim, reps)
    for a, b in zip((m, x, y, z, q, sp), (ref.m, ref.x, ref.y, ref.z, ref.q, ref.species)):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        aa.supercell_replication(prim.m, -1.0)


def test_peng_robinson_equals_oracle():
    for bar in (0.1, 0.2259, 1.1525, 5.8799, 30.0):
        got = aa.peng_robinson_fugacity(304.1282, 7377300.0, 0.22394, 298.0, bar * 1e5)
        ref = ol.oracle().orc_pr_fugacity(C.c_double(304.1282), C.c_double(7377300.0), C.c_double(0.22394),
                                          C.c_double(298.0), C.c_double(bar * 1e5))
        assert got == ref


def test_tail_insertion_delta_equals_oracle_difference():
    fw, lj = ol.c0_system()
    site = [(0.0, 0.0, 0.0, 4, 0.0)]
    got = aa.lj_tail_insertion_delta(lj.eps, lj.sigma, lj.cutoff, fw.m, fw.species, site)
    l, keep = ol.lj_struct(lj)
    lat = ol.lattice(fw.m)
    p0, k0 = ol.particles(fw)
    with_guest = ol.Framework(fw.m, np.append(fw.x, 0.0), np.append(fw.y, 0.0), np.append(fw.z, 0.0),
                              np.append(fw.q, 0.0), np.append(fw.species, np.uint32(4)))
    p1, k1 = ol.particles(with_guest)
    ref = ol.oracle().orc_lj_tail_energy(C.byref(l), C.byref(p1), C.byref(lat)) - \
        ol.oracle().orc_lj_tail_energy(C.byref(l), C.byref(p0), C.byref(lat))
    assert got == ref and got < 0.0


def test_device_calls_without_a_context_are_rejected():
    lib = C.CDLL(aa.library_path())
    out = C.c_void_p()
    assert lib.aleator_gpu_framework_create(None, None, None, None, None, C.byref(out)) == 1
    lib.aleator_gpu_last_error.restype = C.c_char_p
    assert b"null" in lib.aleator_gpu_last_error()
    assert lib.aleator_gpu_ewald_state_propose(None, None, 0, None, 0, None) == 1


def _fma_horner(c, x):
    """Horner's rule with each step one correctly rounded fused multiply-add (as the device does)."""
    import mpmath
    mpmath.mp.prec = 200
    p = mpmath.mpf(c[-1])
    out = float(p)
    for k in range(len(c) - 2, -1, -1):
        out = float(mpmath.mpf(out) * mpmath.mpf(x) + mpmath.mpf(c[k]))
    return out


@pytest.mark.parametrize("alpha,rc,tol,deg", [(0.265, 12.0, 1e-17, 24), (0.265, 12.0, 1e-12, 18),
                                              (0.2, 10.0, 1e-17, 18), (0.3, 13.3, 1e-17, 28)])
def test_erf_kernel_polynomial_against_mpmath(alpha, rc, tol, deg):
    """erf(alpha r)/r from the polynomial the device evaluates, against 50-digit mpmath.  Pins the
    Chebyshev fit, its degree choice and the monomial conversion: the FP64 bound is a few ulps of
    2 alpha/sqrt(pi) (the value at r = 0), the MIXED bound its 1e-12."""
    import mpmath
    mpmath.mp.dps = 50
    c, xs = aa.erf_kernel_poly(alpha, rc, tol)
    assert len(c) == deg + 1
    scale = 2 * alpha / np.sqrt(np.pi)
    worst = 0.0
    for r in np.concatenate([np.linspace(0.0, rc, 241), [1e-9, 0.5 * rc, rc]]):
        s = r * r
        x = float(mpmath.mpf(s) * xs - 1)  # fma(s, xs, -1), one rounding
        got = _fma_horner(c, x)
        ref = scale if r == 0 else float(mpmath.erf(alpha * mpmath.mpf(r)) / mpmath.mpf(r))
        worst = max(worst, abs(got - ref) / scale)
    bound = 8e-16 if tol <= 1e-17 else 2e-12
    assert worst <= bound, worst


def test_erf_kernel_polynomial_planted_mistake_fails():
    """A dropped highest term (the degree cut one step early) must fail the FP64 bound."""
    import mpmath
    mpmath.mp.dps = 50
    c, xs = aa.erf_kernel_poly(0.265, 12.0, 1e-17)
    c = c[:-2]
    scale = 2 * 0.265 / np.sqrt(np.pi)
    r = 11.5
    got = _fma_horner(c, float(np.float64(r * r) * np.float64(xs) - 1.0))
    ref = float(mpmath.erf(0.265 * mpmath.mpf(r)) / mpmath.mpf(r))
    assert abs(got - ref) / scale > 8e-16

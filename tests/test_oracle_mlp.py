"""Pins of the MLP oracle (oracle/mlp.c) that the device network is checked against.

The FP64 gradient is pinned by central finite differences of the FP64 loss; the BF16-emulating
mode is pinned to the FP64 mode (it must be a small perturbation of it) and by the exact
bfloat16 rounding of a few known values.  The frequency table of the library (host function,
no device needed) must equal the oracle's."""
import numpy as np
import pytest

import oracle_lib as O


def _rows(n, seed=5):
    rng = np.random.default_rng(seed)
    frac = rng.random((n, 3))
    target = np.sin(2 * np.pi * frac[:, 0]) * np.cos(2 * np.pi * frac[:, 1]) + 0.3 * frac[:, 2]
    return frac, target


def test_frequency_table():
    k = O.mlp_frequencies()
    assert k.shape == (32, 3)
    assert len({tuple(v) for v in k}) == 32  # distinct
    # every axis carries 2^0 .. 2^5 on its own
    for a in range(3):
        for j in range(6):
            v = [0, 0, 0]
            v[a] = 2 ** j
            assert tuple(v) in {tuple(r) for r in k}
    aleator_amd = pytest.importorskip("aleator_amd")
    assert np.array_equal(aleator_amd.Mlp.frequencies(), k)
    assert aleator_amd.Mlp.num_params() == O.MLP_PARAMS == 214273


def test_features_are_periodic_in_the_cell():
    p = O.mlp_he_init(3).astype(np.float64)
    frac, _ = _rows(16)
    shifted = frac + np.array([1.0, -2.0, 3.0])
    assert np.allclose(O.mlp_forward(p, frac, False), O.mlp_forward(p, shifted, False), rtol=0, atol=1e-11)


def test_fp64_gradient_matches_central_differences():
    p = O.mlp_he_init(11).astype(np.float64)
    # non-zero biases so that their gradients are exercised away from the symmetric start
    rng = np.random.default_rng(2)
    for name, o, n in O.mlp_offsets():
        if name.startswith("b"):
            p[o:o + n] = rng.normal(0, 0.1, n)
    frac, target = _rows(6)
    loss, g = O.mlp_loss_grad(p, frac, target, False)
    assert loss > 0
    h = 1e-5
    picks = []
    for name, o, n in O.mlp_offsets():
        picks += list(o + rng.choice(n, size=min(n, 4), replace=False))
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.mlp_loss_grad(pp, frac, target, False)[0] - O.mlp_loss_grad(pm, frac, target, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_fp64_forward_matches_numpy_restatement():
    """An independent numpy forward pass of the same network (features, 4 SiLU layers, head)."""
    p = O.mlp_he_init(4).astype(np.float64)
    frac, _ = _rows(32)
    k = O.mlp_frequencies().astype(np.float64)
    ph = frac @ k.T
    x = np.empty((len(frac), 64))
    x[:, 0::2] = np.sin(2 * np.pi * ph)
    x[:, 1::2] = np.cos(2 * np.pi * ph)
    offs = {name: (o, n) for name, o, n in O.mlp_offsets()}
    a = x
    for l in range(1, 5):
        o, n = offs[f"W{l}"]
        W = p[o:o + n].reshape(a.shape[1], 256)
        ob, nb = offs[f"b{l}"]
        z = a @ W + p[ob:ob + nb]
        a = z / (1 + np.exp(-z))
    o, n = offs["w5"]
    y = a @ p[o:o + n] + p[offs["b5"][0]]
    assert np.allclose(O.mlp_forward(p, frac, False), y, rtol=1e-12, atol=1e-12)


def test_bf16_mode_is_a_small_perturbation():
    p = O.mlp_he_init(7).astype(np.float64)
    frac, target = _rows(16)
    y64 = O.mlp_forward(p, frac, False)
    y16 = O.mlp_forward(p, frac, True)
    assert not np.array_equal(y64, y16)
    assert np.max(np.abs(y64 - y16)) <= 0.05 * np.max(np.abs(y64)) + 1e-3
    l64, g64 = O.mlp_loss_grad(p, frac, target, False)
    l16, g16 = O.mlp_loss_grad(p, frac, target, True)
    assert abs(l64 - l16) <= 0.05 * l64
    for name, o, n in O.mlp_offsets():
        rel = np.linalg.norm(g16[o:o + n] - g64[o:o + n]) / np.linalg.norm(g64[o:o + n])
        assert rel < 0.1, (name, rel)


def test_bf16_rounding_of_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even gives 1; 1 + 3 * 2^-8 rounds up to 1 + 2^-6
    p = np.zeros(O.MLP_PARAMS)
    offs = {name: (o, n) for name, o, n in O.mlp_offsets()}
    # route one feature straight through: W1[0][0] = 1 + 2^-8 times feature 0 = sin(2 pi f0)
    p[offs["W1"][0]] = 1.0 + 2.0 ** -8
    p[offs["b1"][0]] = 30.0  # silu(z) ~ z for large z; the bias does not pass through bf16
    # head reads unit 0 of layer 4, which chains unit 0 of every layer through weights of 1
    for l in (2, 3, 4):
        p[offs[f"W{l}"][0]] = 1.0
    p[offs["w5"][0]] = 1.0
    f = np.array([[0.25, 0.0, 0.0]])  # sin(2 pi 0.25) = 1
    y64 = O.mlp_forward(p, f, False)[0]
    y16 = O.mlp_forward(p, f, True)[0]
    # FP64: silu(31 + 2^-8) = 31 + 2^-8 to 1e-11, and silu of that again, four times
    assert y64 == pytest.approx(31.0 + 2.0 ** -8, abs=1e-9)
    # BF16: the weight 1 + 2^-8 rounds to 1 (tie to even), so z = 31 and every silu(31) rounds to 31
    assert y16 == 31.0

"""Device MLP (BF16 MFMA forward/backward, FP32 accumulation, Adam) against the CPU oracle.

Tolerances: against the oracle's BF16-emulating mode (same roundings, FP64 sums) the device
differs by FP32 accumulation order and by rare one-ulp bfloat16 flips of an activation that
sits on a rounding boundary: outputs within 2e-3 of the output scale, gradient tensors within
2e-2 relative L2.  Against plain FP64 the BF16 network itself differs by a few percent."""
import numpy as np
import pytest

import aleator_amd as aa
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rows(n, seed=9):
    rng = np.random.default_rng(seed)
    frac = rng.random((n, 3)) * 2.0 - 0.5  # outside [0, 1) too: the features are periodic
    target = np.sin(2 * np.pi * frac[:, 0]) * np.cos(2 * np.pi * frac[:, 1]) + 0.5 * np.cos(4 * np.pi * frac[:, 2])
    return frac, target


def test_init_matches_documented_initialisation(ctx):
    net = aa.Mlp(ctx, max_rows=256, seed=17)
    assert np.array_equal(net.params, ol.mlp_he_init(17))


def test_forward_matches_oracle(ctx):
    print("mlp forward: 1000 rows, 3 passes of max_rows 400", flush=True)
    net = aa.Mlp(ctx, max_rows=400, seed=3)
    p = net.params.astype(np.float64)
    frac, _ = rows(1000)
    y = net.predict(frac).astype(np.float64)
    y16 = ol.mlp_forward(p, frac, True)
    y64 = ol.mlp_forward(p, frac, False)
    scale = np.max(np.abs(y16))
    err16 = np.max(np.abs(y - y16)) / scale
    err64 = np.max(np.abs(y - y64)) / scale
    print(f"mlp forward: max err vs bf16 oracle {err16:.2e}, vs fp64 oracle {err64:.2e} (of scale {scale:.3f})")
    assert err16 < 2e-3
    assert err64 < 5e-2


def test_gradients_match_oracle(ctx):
    print("mlp gradient: 256 rows", flush=True)
    net = aa.Mlp(ctx, max_rows=256, seed=5)
    p = net.params.astype(np.float64)
    frac, target = rows(256, seed=4)
    loss = net.loss_grad(frac, target)
    g = net.grads().astype(np.float64)
    l16, g16 = ol.mlp_loss_grad(p, frac, target, True)
    l64, g64 = ol.mlp_loss_grad(p, frac, target, False)
    assert loss == pytest.approx(l16, rel=2e-3)
    for name, o, n in ol.mlp_offsets():
        r16 = np.linalg.norm(g[o:o + n] - g16[o:o + n]) / np.linalg.norm(g16[o:o + n])
        r64 = np.linalg.norm(g[o:o + n] - g64[o:o + n]) / np.linalg.norm(g64[o:o + n])
        print(f"  {name}: rel L2 err vs bf16 oracle {r16:.2e}, vs fp64 {r64:.2e}")
        assert r16 < 2e-2, name
        assert r64 < 0.1, name


def test_one_adam_step_matches_host_update(ctx):
    """train(1 iteration): rows drawn by Philox(seed, stream) at calls 0..batch-1 as floor(u n),
    the oracle gradient on those rows, and Adam (lr, 0.9, 0.999, 1e-8, bias-corrected, t = 1)."""
    net = aa.Mlp(ctx, max_rows=128, seed=8)
    p0 = net.params.astype(np.float64)
    frac, target = rows(300, seed=6)
    d_f = aa.DeviceBuffer.from_array(frac)
    d_t = aa.DeviceBuffer.from_array(target)
    lr = 1e-3
    losses = net.train_device(int(d_f), int(d_t), len(frac), 128, 1, lr, seed=21, stream=4)
    idx = np.minimum((ol.uniforms(21, 4, 0, 128) * len(frac)).astype(np.int64), len(frac) - 1)
    l16, g16 = ol.mlp_loss_grad(p0, frac[idx], target[idx], True)
    assert losses[0] == pytest.approx(l16, rel=2e-3)
    g_dev = net.grads().astype(np.float64)
    # Adam at t = 1: m = 0.1 g / 0.1, v = 0.001 g^2 / 0.001, step = lr g / (|g| + eps)
    expect = p0 - lr * g_dev / (np.abs(g_dev) + 1e-8)
    p1 = net.params.astype(np.float64)
    assert np.allclose(p1, expect.astype(np.float32), rtol=0, atol=2e-7)
    # and the device gradient of that minibatch is the oracle's
    for name, o, n in ol.mlp_offsets():
        r = np.linalg.norm(g_dev[o:o + n] - g16[o:o + n]) / np.linalg.norm(g16[o:o + n])
        assert r < 2e-2, name


def test_training_fits_a_smooth_periodic_target(ctx):
    print("mlp training: 600 iterations, batch 512", flush=True)
    net = aa.Mlp(ctx, max_rows=512, seed=1)
    frac, target = rows(20000, seed=2)
    d_f = aa.DeviceBuffer.from_array(frac)
    d_t = aa.DeviceBuffer.from_array(target)
    losses = net.train_device(int(d_f), int(d_t), len(frac), 512, 600, 1e-3, seed=3, stream=0)
    first, last = float(np.mean(losses[:20])), float(np.mean(losses[-50:]))
    print(f"mlp training: loss {first:.4f} -> {last:.5f}")
    assert np.all(np.isfinite(losses))
    assert last < 0.02 * first
    test_f, test_t = rows(2000, seed=99)
    err = np.sqrt(np.mean((net.predict(test_f) - test_t) ** 2))
    print(f"mlp training: held-out rms error {err:.4f} (target rms {np.sqrt(np.mean(test_t ** 2)):.3f})")
    assert err < 0.1


def test_graph_replay_equals_eager_launches(ctx):
    """Training from replayed hipGraphs (eight iterations per replay, single ones for the rest:
    25 = 3 x 8 + 1 and 30 = 3 x 8 + 6) gives the eager launches' results bit for bit, over two
    train calls with other device work (a prediction) between them."""
    print("mlp training: graph replay vs eager launches", flush=True)
    frac, target = rows(5000, seed=12)
    d_f = aa.DeviceBuffer.from_array(frac)
    d_t = aa.DeviceBuffer.from_array(target)
    runs = []
    for graph in (True, False):
        net = aa.Mlp(ctx, max_rows=512, seed=4)
        net.use_graph(graph)
        l1 = net.train_device(int(d_f), int(d_t), len(frac), 512, 25, 1e-3, seed=5, stream=1)
        pred = net.predict(frac[:100])
        l2 = net.train_device(int(d_f), int(d_t), len(frac), 512, 30, 1e-3, seed=6, stream=1)
        runs.append((l1, pred, l2, net.params))
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_errors(ctx):
    net = aa.Mlp(ctx, max_rows=16, seed=0)
    frac, target = rows(32)
    with pytest.raises(ValueError):
        net.loss_grad(frac, target)  # more rows than max_rows
    with pytest.raises(ValueError):
        aa.Mlp(ctx, max_rows=0)
    with pytest.raises(ValueError):
        net.params = np.zeros(10, dtype=np.float32)

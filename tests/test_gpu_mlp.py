"""The warm-start network's gfx950 kernels (BF16 MFMA forward / backward,
Adam) against the CPU restatement oracle/mlp_oracle.py with the same bf16
rounding points.  Tolerances: the device accumulates in fp32 in MFMA order
and the oracle in float64, so an activation can land on the other side of
a bf16 rounding boundary now and then; errors are bounded relative to each
layer's gradient norm."""
import sys

import numpy as np
import pytest

from optimisation2forme_amd import capi
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    sys.stdout.flush()
    yield


def sample(n, rng):
    return np.stack([rng.uniform(-5, 7, n), rng.uniform(-6, 6, n), rng.uniform(1e-4, 10, n),
                     rng.uniform(0, 0.06, n), rng.uniform(0.2, 0.6, n), rng.uniform(0.08, 0.15, n),
                     rng.uniform(0.5, 0.85, n), rng.uniform(0, 4, n)], 1).astype(np.float32)


def smooth_target(raw):
    x, y, d = raw[:, 0] - 0.855, raw[:, 1], raw[:, 2]
    return np.stack([0.1 * np.exp(-d) * (1 + raw[:, 3] * 10), 0.2 * np.tanh(x) * raw[:, 6],
                     0.05 * np.sin(y) + 0.01 * raw[:, 7], 0.3 * np.exp(-(x * x + y * y) / 4)], 1).astype(np.float32)


def test_forward_matches_the_oracle():
    net = capi.WarmStartNet(seed=3)
    w, fb = net.weights()
    rng = np.random.default_rng(1)
    raw = sample(1000, rng)                   # not a multiple of 64: padded rows are masked
    y = net.forward(raw)
    yo, _, _ = MO.forward(w, fb, raw, net.layout, emulate=True)
    err = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    rms = np.sqrt(np.mean((y - yo) ** 2) / np.mean(yo ** 2))
    print("forward max err / max |y| = %.3g, rms %.3g" % (err, rms), flush=True)
    assert err < 1e-2 and rms < 1e-3


def test_gradient_matches_the_oracle():
    net = capi.WarmStartNet(seed=4)
    w, fb = net.weights()
    rng = np.random.default_rng(2)
    raw = sample(2048, rng)
    tgt = smooth_target(raw)
    loss, g = net.loss_grad(raw, tgt)
    lo, go = MO.loss_grad(w, fb, raw, tgt, net.layout, emulate=True)
    print("loss %.6g oracle %.6g" % (loss, lo), flush=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for l, (K, N, wo, bo) in enumerate(net.layout):
        for name, sl in (("W", slice(wo, wo + K * N)), ("b", slice(bo, bo + N))):
            a, b = g[sl], go[sl]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print("layer %d %s rel L2 err %.3g" % (l, name, rel), flush=True)
            assert rel < 1e-2


def test_every_gemm_kernel_gives_the_same_bits():
    """The seven GEMM kernels (single / double-buffered, four tile shapes)
    accumulate each output in the same k order: forward outputs, loss and
    gradients are bit-identical whichever runs (a forced kernel runs where
    its tile divides the shape, the 64 x 64 one elsewhere)."""
    net = capi.WarmStartNet(seed=6)
    rng = np.random.default_rng(4)
    raw = sample(1024, rng)
    tgt = smooth_target(raw)
    ref = None
    for v in range(7):
        net.set_gemm_variant(v)
        y = net.forward(raw)
        loss, g = net.loss_grad(raw, tgt)
        if ref is None:
            ref = (y, loss, g)
        else:
            assert np.array_equal(y, ref[0]) and loss == ref[1] and np.array_equal(g, ref[2]), v
    net.set_gemm_variant(-1)
    assert np.array_equal(net.forward(raw), ref[0])
    assert np.array_equal(net.loss_grad(raw, tgt)[1], ref[2])


def test_adam_training_reduces_the_loss():
    net = capi.WarmStartNet(seed=5)
    rng = np.random.default_rng(3)
    raw = sample(50000, rng)
    tgt = smooth_target(raw)
    net.set_data(raw, tgt)
    losses = net.train(300, batch=4096, lr=1e-3, seed=7)
    print("loss %.4g -> %.4g" % (losses[0], losses[-1]), flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    # held-out error tracks the training loss
    test = sample(4000, rng)
    err = np.mean((net.forward(test) - smooth_target(test)) ** 2)
    assert err < 3 * losses[-20:].mean()
    # training is deterministic (no atomics; seeded batches)
    net2 = capi.WarmStartNet(seed=5)
    net2.set_data(raw, tgt)
    l2 = net2.train(300, batch=4096, lr=1e-3, seed=7)
    assert np.array_equal(losses, l2)
    assert np.array_equal(net.weights()[0], net2.weights()[0])


def test_warm_start_reduces_the_steps_to_tolerance():
    """Ablation: the network trained on solver fields (8 candidates) against
    the freestream cold start on 2 held-out candidates; both iterate to the
    same density-residual tolerance."""
    from optimisation2forme_amd import warmstart as W
    fields = W.solve_fields(W.candidates(8, seed=11), tol=0.25, max_steps=40000)
    assert all(f[3] > 1000 for f in fields)     # developed fields (to 0.25 or 40000 steps)
    net, losses, _ = W.train(fields, steps=600, batch=8192, lr=1e-3, seed=0)
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()
    rows = W.compare(net, W.candidates(2, seed=12), tol=0.4, max_steps=40000)
    for r in rows:
        print("cells %d cold %d warm %d steps" % (r["cells"], r["cold_steps"], r["warm_steps"]), flush=True)
        assert r["cold_res"] <= 0.4 and r["warm_res"] <= 0.4
    assert sum(r["warm_steps"] for r in rows) < 0.7 * sum(r["cold_steps"] for r in rows)

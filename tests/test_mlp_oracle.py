"""CPU pins of the warm-start network's restatement (oracle/mlp_oracle.py):
its float64 forward / loss / gradient against PyTorch-CPU autograd on the
same weights, and its bf16 rounding against the bit-level definition."""
import numpy as np
import torch

from oracle import mlp_oracle as MO

LAYOUT = []
_off = 0
for _K, _N in [(64, 256), (256, 256), (256, 256), (256, 256), (256, 64)]:
    LAYOUT.append([_K, _N, _off, 0])
    _off += _K * _N
for _l, (_K, _N) in enumerate([(64, 256), (256, 256), (256, 256), (256, 256), (256, 64)]):
    LAYOUT[_l][3] = _off
    _off += _N
LAYOUT = [tuple(x) for x in LAYOUT]
NPARAM = _off


def sample(n, rng):
    raw = np.stack([rng.uniform(-5, 7, n), rng.uniform(-6, 6, n), rng.uniform(1e-4, 10, n),
                    rng.uniform(0, 0.06, n), rng.uniform(0.2, 0.6, n), rng.uniform(0.08, 0.15, n),
                    rng.uniform(0.5, 0.85, n), rng.uniform(0, 4, n)], 1)
    return raw


def test_bf16_rounding_is_round_to_nearest_even():
    """bf16 keeps 7 mantissa bits: spacing 2^-7 at 1.0."""
    u = 2.0 ** -7
    x = np.array([1.0 + u, 1.0 + u / 2, 1.0 + 1.5 * u, 1.0 + u / 4, 1.0 + 0.75 * u, -2.5], np.float32)
    r = MO.bf16(x)
    assert r[0] == 1.0 + u                 # representable
    assert r[1] == 1.0                     # tie -> even mantissa (0)
    assert r[2] == 1.0 + 2 * u             # tie -> even mantissa (2)
    assert r[3] == 1.0 and r[4] == 1.0 + u  # below / above the half-way point
    assert r[5] == -2.5
    assert np.array_equal(MO.bf16(r), r)   # idempotent


def test_float64_gradient_matches_torch_autograd():
    rng = np.random.default_rng(0)
    w = np.zeros(NPARAM)
    for K, N, wo, bo in LAYOUT:
        w[wo:wo + K * N] = rng.uniform(-1, 1, K * N) * np.sqrt(6.0 / K)
        w[bo:bo + N] = rng.uniform(-0.1, 0.1, N)
    fb = rng.standard_normal((2, 16))
    raw = sample(37, rng)
    tgt = rng.standard_normal((37, 4))
    loss, g = MO.loss_grad(w, fb, raw, tgt, LAYOUT, emulate=False)
    # torch: the same network from the same parameters
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(MO.features(raw, fb, emulate=False), dtype=torch.float64)
    for l, (K, N, wo, bo) in enumerate(LAYOUT):
        z = a @ wt[wo:wo + K * N].reshape(K, N) + wt[bo:bo + N]
        a = torch.nn.functional.silu(z) if l < 4 else z
    y = a[:, :4]
    lt = ((y - torch.tensor(tgt)) ** 2).mean()
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-12 * abs(loss)
    gt = wt.grad.numpy()
    assert np.max(np.abs(g - gt)) <= 1e-10 * np.max(np.abs(gt))
    # a plausible slip (SiLU' without the z (1 - s) term) would not pass
    assert np.max(np.abs(gt)) > 0

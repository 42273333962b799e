"""CPU restatement of the learned warm-start network (test infrastructure).

Only tests/ may use this module: it checks the gfx950 network kernels
(optimisation2forme_amd/csrc/o2f_mlp.hip).  The network has no counterpart
in the reference; its arithmetic is fixed here in numpy:

  features  the 8 raw inputs (x, y, wall distance, m, p, t, Mach, AoA deg)
            scaled, then sin / cos (2 pi b_k . (x - 0.855, y)) for 16
            frequencies, zero padded to 64 columns
  layers    z_l = a_l W_l + b_l, a_{l+1} = silu(z_l) for 4 hidden layers,
            y = a_4 W_4 + b_4 (first 4 of 64 columns used)
  loss      mean over n samples and 4 outputs of (y - t)^2

With emulate=True the bf16 roundings of the device are reproduced at the
same points (inputs, weights, activations, output gradient, hidden
gradients), accumulation in float64 then rounded to float32; with
emulate=False everything is float64 (the form PyTorch's autograd checks).
"""
import numpy as np

K_IN, K_HID, K_OUT, N_FOURIER = 64, 256, 64, 16


def bf16(x):
    """float32 -> bf16 -> float32, round to nearest even."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


def features(raw, fb, emulate=True):
    raw = np.asarray(raw, dtype=np.float64 if not emulate else np.float32)
    dt = raw.dtype
    n = raw.shape[0]
    f = np.zeros((n, K_IN), dtype=dt)
    x = raw[:, 0] - dt.type(0.855)
    y = raw[:, 1]
    f[:, 0] = x / dt.type(20.0)
    f[:, 1] = y / dt.type(20.0)
    f[:, 2] = np.log10(raw[:, 2] + dt.type(1e-4)) / dt.type(4.0)
    f[:, 3] = raw[:, 3] / dt.type(0.06)
    f[:, 4] = (raw[:, 4] - dt.type(0.4)) / dt.type(0.2)
    f[:, 5] = (raw[:, 5] - dt.type(0.115)) / dt.type(0.035)
    f[:, 6] = (raw[:, 6] - dt.type(0.675)) / dt.type(0.175)
    f[:, 7] = raw[:, 7] / dt.type(4.0)
    fb = np.asarray(fb, dtype=dt)
    for k in range(N_FOURIER):
        a = dt.type(6.283185307179586) * (fb[0, k] * x + fb[1, k] * y)
        f[:, 8 + 2 * k] = np.sin(a)
        f[:, 9 + 2 * k] = np.cos(a)
    return bf16(f) if emulate else f


def unpack(w, layout):
    """[(W_l (K x N), b_l)] from the flat parameter vector."""
    return [(w[wo:wo + K * N].reshape(K, N), w[bo:bo + N]) for K, N, wo, bo in layout]


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(w, fb, raw, layout, emulate=True):
    layers = unpack(np.asarray(w, np.float64 if not emulate else np.float32), layout)
    a = features(raw, fb, emulate).astype(np.float64)
    acts, zs = [a], []
    for l, (W, b) in enumerate(layers):
        Wd = (bf16(W) if emulate else W).astype(np.float64)
        z = a @ Wd
        if emulate:
            z = z.astype(np.float32) + b.astype(np.float32)
        else:
            z = z + b
        if l < len(layers) - 1:
            zs.append(z)
            s = _sig(z)
            a = z * s
            a = bf16(a).astype(np.float64) if emulate else a
            acts.append(a)
        else:
            return z[:, :4].astype(np.float64), acts, zs


def loss_grad(w, fb, raw, targets, layout, emulate=True):
    """(loss, gradient) with the parameter layout of the device network."""
    w = np.asarray(w)
    n = raw.shape[0]
    y, acts, zs = forward(w, fb, raw, layout, emulate)
    t = np.asarray(targets, dtype=np.float64)
    d = y - t
    loss = float(np.sum(d * d) / (4.0 * n))
    layers = unpack(w.astype(np.float32 if emulate else np.float64), layout)
    g = np.zeros(w.shape[0], dtype=np.float64)
    dZ = np.zeros((n, K_OUT))
    if emulate:
        dZ[:, :4] = bf16((2.0 / (4.0 * n)) * d.astype(np.float32)).astype(np.float64)
    else:
        dZ[:, :4] = 2.0 * d / (4.0 * n)
    for l in range(len(layers) - 1, -1, -1):
        K, N, wo, bo = layout[l]
        g[wo:wo + K * N] = (acts[l].T @ dZ).reshape(-1)
        g[bo:bo + N] = dZ.sum(axis=0)
        if l == 0:
            break
        W = layers[l][0]
        Wd = (bf16(W) if emulate else W).astype(np.float64)
        dA = dZ @ Wd.T
        z = zs[l - 1]
        s = _sig(z)
        dz = dA * (s * (1.0 + z * (1.0 - s)))
        dZ = bf16(dz.astype(np.float32)).astype(np.float64) if emulate else dz
    return loss, g

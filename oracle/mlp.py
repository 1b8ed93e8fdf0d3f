"""CPU oracle of the learned traveltime surrogate.  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py import this module, to check the
GPU network (kirch_datuming_amd/csrc/kmig_nn.hip) or to build its training
records.  The reference has no network; this restates the definition in
include/kmig_nn.h in float64 numpy:

  features  u = (c - c_min) / (c_max - c_min) for c in (s, z, x), then
            sin(pi 2^k u), cos(pi 2^k u), k = 0..7, per coordinate;
  layers    4 x (128 wide, SiLU) and a linear output y[4];
  outputs   T = d S0 (1 + y0), dT/dx = S0 y1, dT/dz = S0 y2, dT/ds = S0 y3,
            d = sqrt((x - s)^2 + (z - z_src)^2);
  loss      mean over records of ((T' - T)/sigma_t)^2 + sum((g' - g)/sigma_g)^2.

``bf16=True`` rounds to bfloat16 (nearest even) wherever the GPU stores a
BF16 operand: the features, the weights, the hidden activations, dL/dy and
the back-propagated dZ.  Sums stay float64, so the GPU result differs from
it only by FP32 accumulation order and rounding-boundary flips.

Pinning: the gradient is checked against central finite differences of the
float64 loss, and the forward pass against the closed form of a one-neuron
network (tests/test_nn_oracle.py).
"""

from __future__ import annotations

import numpy as np

FEAT, WIDTH, OUT = 48, 128, 4
LK = (FEAT, WIDTH, WIDTH, WIDTH, WIDTH)
LM = (WIDTH, WIDTH, WIDTH, WIDTH, OUT)
NPARAMS = sum(m * k + m for m, k in zip(LM, LK))


def bf16(a):
    """Round float64/float32 values to bfloat16 (nearest even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """Parameter vector -> [(W [M][K], b [M])] per layer (float64)."""
    p = np.asarray(params, dtype=np.float64)
    assert p.size == NPARAMS
    out, o = [], 0
    for m, k in zip(LM, LK):
        W = p[o:o + m * k].reshape(m, k)
        o += m * k
        b = p[o:o + m]
        o += m
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def features(cfg, s, z, x):
    cols = []
    for c, lo, hi in ((s, cfg["s_min"], cfg["s_max"]), (z, cfg["z_min"], cfg["z_max"]),
                      (x, cfg["x_min"], cfg["x_max"])):
        u = (np.asarray(c, np.float64) - lo) / (hi - lo if hi > lo else 1.0)
        for k in range(8):
            f = np.pi * (2.0 ** k) * u
            cols += [np.sin(f), np.cos(f)]
    return np.stack(cols, axis=1)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, cfg, pts, bf16_ops=False):
    """pts [n][3] (s, z, x) -> outputs [n][4] (T, dT/dx, dT/dz, dT/ds), raw y
    [n][4] and the activations (for backward)."""
    pts = np.asarray(pts, np.float64)
    s, z, x = pts[:, 0], pts[:, 1], pts[:, 2]
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    X = r(features(cfg, s, z, x))
    acts, zs = [X], []
    h = X
    for W, b in layers[:4]:
        zl = h @ r(W).T + b
        zs.append(zl)
        h = r(_silu(zl))
        acts.append(h)
    W5, b5 = layers[4]
    y = h @ r(W5).T + b5
    d = np.sqrt((x - s) ** 2 + (z - cfg["z_src"]) ** 2)
    S0 = cfg["slowness"]
    out = np.stack([d * S0 * (1.0 + y[:, 0]), S0 * y[:, 1], S0 * y[:, 2], S0 * y[:, 3]], axis=1)
    return out, y, (acts, zs, d)


def loss_grad(params, cfg, records, bf16_ops=False):
    """Mean loss over records [n][7] and its gradient (parameter layout)."""
    rec = np.asarray(records, np.float64)
    n = rec.shape[0]
    out, y, (acts, zs, d) = forward(params, cfg, rec[:, :3], bf16_ops)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    S0 = cfg["slowness"]
    ist2, isg2 = 1.0 / cfg["sigma_t"] ** 2, 1.0 / cfg["sigma_g"] ** 2
    res = out - rec[:, 3:7]
    inv2 = np.array([ist2, isg2, isg2, isg2])
    loss = float(np.sum(res ** 2 * inv2) / n)
    dpdy = np.stack([d * S0, np.full(n, S0), np.full(n, S0), np.full(n, S0)], axis=1)
    dy = r(2.0 * res * inv2 * dpdy / n)
    layers = unpack(params)
    grads = [None] * 5
    W5 = r(layers[4][0])
    grads[4] = (dy.T @ acts[4], dy.sum(axis=0))
    dz = r((dy @ W5) * _silu_grad(zs[3]))
    for l in (3, 2, 1, 0):
        grads[l] = (dz.T @ acts[l], dz.sum(axis=0))
        if l > 0:
            dz = r((dz @ r(layers[l][0])) * _silu_grad(zs[l - 1]))
    return loss, pack(grads)


def init_params(seed, scale_out=0.1):
    """A float64 random parameter vector (uniform, variance 1/fan-in)."""
    rng = np.random.default_rng(seed)
    layers = []
    for i, (m, k) in enumerate(zip(LM, LK)):
        a = np.sqrt(3.0 / k) * (scale_out if i == 4 else 1.0)
        layers.append((rng.uniform(-a, a, (m, k)), rng.uniform(-0.1, 0.1, m)))
    return pack(layers)


def vz_time(s, z, x, v0, g, zs=0.0):
    """First-arrival time in v(z) = v0 + g z from a source at (s, zs):
    acosh(1 + g^2 r^2 / (2 v(zs) v(z))) / g (straight-ray r/v0 for g = 0)."""
    s, z, x = (np.asarray(a, np.float64) for a in (s, z, x))
    r2 = (x - s) ** 2 + (z - zs) ** 2
    if g == 0.0:
        return np.sqrt(r2) / v0
    vs, vz = v0 + g * zs, v0 + g * z
    return np.arccosh(1.0 + g * g * r2 / (2.0 * vs * vz)) / g


def vz_records(s, z, x, v0, g, zs=0.0, h=1e-5):
    """Records [n][7] (s, z, x, T, dT/dx, dT/dz, dT/ds) of the v(z) closed form,
    derivatives by central differences of step h (km)."""
    T = vz_time(s, z, x, v0, g, zs)
    tx = (vz_time(s, z, x + h, v0, g, zs) - vz_time(s, z, x - h, v0, g, zs)) / (2 * h)
    tz = (vz_time(s, z + h, x, v0, g, zs) - vz_time(s, z - h, x, v0, g, zs)) / (2 * h)
    ts = (vz_time(s + h, z, x, v0, g, zs) - vz_time(s - h, z, x, v0, g, zs)) / (2 * h)
    return np.stack([np.asarray(s, np.float64) + 0 * T, np.asarray(z, np.float64) + 0 * T,
                     np.asarray(x, np.float64) + 0 * T, T, tx, tz, ts], axis=1)

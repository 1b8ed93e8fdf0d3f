"""CPU checks of the surrogate network's oracle (oracle/mlp.py): the loss
gradient against central finite differences, the forward pass against closed
forms, the BF16 rounding, the parameter count of include/kmig_nn.h, and the
v(z) traveltime used for training records against the eikonal equation."""
import math
import os
import re

import numpy as np

from oracle import mlp as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = dict(s_min=0.0, s_max=3.0, z_min=0.0, z_max=2.0, x_min=0.0, x_max=3.0, z_src=0.0,
           slowness=0.5, sigma_t=0.01, sigma_g=0.1)


def test_parameter_count_matches_header():
    txt = open(os.path.join(ROOT, "include", "kmig_nn.h")).read()
    n = int(re.search(r"#define KMIG_NN_PARAMS (\d+)", txt).group(1))
    assert n == M.NPARAMS == 56324


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    assert M.bf16(one) == one
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0                  # tie -> even (1.0)
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6   # tie -> even (odd below)
    assert M.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert M.bf16(-2.5) == -2.5 and M.bf16(0.0) == 0.0


def test_forward_closed_forms():
    layers = M.unpack(np.zeros(M.NPARAMS))
    layers[4] = (layers[4][0], np.array([0.2, 0.3, -0.1, 0.05]))
    p = M.pack(layers)
    pts = np.array([[0.5, 1.0, 2.0], [1.0, 0.3, 1.0]])
    out, _, _ = M.forward(p, CFG, pts)
    d = np.hypot(pts[:, 2] - pts[:, 0], pts[:, 1])
    np.testing.assert_allclose(out[:, 0], d * 0.5 * 1.2, rtol=1e-15)
    np.testing.assert_allclose(out[:, 1:], np.tile([0.15, -0.05, 0.025], (2, 1)), rtol=1e-15)
    # one path: feature 0 = sin(pi u_s) through four SiLUs into y0
    layers = M.unpack(np.zeros(M.NPARAMS))
    for l in range(5):
        layers[l][0][0, 0] = 1.0
    out, y, _ = M.forward(M.pack(layers), CFG, pts)
    silu = lambda v: v / (1.0 + math.exp(-v))
    for k in range(2):
        v = math.sin(math.pi * pts[k, 0] / 3.0)
        for _ in range(4):
            v = silu(v)
        assert abs(y[k, 0] - v) < 1e-14


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = M.init_params(3, scale_out=1.0)
    s, z, x = rng.uniform(0, 3, 6), rng.uniform(0.1, 2, 6), rng.uniform(0, 3, 6)
    rec = M.vz_records(s, z, x, 1.5, 0.5)
    loss, g = M.loss_grad(p, CFG, rec)
    idx = np.concatenate([rng.choice(M.NPARAMS, 60, replace=False),
                          np.arange(M.NPARAMS - 4, M.NPARAMS)])  # output biases
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_grad(q, CFG, rec)
        q[i] -= 2 * h
        lm, _ = M.loss_grad(q, CFG, rec)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_vz_time_solves_the_eikonal_equation():
    v0, g, h = 1.5, 0.5, 1e-5
    s = 1.0
    z = np.linspace(0.05, 2.0, 9)
    x = np.linspace(0.0, 3.0, 9)
    Z, X = np.meshgrid(z, x)
    rec = M.vz_records(np.full(Z.size, s), Z.ravel(), X.ravel(), v0, g, h=h)
    slow = np.hypot(rec[:, 4], rec[:, 5])
    np.testing.assert_allclose(slow * (v0 + g * rec[:, 1]), 1.0, rtol=1e-6)
    # lateral invariance: dT/ds = -dT/dx
    np.testing.assert_allclose(rec[:, 6], -rec[:, 4], atol=1e-6)
    # homogeneous limit
    t0 = M.vz_time(0.0, 1.0, 1.0, 2.0, 0.0)
    assert abs(t0 - math.sqrt(2.0) / 2.0) < 1e-15

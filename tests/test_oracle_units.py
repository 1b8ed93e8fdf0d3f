"""Building blocks of the CPU oracle against closed forms and brute force.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign, swapped basis function, off-by-one window) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
fp = C.POINTER(C.c_float)


def _p(a):
    return a.ctypes.data_as(fp)


def test_doubint_impulse_and_constant():
    n, k = 50, 17
    tr = np.zeros(n, np.float32)
    tr[k] = 1.0
    L.kmo_doubint_f32(n, _p(tr))
    j = np.arange(n)
    assert np.array_equal(tr, np.where(j <= k, n - k, n - j).astype(np.float32))
    a = 0.5
    tr = np.full(n, a, np.float32)
    L.kmo_doubint_f32(n, _p(tr))
    exp = a * (n * (n + 1) / 2 - j * (j + 1) / 2)
    assert np.array_equal(tr, exp.astype(np.float32))


def test_pick_linear_trace_is_zero_and_impulse_weights():
    nt, t0, dt = 200, 0.0, 0.004
    tr = (0.3 + 0.01 * np.arange(nt)).astype(np.float32)
    for ti, dlt in ((0.2, 0.0), (0.3137, 0.0051), (0.5, 0.02)):
        assert abs(L.kmo_pick_f32(nt, t0, dt, ti, dlt, _p(tr))) < 1e-5
    # impulse at k, ti on sample k, no anti-alias widening: taps k-1, k, k+1
    k = 60
    tr = np.zeros(nt, np.float32)
    tr[k] = 1.0
    v = L.kmo_pick_f32(nt, t0, dt, np.float32(k * dt), 0.0, _p(tr))
    assert abs(v - 2.0 / 9.0) < 1e-6  # imp = (dt / 3dt)^2, 2*tr[k]
    # an impulse one sample later is picked by the outer tap with weight -1
    v = L.kmo_pick_f32(nt, t0, dt, np.float32((k - 1) * dt), 0.0, _p(tr))
    assert abs(v + 1.0 / 9.0) < 1e-6
    # widening delta moves the side taps: imp = (dt/(3dt + 2 delta))^2
    dlt = 2 * dt
    v = L.kmo_pick_f32(nt, t0, dt, np.float32(k * dt), np.float32(dlt), _p(tr))
    assert abs(v - 2.0 * (1.0 / 7.0) ** 2) < 1e-6
    # window rejections: it >= nt-1, itm < 0, itp >= nt-1
    assert L.kmo_pick_f32(nt, t0, dt, np.float32((nt - 1) * dt), 0.0, _p(tr)) == 0.0
    assert L.kmo_pick_f32(nt, t0, dt, np.float32(0.5 * dt), 0.0, _p(tr)) == 0.0
    assert L.kmo_pick_f32(nt, t0, dt, np.float32((nt - 2.5) * dt), 0.0, _p(tr)) == 0.0


def _cubic(p):
    return 0.3 + 0.8 * p - 0.5 * p ** 2 + 0.2 * p ** 3, 0.8 - 1.0 * p + 0.6 * p ** 2


def test_hermite_exact_on_cubic_and_identity_at_nodes():
    nz, nx, npos, o, d = 3, 4, 9, 0.1, 0.25
    pos_nodes = o + d * np.arange(npos)
    scale = np.linspace(0.5, 1.5, nz * nx)
    tab = np.concatenate([(_cubic(p)[0] * scale) for p in pos_nodes]).astype(np.float32)
    der = np.concatenate([(_cubic(p)[1] * scale) for p in pos_nodes]).astype(np.float32)
    T = np.zeros(nz * nx, np.float32)
    Tx = np.zeros(nz * nx, np.float32)
    ino = C.c_int()
    for pos in (0.2, 0.77, 1.3, 2.04):
        L.kmo_interp_side_f32(ord("h"), nz, nx, 0.01, npos, o, d, pos, _p(tab), _p(der), _p(T),
                              _p(Tx), C.byref(ino))
        v, dv = _cubic(np.float64(np.float32(pos)))
        assert np.allclose(T, v * scale, rtol=2e-6, atol=1e-6)
        assert np.allclose(Tx, dv * scale, rtol=2e-5, atol=1e-5)
    # linear interpolation is not exact on a cubic: its error must show
    L.kmo_interp_side_f32(ord("l"), nz, nx, 0.01, npos, o, d, 0.77, _p(tab), _p(der), _p(T),
                          _p(Tx), C.byref(ino))
    assert not np.allclose(T, _cubic(0.77)[0] * scale, rtol=1e-4)
    # node positions return the node slice; outside the axis clamps to the ends
    for pos, k in ((o + 3 * d, 3), (-1.0, 0), (o + d * (npos - 1) + 0.3, npos - 1)):
        L.kmo_interp_side_f32(ord("h"), nz, nx, 0.01, npos, o, d, pos, _p(tab), _p(der),
                              _p(T), _p(Tx), C.byref(ino))
        assert ino.value == k
        assert np.allclose(T, tab[k * nz * nx:(k + 1) * nz * nx], rtol=1e-6)


def test_fd_gradient_exact_on_linear_field():
    nz, nx, dz, dx = 9, 11, 0.02, 0.03
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="xy")  # (nx, nz)
    a, b = 0.7, -0.4
    tab = (a * iz * dz + b * ix * dx).astype(np.float32).ravel()
    gx = np.zeros_like(tab)
    gz = np.zeros_like(tab)
    for radius in (1, 2, 3):
        L.kmo_grad_plane_f32(_p(tab), nz, nx, dz, dx, radius, _p(gx), _p(gz))
        assert np.allclose(gx, b, atol=2e-5) and np.allclose(gz, a, atol=2e-5)


def _box(f, axis, r):
    n = f.shape[axis]
    out = np.empty_like(f, dtype=np.float64)
    for i in range(n):
        lo, hi = max(0, i - r), min(n - 1, i + r)
        sl = [slice(None)] * 2
        sl[axis] = slice(lo, hi + 1)
        osl = [slice(None)] * 2
        osl[axis] = i
        out[tuple(osl)] = f[tuple(sl)].astype(np.float64).mean(axis=axis)
    return out


def test_triangular_smoothing_brute_force():
    rng = np.random.default_rng(3)
    nz, nx = 13, 9
    f = rng.normal(size=(nx, nz)).astype(np.float32)  # memory (x, z): z fastest
    for r in (1, 2, 3):
        g = f.copy().ravel()
        L.kmo_smooth_f32(_p(g), nz, nx, r)
        ref = f.astype(np.float64)
        for _ in range(2):
            ref = _box(_box(ref, 1, r), 0, r)  # z (axis 1) then x (axis 0)
        assert np.allclose(g.reshape(nx, nz), ref, atol=1e-5)
    g = np.full(nz * nx, 2.5, np.float32)
    L.kmo_smooth_f32(_p(g), nz, nx, 2)
    assert np.allclose(g, 2.5, atol=1e-6)


def test_angle_closed_form():
    for a, b in ((0.3, 1.2), (-2.0, 2.5), (0.0, math.pi), (1.0, 1.0)):
        got = L.kmo_angle_f32(math.cos(a), math.sin(a), 2 * math.cos(b), 2 * math.sin(b), 0)
        exp = abs(a - b) % (2 * math.pi)
        exp = math.degrees(min(exp, 2 * math.pi - exp))
        assert abs(got - exp) < 1e-4
        half = L.kmo_angle_f32(math.cos(a), math.sin(a), math.cos(b), math.sin(b), 1)
        assert abs(half - exp / 2) < 1e-4
    assert math.isnan(L.kmo_angle_f32(0.0, 1e-13, 1.0, 0.0, 0))


def test_linear_bins_and_weights():
    ig0, ig1, w0, w1 = C.c_int(), C.c_int(), C.c_float(), C.c_float()
    args = (C.byref(ig0), C.byref(ig1), C.byref(w0), C.byref(w1))
    # the reference's own worked example, src/kirchmig2d_cig.cpp:45-49
    assert L.kmo_linear_bins_f32(25.4, 0.0, 1.0, 61, *args)
    assert (ig0.value, ig1.value) == (25, 26)
    assert abs(w0.value - 0.6) < 1e-5 and abs(w1.value - 0.4) < 1e-5
    assert L.kmo_linear_bins_f32(60.0, 0.0, 1.0, 61, *args) and (ig0.value, w0.value) == (60, 1)
    assert L.kmo_linear_bins_f32(-1e-6, 0.0, 1.0, 61, *args) and (ig0.value, w0.value) == (0, 1)
    assert not L.kmo_linear_bins_f32(60.01, 0.0, 1.0, 61, *args)
    assert not L.kmo_linear_bins_f32(-0.01, 0.0, 1.0, 61, *args)
    # cosine ramps: 1 inside, 0 outside, 0.5 at mid band
    assert L.kmo_leg_weight_f32(75.0, 90.0, 10.0) == 1.0
    assert abs(L.kmo_leg_weight_f32(85.0, 90.0, 10.0) - 0.5) < 1e-6
    assert L.kmo_leg_weight_f32(90.0, 90.0, 10.0) == 0.0
    assert L.kmo_leg_weight_f32(40.0, 40.0, 0.0) == 1.0  # hard cone keeps the edge
    assert L.kmo_floor_weight_f32(0.3, 0.0, 0.05) == 0.0  # rmin 0: off
    assert abs(L.kmo_floor_weight_f32(0.125, 0.15, 0.05) - 0.5) < 1e-6
    assert abs(L.kmo_floor_weight_f32(0.05, 0.1, 0.2) - 0.5) < 1e-6  # inner <= 0 branch
    # depth mute: cutoff runs from amin at iz=0 to amax at iz=nz-1
    assert abs(L.kmo_depth_mute_f32(85.0, 0, 11, 80.0, 50.0, 10.0) - 0.5) < 1e-6
    assert abs(L.kmo_depth_mute_f32(55.0, 10, 11, 80.0, 50.0, 10.0) - 0.5) < 1e-6
    assert L.kmo_depth_mute_f32(64.9, 5, 11, 80.0, 50.0, 10.0) == 1.0
    assert abs(L.kmo_angle_taper_f32(55.0, 50.0, 60.0) - 0.5) < 1e-6
    assert L.kmo_angle_taper_f32(55.0, -1.0, 60.0) == 1.0


def test_direct_mute_ramp():
    nt, t0, dt = 101, 0.0, 0.01
    tr = np.ones(nt, np.float32)
    L.kmo_mute_f32(_p(tr), nt, t0, dt, 1.0, 1.4, 0.5, 0.1, 0.2, 0)  # boundary 0.3
    t = t0 + np.arange(nt, dtype=np.float32) * np.float32(dt)
    assert np.all(tr[t <= 0.3 - 1e-6] == 0)
    assert np.all(tr[t >= 0.5 + 1e-6] == 1)
    assert abs(tr[40] - 0.5) < 1e-5  # mid ramp (t = 0.4)
    tr = np.ones(nt, np.float32)
    L.kmo_mute_f32(_p(tr), nt, t0, dt, 1.0, 1.4, 0.5, 0.5, 0.2, 1)  # inverted: 0.3
    assert abs(tr[40] - 0.5) < 1e-5


def _const_velocity_case(nz=21, nx=31, v=2.0, d=0.02):
    xs = np.arange(nx) * d
    zs = np.arange(nz) * d
    X, Z = np.meshgrid(xs, zs, indexing="ij")  # (nx, nz)
    return X, Z, v


def test_fd_smoothed_angle_approaches_closed_form():
    """Constant velocity: grad T_s is parallel to (x - s, z) and grad T_r to
    (x - r, z), so the opening angle is the angle between those vectors."""
    X, Z, v = _const_velocity_case()
    nx, nz = X.shape
    s, r = 0.1, 0.5
    Ts = (np.hypot(X - s, Z) / v).astype(np.float32).ravel()
    Tr = (np.hypot(X - r, Z) / v).astype(np.float32).ravel()
    gsx, gsz, grx, grz = (np.zeros_like(Ts) for _ in range(4))
    L.kmo_grad_plane_f32(_p(Ts), nz, nx, 0.02, 0.02, 2, _p(gsx), _p(gsz))
    L.kmo_grad_plane_f32(_p(Tr), nz, nx, 0.02, 0.02, 2, _p(grx), _p(grz))
    for g in (gsx, gsz, grx, grz):
        L.kmo_smooth_f32(_p(g), nz, nx, 1)
    ang = np.array([L.kmo_angle_f32(a, b, c, d, 0) for a, b, c, d in zip(gsx, gsz, grx, grz)])
    ex = np.degrees(np.abs(np.arctan2(X - s, Z) - np.arctan2(X - r, Z))).ravel()
    # interior only: the one-sided FD and renormalised smoothing bias the edges
    far = ((Z.ravel() > 0.2) & (Z.ravel() < 0.32) & (np.abs(X.ravel() - 0.3) < 0.2))
    assert np.max(np.abs(ang[far] - ex[far])) < 1.0


def test_f64_oracle_linear_in_data_and_zero_for_zero_data():
    from tests import golden_cases as gc
    kv, run, _ = gc.case("mute_qc")
    d1 = run["data"]
    d2 = np.roll(d1, 7, axis=-1)
    args = (run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"], run["rder"],
            run["r_axes"])
    c1 = O.migrate(kv, d1, *args, precision="f64")[0]
    c2 = O.migrate(kv, d2, *args, precision="f64")[0]
    c12 = O.migrate(kv, (0.5 * d1 - 2.0 * d2).astype(np.float32), *args, precision="f64")[0]
    assert O.rel_l2(c12, 0.5 * c1 - 2.0 * c2) < 1e-6
    c0 = O.migrate(kv, np.zeros_like(d1), *args, precision="f32")[0]
    assert not np.any(c0)


def test_shot_subset_sums_to_full_run_f64():
    from tests import golden_cases as gc
    kv, run, _ = gc.case("angle_base")
    args = (run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"],
            run["rder"], run["r_axes"])
    full = O.migrate(kv, *args, precision="f64")[0]
    a = O.migrate(kv, *args, shots=(0, 3), precision="f64")[0]
    b = O.migrate(kv, *args, shots=(3, 6), precision="f64")[0]
    assert O.rel_l2(a + b, full) < 1e-12

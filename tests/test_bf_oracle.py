"""The segmented-butterfly oracle (oracle/kbf_oracle.c, kirchdat2d_freq's
use_bf=1 operator, SERECKIRCH/src/bf1d.cpp) pinned.

* Against fixtures written by the reference's own bf1d.cpp compiled from
  source (tests/golden/make_bf_golden.py): the panel list (positions, sizes
  and the butterfly / exact choice) exactly, the applied vectors to 1e-6
  relative L2 (they agree bitwise where this container built both).
* Against the compiled reference directly, on random operators, when
  oracle/_ref holds it.
* By closed forms: at omega = 0 with a bilinear amplitude every stage of the
  factor interpolates exactly, and a separable traveltime with a constant
  amplitude is a rank-one kernel whose phases cancel, so in both cases the
  butterfly reproduces the exact operator to float rounding; the Lagrange
  basis reproduces polynomials of degree < p; the panels tile the active
  entries once each.
* The datuming oracle's use_bf=1 path: with no legal butterfly panel
  (bf_panel_levels=0) it is the direct path to float rounding; with the
  defaults on a 64-trace geometry it departs from the direct path by the
  butterfly's own small error.
"""
import json
import os

import numpy as np
import pytest

from oracle import bf
from oracle import datum as D
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
INFO = json.load(open(os.path.join(HERE, "golden", "bf_cases.json")))
GOLD = np.load(os.path.join(HERE, "golden", "bf.npz"))
TOL = 1e-6


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("name", sorted(INFO["cases"]))
def test_against_the_reference_fixtures(name):
    c = INFO["cases"][name]
    tau, amp, u = GOLD[name + "/tau"], GOLD[name + "/amp"], GOLD[name + "/u"]
    ref = GOLD[name + "/out"]
    if c["kind"] == "strict":
        out = bf.strict_apply(tau, amp, c["omega"], u, p=c["p"], leaf=c["leaf"])
    else:
        par = {k: c[k] for k in ("p", "leaf", "levels", "amp_eps", "phase_tol")}
        out, pan = bf.seg_apply(tau, amp, c["omega"], u, **par)
        assert np.array_equal(pan, GOLD[name + "/panels"])
    assert np.abs(ref).max() > 0
    assert _rel(out, ref) <= TOL, (name, _rel(out, ref))


def test_fixtures_hold_butterfly_panels():
    """Every segmented fixture exercises the interpolated panels, and the
    reference's own self-test geometry stays close to the exact operator
    (bf1d.cpp:1055-1107 prints this error)."""
    for name, c in INFO["cases"].items():
        if c["kind"] == "segmented":
            assert (GOLD[name + "/panels"][:, 4] == bf.BUTTERFLY).any(), name
    tau, amp, om, u = bf.selftest_case()
    ex = bf.exact_apply(tau, amp, om, u)
    assert 1e-7 < _rel(GOLD["selftest/out"], ex) < 1e-3


def _smooth_case(rng, n):
    r = np.arange(n, dtype=np.float64)
    tau = (0.4 + 0.2 * np.sqrt(0.09 + ((r[:, None] - r[None, :]) * 0.01) ** 2)
           + 0.002 * np.sin(0.07 * r[None, :])).astype(np.float32)
    amp = ((1.0 + 0.2 * rng.standard_normal((n, n))) *
           np.exp(1j * 0.2 * rng.standard_normal((n, n)))).astype(np.complex64)
    return tau, amp


@pytest.mark.skipif(bf.ref_lib() is None, reason="oracle/_ref/libbf1d_ref.so not built")
def test_against_the_compiled_reference_on_random_operators():
    rng = np.random.default_rng(11)
    for n, leaf, p, lev, aper, om in ((96, 16, 12, 1, 50, 10.0), (128, 16, 12, 2, 127, 20.0),
                                      (80, 8, 6, 2, 30, 35.0), (64, 16, 15, 1, 63, 5.0),
                                      (200, 16, 12, 3, 150, 8.0)):
        tau, amp = _smooth_case(rng, n)
        d = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
        amp[d > aper] = 0
        amp[rng.random((n, n)) < 0.0003] = 0  # isolated zeros split panels
        u = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
        kw = dict(p=p, leaf=leaf, levels=lev)
        o, po = bf.seg_apply(tau, amp, om, u, **kw)
        r, pr = bf.seg_apply(tau, amp, om, u, ref=True, **kw)
        assert np.array_equal(po, pr), (n, leaf, p, lev)
        assert (po[:, 4] == bf.BUTTERFLY).any()
        assert _rel(o, r) <= TOL


def test_bilinear_amplitude_at_zero_frequency_is_reproduced():
    n = 128
    r = np.arange(n, dtype=np.float32)
    amp = ((1.0 + 0.01 * r[:, None] + 0.02 * r[None, :]) * (1 + 0.5j)).astype(np.complex64)
    rng = np.random.default_rng(3)
    tau = rng.random((n, n)).astype(np.float32)  # irrelevant at omega = 0
    u = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
    for lev in (1, 2, 3):
        out, pan = bf.seg_apply(tau, amp, 0.0, u, levels=lev)
        assert (pan[:, 4] == bf.BUTTERFLY).all() and len(pan) > 0
        assert _rel(out, bf.exact_apply(tau, amp, 0.0, u)) < 1e-5


def test_separable_phase_rank_one_kernel_is_reproduced():
    n = 128
    r = np.arange(n, dtype=np.float64)
    tau = (0.3 + 0.002 * r[:, None] + 0.0005 * (r[None, :] - 40.0) ** 2 / 40).astype(np.float32)
    amp = np.full((n, n), 0.7 - 0.2j, np.complex64)
    rng = np.random.default_rng(4)
    u = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
    for om in (5.0, 60.0, 400.0):
        out, pan = bf.seg_apply(tau, amp, om, u, levels=2)
        # a separable float tau leaves a rounding-level residual only
        assert (pan[:, 4] == bf.BUTTERFLY).all()
        assert _rel(out, bf.exact_apply(tau, amp, om, u)) < 1e-4


def test_lagrange_basis_reproduces_polynomials():
    rng = np.random.default_rng(5)
    for p in (3, 7, 12):
        nodes = 10.0 + 6.0 * np.cos((2 * np.arange(p) + 1) * np.pi / (2 * p))
        coef = rng.standard_normal(p)
        P = lambda x: np.polyval(coef, (x - 10.0) / 6.0)  # degree p - 1 on the box
        for x in (4.3, 10.0, 15.99, nodes[2]):
            b = bf.lagrange(nodes, x)
            assert abs(b.sum() - 1.0) < 1e-12
            assert abs(b @ P(nodes) - P(x)) < 1e-10
        assert np.array_equal(bf.lagrange(nodes, nodes[1]), np.eye(p)[1])


def test_panels_tile_the_active_entries_once():
    rng = np.random.default_rng(6)
    n = 150
    tau, amp = _smooth_case(rng, n)
    d = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
    amp[d > 70] = 0
    u = np.ones((1, n), np.complex64)
    for lev in (0, 1, 2, -1):
        _, pan = bf.seg_apply(tau, amp, 12.0, u, levels=lev)
        cover = np.zeros((n, n), int)
        for r0, c0, nr, nc, kind in pan:
            cover[r0:r0 + nr, c0:c0 + nc] += 1
            if kind == bf.BUTTERFLY:
                assert nr == nc and (amp[r0:r0 + nr, c0:c0 + nc] != 0).all()
        assert cover.max() == 1
        assert (cover[amp != 0] == 1).all()
        if lev == 0:
            assert (pan[:, 4] == bf.EXACT).all()  # leaf-sized panels are never legal


def test_invalid_parameters_are_rejected():
    tau = np.zeros((32, 32), np.float32)
    amp = np.ones((32, 32), np.complex64)
    with pytest.raises(ValueError):
        bf.seg_apply(tau, amp, 1.0, np.ones((1, 32)), p=16, leaf=16)
    with pytest.raises(ValueError):
        bf.seg_apply(tau, amp, 1.0, np.ones((1, 32)), p=0)


# ---------------------------------------------------------------- datuming --

DINFO = json.load(open(os.path.join(HERE, "golden", "datum_cases.json")))
DGOLD = np.load(os.path.join(HERE, "golden", "datum.npz"))
M = DINFO["model"]
SDATUM = M["should_datum"] * M["d"]
POS = (0.0, M["d"])


def _datum_case(ns=64, nh=64, nt=256, seed=7):
    """Random traces of 256 samples (the two datum delays add up to about
    0.35 s, so shorter traces keep only wrapped, cancelling energy) on the
    64-position table, shots and receivers from position 0."""
    rng = np.random.default_rng(seed)
    data = rng.normal(size=(ns, nh, nt)).astype(np.float32)
    ax = [[nt, 0.0, 0.004], [nh, 0.0, M["d"]], [ns, 0.0, M["d"]]]
    return data, ax


@pytest.mark.parametrize("cmp", [0, 1])
def test_datum_without_legal_panels_is_the_direct_path(cmp):
    data, ax = _datum_case(40, 40) if cmp == 0 else _datum_case(20, 40)
    g = DGOLD["green"]
    kv = dict(cmp=cmp, aperture=39, taper=3, antialias=1.0)
    a = D.datum(dict(kv, use_bf=1, bf_panel_levels=0), data, ax, g, POS, g, POS, SDATUM)
    b = D.datum(dict(kv, use_bf=0), data, ax, g, POS, g, POS, SDATUM)
    assert np.abs(b).max() > 0
    assert O.rel_l2(a, b) < 1e-6


def test_datum_butterfly_departs_from_the_direct_path_by_its_own_error():
    data, ax = _datum_case()
    g = DGOLD["green"]
    kv = dict(cmp=0, aperture=63, taper=0, antialias=1.0)
    a, am = D.datum(dict(kv, use_bf=1), data, ax, g, POS, g, POS, SDATUM, interm=True)
    b, bm = D.datum(dict(kv, use_bf=0), data, ax, g, POS, g, POS, SDATUM, interm=True)
    for x, y in ((a, b), (am, bm)):
        r = O.rel_l2(x, y)
        assert 1e-6 < r < 1e-2, r
    with pytest.raises(ValueError):
        D.datum(dict(kv, use_bf=1), data, ax, g, POS, g, POS, SDATUM, precision="f64")

"""GPU butterfly datuming (use_bf=1) against the oracle's restatement
of the reference's segmented butterfly (oracle/kbf_oracle.c, pinned bitwise
to SERECKIRCH/src/bf1d.cpp compiled from source, tests/test_bf_oracle.py).

The GPU adds (B_P - K_P) u over the panels the reference factors at each
frequency to its exact sums (banded for cmp=0, pair lists for cmp=1, where
the matrices are per shot on the receiver side and per common-receiver
gather on the source side).  Held, per geometry:
* the output and the receiver-side cube against the float oracle with
  use_bf=1: relative L2 <= 2e-5 (the exact path's bound against the FP64
  oracle), and the butterfly moves the output by at least ten times the
  GPU's distance from the oracle;
* the butterfly's own contribution, GPU(use_bf=1) - GPU(use_bf=0) against
  oracle(use_bf=1) - oracle(use_bf=0): relative L2 <= 2e-2 (the difference
  is 2e-5 .. 1e-3 of the output, so float rounding of the outputs is a few
  1e-3 of it);
* repeat runs bitwise equal.
Geometries: the workflow's keys on the salt model (128 shots x 128 receivers
x 1001 samples, aperture 200, data_taper 20, antialias 1: bf_p 12, leaf 16,
one panel level), and on the 64-position golden table leaf 8 with p 5 or 6
and panel levels 1-3 (panels of 16, 32 and 64 points), a narrow aperture
(panels at the band's edge stay exact) and a 50 x 60 cube (panels cut by the
matrix edge); cmp=1 with leaf 8 (panels on both sides), with panel levels 2
(source-side gathers only) and with shots at half the offset spacing (jump
2: receiver side only).
"""
import json
import os
import zlib

import numpy as np
import pytest

from oracle import datum as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INFO = json.load(open(os.path.join(HERE, "golden", "datum_cases.json")))
GOLD = np.load(os.path.join(HERE, "golden", "datum.npz"))
M = INFO["model"]
SDATUM = M["should_datum"] * M["d"]
POS = (0.0, M["d"])
TOL = 2e-5
DTOL = 2e-2

CASES = {
    "levels1_leaf8": (64, 64, dict(cmp=0, aperture=63, taper=0, antialias=1.0, data_taper=2,
                                   bf_n_leaf=8, bf_p=5, bf_panel_levels=3)),
    "levels2_leaf8": (64, 64, dict(cmp=0, aperture=63, taper=3, antialias=1.0, bf_n_leaf=8,
                                   bf_p=6, bf_panel_levels=2)),
    "narrow": (64, 64, dict(cmp=0, aperture=30, taper=2, antialias=1.0, bf_n_leaf=8, bf_p=5)),
    "uneven": (50, 60, dict(cmp=0, aperture=63, taper=0, antialias=0.5, bf_n_leaf=8, bf_p=5,
                            bf_panel_levels=2)),
}

# name: (ns, ds, nh, dh, keys) on the 64-position table, shots and offsets from 0
CMP1 = {
    "cmp1_leaf8": (20, 1.0, 40, 1.0, dict(cmp=1, aperture=39, taper=0, antialias=1.0,
                                          bf_n_leaf=8, bf_p=5)),
    "cmp1_levels2": (36, 1.0, 28, 1.0, dict(cmp=1, aperture=35, taper=2, antialias=1.0,
                                            bf_n_leaf=8, bf_p=6, bf_panel_levels=2)),
    "cmp1_jump2": (24, 0.5, 40, 1.0, dict(cmp=1, aperture=20, taper=0, antialias=1.0,
                                          bf_n_leaf=8, bf_p=5)),
}


def _dat():
    from kirch_datuming_amd import dat
    return dat


def _check(name, kv, data, ax, sg, sax, datum, sides=(True, True)):
    dat = _dat()
    print(f"datum butterfly {name}: GPU use_bf=1 and 0", flush=True)
    out, mid, st = dat.datum(dict(kv, use_bf=1), data, ax, sg, sax, sg, sax, datum, interm=True,
                             stats=True)
    again = dat.datum(dict(kv, use_bf=1), data, ax, sg, sax, sg, sax, datum)
    ex, exm = dat.datum(dict(kv, use_bf=0), data, ax, sg, sax, sg, sax, datum, interm=True)
    assert st["butterfly"] == 1
    assert (st["bf_panels_receiver"] > 0, st["bf_panels_source"] > 0) == sides
    assert np.array_equal(out.view(np.uint32), again.view(np.uint32))
    print(f"datum butterfly {name}: oracle use_bf=1 and 0", flush=True)
    ref, rmid = D.datum(dict(kv, use_bf=1), data, ax, sg, sax, sg, sax, datum, interm=True)
    rex, rexm = D.datum(dict(kv, use_bf=0), data, ax, sg, sax, sg, sax, datum, interm=True)
    r_out, r_mid = O.rel_l2(out, ref), O.rel_l2(mid, rmid)
    effect = O.rel_l2(ref, rex)
    d_out = O.rel_l2(out - ex, ref - rex)
    d_mid = O.rel_l2(mid - exm, rmid - rexm)
    print(f"datum butterfly {name}: {st['bf_panels_receiver']}+{st['bf_panels_source']} panels, "
          f"{st['bf_items_receiver']}+{st['bf_items_source']} items, "
          f"{st['ms_bf_receiver']:.2f}+{st['ms_bf_source']:.2f} ms; relL2 out {r_out:.2e} "
          f"interm {r_mid:.2e}; butterfly effect {effect:.2e}, its GPU-vs-oracle error "
          f"out {d_out:.2e} interm {d_mid:.2e}")
    assert r_out <= TOL and r_mid <= TOL
    assert effect > 10.0 * max(r_out, r_mid)  # the panels matter at this distance
    assert d_out <= DTOL
    if sides[0]:  # without receiver-side panels the interm cube's difference is rounding only
        assert d_mid <= DTOL


@pytest.mark.parametrize("name", sorted(CASES))
def test_butterfly_against_the_oracle(name):
    ns, nh, kv = CASES[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    nt = 256  # holds the datumed energy of the 64-position table
    data = rng.normal(size=(ns, nh, nt)).astype(np.float32)
    ax = [[nt, 0.0, 0.004], [nh, 0.0, M["d"]], [ns, 0.0, M["d"]]]
    _check(name, kv, data, ax, GOLD["green"], POS, SDATUM)


def test_butterfly_salt_workflow_keys():
    from kirch_datuming_amd import configs
    print("datum butterfly salt: building 128 x 128 x 1001", flush=True)
    w = configs.datum(threads=16, npos=128, nt=1001)
    assert w.kv["use_bf"] == 1 and w.kv["cmp"] == 0
    _check("salt", w.kv, w.data, w.data_axes, w.green, w.pos_axis, w.datum)


@pytest.mark.parametrize("name", sorted(CMP1))
def test_cmp1_butterfly_against_the_oracle(name):
    ns, ds, nh, dh, kv = CMP1[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    nt = 256
    data = rng.normal(size=(ns, nh, nt)).astype(np.float32)
    ax = [[nt, 0.0, 0.004], [nh, 0.0, dh * M["d"]], [ns, 0.0, ds * M["d"]]]
    sides = {"cmp1_leaf8": (True, True), "cmp1_levels2": (False, True),
             "cmp1_jump2": (True, False)}[name]
    _check(name, kv, data, ax, GOLD["green"], POS, SDATUM, sides)

"""Writes tests/golden/bf.npz: the reference's segmented butterfly
(SERECKIRCH/src/bf1d.cpp, compiled from its own source by oracle/ref.mk into
oracle/_ref/libbf1d_ref.so) applied to fixed operators.

Cases (inputs and outputs are both stored):
  selftest           the geometry of bf1d.cpp's own self-test (:1055-1107):
                     n 64, leaf 16, p 12, panel levels 2, amp_eps 0
  selftest_workflow  the same operator with kirchdat2d_freq's defaults
                     (p 12, leaf 16, levels 1, amp_eps 1e-20, phase_tol 2)
  banded_*           a datuming-like operator: traveltimes from a point at
                     depth to a surface line, the amplitude with its bulk
                     phase removed (amp = w exp(i omega tau) as
                     kirchdat2d_freq forms it), an aperture band and
                     antialias-style zeros; low and high frequency, panel
                     levels 1 and 2, and an edge geometry (n not a multiple
                     of the panel)
  strict_L2, _L3     single strict factors of 64 and 128 points
  small_leaf         leaf 8, p 5, levels 2 on 72 points

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_bf_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bf  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")


def banded(n, aper, omega, cut=None, seed=0, z=0.4):
    """Datuming-like operator: tau(r, c) = sqrt(z^2 + ((r - c) dx)^2) / v plus
    a smooth lateral term, amp = w(r, c) exp(i omega tau) inside |r - c| <=
    aper, w zero where |r - c| > cut (the antialias cut of steep pairs)."""
    r = np.arange(n, dtype=np.float64)
    d = r[:, None] - r[None, :]
    dx, v = 0.01, 2.0
    tau = np.sqrt(z * z + (d * dx) ** 2) / v + 0.002 * np.sin(0.05 * r[None, :]) \
        + 0.001 * np.cos(0.03 * r[:, None])
    tau = tau.astype(np.float32)
    w = (1.0 / np.pi * dx * z * tau / (z * z + (d * dx) ** 2)).astype(np.float32)
    w[np.abs(d) > aper] = 0.0
    if cut is not None:
        w[np.abs(d) > cut] = 0.0
    ph = np.float32(omega) * tau
    amp = (w * (np.cos(ph) + 1j * np.sin(ph))).astype(np.complex64)
    tau[w == 0] = 0.0
    rng = np.random.default_rng(seed)
    u = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    return tau, amp, np.float32(omega), u


def smooth(n, omega, seed):
    rng = np.random.default_rng(seed)
    r = np.arange(n, dtype=np.float64)
    tau = (0.5 + 0.3 * np.sqrt(0.04 + ((r[:, None] - r[None, :]) * 0.01) ** 2)
           + 0.001 * r[None, :]).astype(np.float32)
    amp = ((1.0 + 0.1 * rng.standard_normal((n, n))) *
           np.exp(1j * 0.3 * rng.standard_normal((n, n)))).astype(np.complex64)
    u = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
    return tau, amp, np.float32(omega), u


def main():
    if bf.ref_lib() is None:
        raise SystemExit("build the reference first: make -C oracle ref")
    cases, arrays = {}, {}

    def seg(name, tau, amp, omega, u, **kw):
        par = dict(bf.DEFAULTS, **kw)
        out, pan = bf.seg_apply(tau, amp, omega, u, ref=True, **par)
        cases[name] = dict(kind="segmented", omega=float(omega), **par)
        arrays.update({f"{name}/tau": tau, f"{name}/amp": amp, f"{name}/u": u,
                       f"{name}/out": out, f"{name}/panels": pan})
        nb = int((pan[:, 4] == bf.BUTTERFLY).sum())
        print(f"{name}: n {tau.shape[0]}, {len(pan)} panels ({nb} butterfly)")

    tau, amp, om, u = bf.selftest_case()
    seg("selftest", tau, amp, om, u, levels=2, amp_eps=0.0)
    seg("selftest_workflow", tau, amp, om, u)
    for name, n, aper, omega, cut, z, lev in (("banded_low", 128, 60, 15.0, None, 0.4, 1),
                                             ("banded_high", 128, 127, 45.0, 90, 1.5, 1),
                                             ("banded_levels2", 128, 127, 12.0, None, 1.0, 2),
                                             ("banded_edge", 100, 99, 25.0, None, 1.0, 1)):
        tau, amp, om, u = banded(n, aper, omega, cut, seed=n + aper, z=z)
        seg(name, tau, amp, om, u, levels=lev)
    tau, amp, om, u = smooth(72, 12.0, 5)
    seg("small_leaf", tau, amp, om, u, p=5, leaf=8, levels=2)
    for name, n in (("strict_L2", 64), ("strict_L3", 128)):
        tau, amp, om, u = smooth(n, 30.0, n)
        out = bf.strict_apply(tau, amp, om, u, p=12, leaf=16, ref=True)
        cases[name] = dict(kind="strict", omega=float(om), p=12, leaf=16)
        arrays.update({f"{name}/tau": tau, f"{name}/amp": amp, f"{name}/u": u,
                       f"{name}/out": out})
        print(f"{name}: n {n}")
    np.savez_compressed(os.path.join(OUT, "bf.npz"), **arrays)
    with open(os.path.join(OUT, "bf_cases.json"), "w") as f:
        json.dump(dict(source="SERECKIRCH/src/bf1d.cpp compiled by oracle/ref.mk",
                       cases=cases), f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Write the golden fixtures for the fast-marching table source.

Runs the reference's eikods2d, compiled from the reference's own sources into
oracle/_ref (``make -C oracle ref``), on small seeded velocity models and
stores inputs and outputs under tests/golden/:

  eik_cases.json  per case: grid, spacing, the shot rows, key=value options
  eik.npz         per case: velocity and the time, tdl1, tds1, gradz, gradx
                  volumes the program writes

The shot file holds one (depth, lateral, cross-line) row per position, as
eikods2d reads it (src/eikods2d.cpp:276-300).
"""

from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from kirch_datuming_amd import rsf  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
SEED = 20261205


def smooth_random(rng, n2, n1, lo, hi, passes=6):
    v = rng.normal(size=(n2, n1))
    for _ in range(passes):  # separable 1-2-1 smoothing
        v = 0.25 * np.roll(v, 1, 0) + 0.5 * v + 0.25 * np.roll(v, -1, 0)
        v = 0.25 * np.roll(v, 1, 1) + 0.5 * v + 0.25 * np.roll(v, -1, 1)
    v = (v - v.min()) / (v.max() - v.min())
    return (lo + (hi - lo) * v).astype(np.float32)


def salt(n2, n1, d1, d2):
    z = np.arange(n1) * d1
    x = np.arange(n2) * d2
    Z, X = np.meshgrid(z, x)
    v = 1.8 + 0.6 * Z
    body = (X - 0.6) ** 2 / 0.09 + (Z - 0.45) ** 2 / 0.03 < 1.0
    v[body] = 4.5
    return v.astype(np.float32)


def cases(rng):
    f = np.float32
    out = {}
    n2, n1 = 64, 48
    out["random_fractional"] = dict(
        vel=smooth_random(rng, n2, n1, 1.5, 3.5), d1=0.02, d2=0.02,
        shots=[[0.0, 0.013, 0.0], [0.0, 0.4170, 0.0], [0.0, 1.2345, 0.0]],
        opts={"b1": 2, "b2": 2})
    n2, n1 = 70, 50
    out["salt_buried_rect"] = dict(
        vel=salt(n2, n1, 0.015, 0.02), d1=0.015, d2=0.02,
        shots=[[0.1, 0.35, 0.0], [0.2275, 0.61, 0.0], [0.0, 1.2, 0.0]],
        opts={"b1": 2, "b2": 2})
    n2, n1 = 64, 48
    out["order1_box1"] = dict(
        vel=smooth_random(rng, n2, n1, 1.5, 3.0), d1=0.02, d2=0.02,
        shots=[[0.0, 0.3, 0.0], [0.0, 0.9, 0.0]], opts={"order": 1, "b1": 1, "b2": 1})
    out["depth_derivative_box3"] = dict(
        vel=smooth_random(rng, n2, n1, 2.0, 4.0), d1=0.02, d2=0.02,
        shots=[[0.05, 0.5, 0.0], [0.0, 0.77, 0.0]], opts={"l": 0, "b1": 3, "b2": 3})
    n2, n1 = 40, 30
    out["constant"] = dict(
        vel=np.full((n2, n1), 2.0, np.float32), d1=0.025, d2=0.025,
        shots=[[0.0, 0.4875, 0.0], [0.1, 0.2, 0.0]], opts={"b1": 2, "b2": 2})
    out["edges"] = dict(
        vel=smooth_random(rng, n2, n1, 1.6, 2.6), d1=0.025, d2=0.025,
        shots=[[0.0, 0.0, 0.0], [0.0, f(n2 - 1) * f(0.025), 0.0], [0.0, -0.03, 0.0],
               [0.0, 1.01, 0.0], [0.7, 0.5, 0.0]], opts={"b1": 2, "b2": 2})
    return out


def main():
    rng = np.random.default_rng(SEED)
    meta, arrays = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, c in cases(rng).items():
            d = os.path.join(tmp, name)
            os.makedirs(d)
            vel = c["vel"]
            n2, n1 = vel.shape
            rsf.write(os.path.join(d, "vel.rsf"), vel, o=[0.0, 0.0], d=[c["d1"], c["d2"]])
            shots = np.asarray(c["shots"], np.float32)
            rsf.write(os.path.join(d, "shots.rsf"), shots)
            args = [os.path.join(REF, "eikods2d"), "in=vel.rsf", "out=time.rsf",
                    "tdl1=tdl1.rsf", "tds1=tds1.rsf", "shotfile=shots.rsf", "gradz=gz.rsf",
                    "gradx=gx.rsf"] + [f"{k}={v}" for k, v in c["opts"].items()]
            r = subprocess.run(args, cwd=d, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"eikods2d failed on {name}: {r.stdout}\n{r.stderr}")
            arrays[f"{name}_vel"] = vel
            for key, f in (("time", "time"), ("tdl1", "tdl1"), ("tds1", "tds1"),
                           ("gradz", "gz"), ("gradx", "gx")):
                a, _ = rsf.read(os.path.join(d, f + ".rsf"))
                arrays[f"{name}_{key}"] = a.astype(np.float32).reshape(len(shots), n2, n1)
            meta[name] = {"n1": n1, "n2": n2, "d1": c["d1"], "d2": c["d2"],
                          "shots": shots.tolist(), "options": c["opts"]}
    np.savez_compressed(os.path.join(HERE, "eik.npz"), **arrays)
    with open(os.path.join(HERE, "eik_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {len(meta)} eikods2d cases")


if __name__ == "__main__":
    main()

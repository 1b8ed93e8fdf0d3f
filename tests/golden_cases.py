"""Loader for the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_CACHE = {}


def load():
    if not _CACHE:
        with open(os.path.join(G, "cases.json")) as f:
            _CACHE["meta"] = json.load(f)
        _CACHE["inp"] = dict(np.load(os.path.join(G, "inputs.npz")))
        _CACHE["out"] = dict(np.load(os.path.join(G, "outputs.npz")))
    return _CACHE["meta"], _CACHE["inp"], _CACHE["out"]


def case_names():
    return [c["name"] for c in load()[0]["cases"]]


def case(name):
    """Return (kv params without file keys, inputs dict for the run, expected)."""
    meta, inp, out = load()
    c = [c for c in meta["cases"] if c["name"] == name][0]
    kv = dict(c["params"])
    grads = None
    if "sgradx" in kv:
        grads = [inp["gx"], inp["gz"], inp["gx"], inp["gz"]]
        for k in ("sgradx", "sgradz", "rgradx", "rgradz"):
            kv.pop(k)
    for k in ("direct_mute_qc_before", "direct_mute_qc_after"):
        kv.pop(k, None)
    ax = meta["axes"]
    run = dict(data=inp[c["data"]], data_axes=ax[c["data"]], stab=inp["time"],
               sder=inp["tds1"], s_axes=ax["time"], rtab=inp["time"], rder=inp["tds1"],
               r_axes=ax["time"], grads=grads)
    exp = dict(cig=out[name], axes=c["axes"], label3=c.get("label3", ""))
    if c.get("qc"):
        exp["qcb"] = out[name + "__qcb"]
        exp["qca"] = out[name + "__qca"]
    return kv, run, exp

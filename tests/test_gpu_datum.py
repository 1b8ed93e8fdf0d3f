"""GPU datuming (libkmdat through the C ABI) against the oracle and the
reference's time-domain program.

The GPU transforms with rocFFT, folds the antialias weight into one table
row per key and sums in float with the reference's order of summation
indices; against the FP64 oracle (the reference's direct path) the output
and the receiver-side cube are held to relative L2 <= 2e-5, with every
option on (antialias, data-edge taper, cosine aperture taper).  With those
options off the GPU is also held to the reference's own time-domain program
(kirchdat2d_auto_cs, the golden fixtures) at the oracle's 1e-5 + its own
2e-5.  Runs are deterministic (no atomics): repeats are bitwise equal.
"""
import json
import os
import subprocess
import zlib

import numpy as np
import pytest

from oracle import datum as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "kirch_datuming_amd", "bin", "kirchdat2d_freq")
INFO = json.load(open(os.path.join(HERE, "golden", "datum_cases.json")))
GOLD = np.load(os.path.join(HERE, "golden", "datum.npz"))
M = INFO["model"]
SDATUM = M["should_datum"] * M["d"]
POS = (0.0, M["d"])
TOL = 2e-5
NAMES = sorted(INFO["cases"])


def _dat():
    from kirch_datuming_amd import dat
    return dat


def _case(name):
    """A golden geometry on the direct path (use_bf=0)."""
    c = INFO["cases"][name]
    return dict(c["params"], use_bf=0), GOLD[name + "/data"], c["data_axes"]


@pytest.mark.parametrize("name", NAMES)
def test_options_on_against_fp64_oracle(name):
    kv, data, ax = _case(name)
    kv.update(antialias=1.0, data_taper=2, taper=2)
    g = GOLD["green"]
    print(f"datum case {name} (antialias, data_taper, taper on)", flush=True)
    out, mid = _dat().datum(kv, data, ax, g, POS, g, POS, SDATUM, interm=True)
    ref, rmid = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64", interm=True)
    r_out, r_mid = O.rel_l2(out, ref), O.rel_l2(mid, rmid)
    print(f"{name}: out {r_out:.2e}, interm {r_mid:.2e}")
    assert np.abs(ref).max() > 0 and r_out < TOL and r_mid < TOL


@pytest.mark.parametrize("name", NAMES)
def test_options_off_against_the_time_domain_reference(name):
    kv, data, ax = _case(name)
    kv.update(antialias=0, data_taper=0)
    g = GOLD["green"]
    print(f"datum case {name} vs kirchdat2d_auto_cs", flush=True)
    out, mid = _dat().datum(kv, data, ax, g, POS, g, POS, SDATUM, interm=True)
    assert O.rel_l2(mid, GOLD[name + "/interm"]) < 1e-5 + TOL
    assert O.rel_l2(out, GOLD[name + "/out"]) < 1e-5 + TOL


@pytest.mark.parametrize("name", ["cmp0", "cmp0_wide"])
def test_matrix_core_kernel_against_fp64_oracle(name, monkeypatch):
    """The banded operator on the matrix cores (KMDAT_MFMA=1, read when a
    context is created) against the same FP64 oracle."""
    monkeypatch.setenv("KMDAT_MFMA", "1")
    kv, data, ax = _case(name)
    kv.update(antialias=1.0, data_taper=2, taper=2)
    g = GOLD["green"]
    print(f"datum case {name} on the matrix cores", flush=True)
    out, mid = _dat().datum(kv, data, ax, g, POS, g, POS, SDATUM, interm=True)
    ref, rmid = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64", interm=True)
    assert O.rel_l2(out, ref) < TOL and O.rel_l2(mid, rmid) < TOL


EDGE = {
    # name: (ns, s0, ds, nh, h0, dh, nt, keys); 256 samples hold the datumed
    # energy (the two datum delays add up to about 0.35 s)
    "one_shot": (1, 0.5, 0.05, 12, 0.2, 0.05, 256, {"cmp": 0, "aperture": 4, "taper": 0}),
    "one_receiver": (9, 0.3, 0.05, 1, 0.6, 0.05, 256, {"cmp": 0, "aperture": 3, "taper": 0}),
    "aperture_0": (6, 0.3, 0.05, 7, 0.3, 0.05, 256, {"cmp": 0, "aperture": 0, "taper": 0}),
    # the default taper (10) gives a lone pair at the aperture edge weight 0
    "aperture_0_tapered": (6, 0.3, 0.05, 7, 0.3, 0.05, 256, {"cmp": 0, "aperture": 0}),
    "short_traces": (5, 0.3, 0.05, 6, 0.3, 0.05, 7, {"cmp": 0, "aperture": 2}),
    "taper_over_aperture": (8, 0.3, 0.05, 9, 0.3, 0.05, 256,
                            {"cmp": 0, "aperture": 3, "taper": 9, "data_taper": 50}),
    "cmp1_tie_jump": (10, 0.3, 0.025, 8, 0.0, 0.05, 256, {"cmp": 1, "aperture": 3}),
    "cmp1_tie_coarse": (8, 0.05, 0.1, 14, 0.0, 0.05, 256, {"cmp": 1, "aperture": 6}),
}


@pytest.mark.parametrize("name", sorted(EDGE))
def test_edge_geometries_against_the_oracle(name):
    """Degenerate axes, zero aperture, traces shorter than the filter, tapers
    wider than the aperture, and cmp=1 geometries whose regrouping indices
    fall on exact .5 ties (the GPU plans them with the reference's float
    arithmetic, as the float oracle does: same traces, so the two agree to
    rounding)."""
    ns, s0, ds, nh, h0, dh, nt, kv = EDGE[name]
    kv = dict(kv, use_bf=0)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    data = rng.normal(size=(ns, nh, nt)).astype(np.float32)
    ax = [[nt, 0.0, 0.004], [nh, h0, dh], [ns, s0, ds]]
    g = GOLD["green"]
    print(f"datum edge case {name}", flush=True)
    out = _dat().datum(kv, data, ax, g, POS, g, POS, SDATUM)
    ref32 = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f32")
    assert np.isfinite(out).all()
    scale = float(np.abs(data).max())
    if np.abs(ref32).max() > 1e-3 * scale:
        assert O.rel_l2(out, ref32) < TOL
    else:
        # (almost) all energy is delayed out of the traces: what is left is
        # rounding-level, held in absolute terms
        assert np.abs(out - ref32).max() <= 1e-6 * scale


def test_repeat_runs_are_bitwise_and_stats():
    dat = _dat()
    kv, data, ax = _case("cmp0")
    g = GOLD["green"]
    cfg = dat.config(dict(kv, data_taper=1), SDATUM, interm=True)
    with dat.Datum(cfg) as dm:
        dm.set_green(0, g, *POS)
        dm.set_green(1, g, *POS)
        dm.set_data(data, ax)
        dm.run()
        dm.sync()
        a, ia = dm.output(), dm.interm()
        dm.run()
        dm.sync()
        b, ib = dm.output(), dm.interm()
        st = dm.stats()
        # new data through the same plan: linear
        dm.set_data(2.0 * data, ax)
        dm.run()
        dm.sync()
        c = dm.output()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))
    assert np.array_equal(c, 2.0 * a)
    assert st["banded"] == 1 and st["nfft"] >= 256 and st["nw"] == st["nfft"] // 2 + 1
    assert st["pairs_receiver"] > 0 and st["pairs_source"] > 0 and st["ms_total"] > 0
    assert abs(st["cmacs"] - (st["pairs_receiver"] + st["pairs_source"]) * st["nw"]) < 1


def test_table_errors_carry_the_reference_messages():
    dat = _dat()
    g = GOLD["green"]
    data = np.zeros((4, 6, 64), np.float32)
    with pytest.raises(dat.DatError, match="Receiver table too small."):
        dat.datum({"cmp": 0}, data, [[64, 0.0, 0.004], [6, 1.5, 0.05], [4, 0.1, 0.05]],
                  g, POS, g, POS, SDATUM)
    with pytest.raises(dat.DatError, match="Source table too small."):
        dat.datum({"cmp": 0}, data, [[64, 0.0, 0.004], [6, 0.1, 0.05], [4, 1.5, 0.05]],
                  g, POS, g, POS, SDATUM)


def test_cli_end_to_end(tmp_path):
    from kirch_datuming_amd import dat, rsf
    kv, data, ax = _case("cmp1_split")
    (nt, _, dt), (nh, h0, dh), (ns, s0, ds) = ax
    rsf.write(str(tmp_path / "in.rsf"), data, o=[0.0, h0, s0], d=[dt, dh, ds])
    g = GOLD["green"]
    # as greensol_auto writes it: n1 = positions, n2 = 1, n1 x n1 samples
    rsf.write(str(tmp_path / "g.rsf"), g.reshape(-1), o=[0.0], d=[M["d"]])
    hdr = (tmp_path / "g.rsf").read_text().replace(f"n1={g.size}", f"n1={g.shape[0]}\n\tn2=1")
    (tmp_path / "g.rsf").write_text(hdr)
    rsf.write(str(tmp_path / "model.rsf"), np.ones((M["nx"], M["nz"]), np.float32),
              o=[0.0, 0.0], d=[M["d"], M["d"]], extra={"should_datum": M["should_datum"]})
    args = [CLI, "input_file=in.rsf", "output_file=out.rsf", "sgreen_file=g.rsf",
            "rgreen_file=g.rsf", "model_file=model.rsf", "interm=mid.rsf", "verb=1"]
    args += [f"{k}={v}" for k, v in kv.items()]
    print("cli kirchdat2d_freq cmp1_split", flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out, h = rsf.read(str(tmp_path / "out.rsf"))
    mid, _ = rsf.read(str(tmp_path / "mid.rsf"))
    assert out.shape == (ns, nh, nt) and h.axes[0].o == 0.0
    ref, rmid = dat.datum(kv, data, ax, g, POS, g, POS, SDATUM, interm=True)
    assert np.array_equal(out, ref) and np.array_equal(mid, rmid)

"""The drop-in program end to end on the GPU: kirch_datuming_amd/bin/
kirchmig2d_cig adj=1 on RSF files, as the reference's workflow calls it
(run_salt_adcig.sh:131-149).

Inputs are the golden fixtures (the reference's own eikods2d tables and the
seeded data cubes) written as RSF files; outputs are read back from the
files the program writes.  Checked against the reference's golden outputs:
the migration= data (relL2 <= 1e-4, 3e-4 for diff=1, as the library tests),
its header (n/o/d of every axis; label1/2/3, unit1 copied from the data's
unit1 and unit3, as src/kirchmig2d_cig.cpp:876-905 writes them) and the
direct-mute QC gathers (before the mute bitwise, after it to 1e-6).  The
workflow's own key set, which no golden case holds, is held to the FP64
oracle the same way."""
import os
import subprocess

import numpy as np
import pytest

from kirch_datuming_amd import rsf
from oracle import oracle as O
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kirch_datuming_amd", "bin", "kirchmig2d_cig")


def _write_inputs(wd, dname):
    meta, inp, _ = gc.load()
    ax = meta["axes"]
    lab = ["Time", "Receiver" if dname == "d0" else "Offset", "Shot"]
    for name in (dname, "time", "tds1", "gx", "gz"):
        a = ax[name]
        rsf.write(os.path.join(wd, f"{name}.rsf"), inp[name], o=[q[1] for q in a],
                  d=[q[2] for q in a], labels=lab if name == dname else None,
                  units=["s"] if name == dname else None)


def _run_cli(wd, dname, params):
    args = [CLI, f"seismic_data={dname}.rsf", "migration=mig.rsf", "stable=time.rsf",
            "sderiv=tds1.rsf", "rtable=time.rsf", "rderiv=tds1.rsf"]
    args += [f"{k}={v}" for k, v in params.items()]
    print("cli:", " ".join(os.path.basename(a) for a in args), flush=True)
    r = subprocess.run(args, cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mig, h = rsf.read(os.path.join(wd, "mig.rsf"))
    return mig, h


def _check_header(h, axes, label3, angle):
    assert h.ndim == 3
    for k, (n, o, d) in enumerate(axes):
        assert h.axes[k].n == n
        assert abs(h.axes[k].o - o) <= 1e-6 * max(1.0, abs(o))
        assert abs(h.axes[k].d - d) <= 1e-6 * max(1.0, abs(d))
    assert h.extra.get("label1") == "Depth" and h.extra.get("label2") == "Lateral"
    assert h.extra.get("unit1") == "s"  # copied from the data's unit1, as the reference does
    if label3:
        assert h.extra.get("label3") == label3
    if angle:
        assert h.extra.get("unit3") == "degree"


@pytest.mark.parametrize("name", ["mute_qc", "angle_cmp1", "half", "gradfiles", "salt_like",
                                  "offset_cig", "stack", "default_axis"])
def test_cli_reproduces_the_golden_case(tmp_path, name):
    meta = gc.load()[0]
    c = [c for c in meta["cases"] if c["name"] == name][0]
    kv, run, exp = gc.case(name)
    wd = str(tmp_path)
    _write_inputs(wd, c["data"])
    mig, h = _run_cli(wd, c["data"], dict(adj=1, **c["params"]))
    assert mig.shape == exp["cig"].shape
    angle = c["params"].get("mode") == "angle"
    _check_header(h, exp["axes"], exp["label3"], angle)
    ref64 = O.migrate(kv, run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"],
                      run["rtab"], run["rder"], run["r_axes"], grads=run["grads"],
                      precision="f64")[0]
    lim = 3e-4 if kv.get("diff") else 1e-4
    rref, r64 = O.rel_l2(mig, exp["cig"]), O.rel_l2(mig, ref64)
    print(f"cli {name}: vs reference {rref:.2e}, vs FP64 {r64:.2e}")
    assert rref <= lim and r64 <= lim
    if c.get("qc"):
        qcb = rsf.read(os.path.join(wd, "qcb.rsf"))[0]
        qca = rsf.read(os.path.join(wd, "qca.rsf"))[0]
        assert np.array_equal(qcb, exp["qcb"])
        assert np.max(np.abs(qca - exp["qca"])) <= 1e-6 * np.max(np.abs(exp["qca"]))


def test_cli_with_the_workflow_key_set(tmp_path):
    """Every key the workflow passes (run_salt_adcig.sh:131-149), with a mute,
    QC gathers, angle taper, anti-alias filter, cone and trace apertures."""
    params = dict(adj=1, cig=1, mode="angle", angle_min=0, angle_max=60, angle_step=2,
                  angle_n=31, angle_interp=1, angle_taper_start=40, angle_taper_end=60,
                  direct_mute=1, direct_mute_invert=0, direct_mute_slope=0.5,
                  direct_mute_intercept=0.05, direct_mute_taper=0.1, direct_mute_qc_shot=2,
                  direct_mute_qc_before="qcb.rsf", direct_mute_qc_after="qca.rsf", cmp=0,
                  tau=0.02, antialias=1, aperture=60, aperture_trace=12)
    wd = str(tmp_path)
    _write_inputs(wd, "d0")
    mig, h = _run_cli(wd, "d0", params)
    kv = {k: v for k, v in params.items() if not k.startswith("direct_mute_qc_") or
          k == "direct_mute_qc_shot"}
    _, run, _ = gc.case("angle_base")
    args = (run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"],
            run["rder"], run["r_axes"])
    ref64, qcb64, qca64 = O.migrate(kv, *args, precision="f64")
    ref32, qcb32, qca32 = O.migrate(kv, *args, precision="f32")
    _check_header(h, [list(a) for a in run["s_axes"][:2]] + [[31, 0.0, 2.0]], "Scattering angle",
                  True)
    r64 = O.rel_l2(mig, ref64)
    print(f"cli workflow keys: vs FP64 {r64:.2e}, float oracle vs FP64 {O.rel_l2(ref32, ref64):.2e}")
    assert np.abs(ref64).max() > 0 and r64 <= 1e-4
    qcb = rsf.read(os.path.join(wd, "qcb.rsf"))[0]
    qca = rsf.read(os.path.join(wd, "qca.rsf"))[0]
    assert np.array_equal(qcb, qcb32)
    assert np.max(np.abs(qca - qca32)) <= 1e-6 * np.max(np.abs(qca32))

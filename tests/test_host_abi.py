"""Host side of the C ABI (no GPU needed): library load, exported symbols,
key=value parsing against the reference's output axes, error messages, the
CLI's argument handling and the RSF helpers."""
import os
import re
import subprocess

import numpy as np
import pytest

from kirch_datuming_amd import kmig, rsf
from oracle import oracle as O
from tests import golden_cases as gc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kirch_datuming_amd", "bin", "kirchmig2d_cig")


def _header_functions(name="kmig.h", prefix="kmig_"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from kirch_datuming_amd import nn
    L = kmig.lib()
    names = _header_functions()
    assert set(names) == set(kmig.EXPORTS)
    nn_names = _header_functions("kmig_nn.h")
    assert set(nn_names) == set(nn.EXPORTS)
    for n in names + nn_names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.kmig_version()
    from kirch_datuming_amd import eik
    E = eik.lib()
    eik_names = _header_functions("kmig_eik.h")
    assert set(eik_names) == set(eik.EXPORTS)
    for n in eik_names:
        assert hasattr(E, n), n
    from kirch_datuming_amd import dat
    Dl = dat.lib()
    dat_names = _header_functions("kmig_dat.h", "kmdat_")
    assert set(dat_names) == set(dat.EXPORTS)
    for n in dat_names:
        assert hasattr(Dl, n), n
    assert set(os.listdir(os.path.join(ROOT, "include"))) == {"kmig.h", "kmig_nn.h",
                                                               "kmig_eik.h", "kmig_dat.h"}


@pytest.mark.parametrize("name", gc.case_names())
def test_params_match_oracle_and_reference_axes(name):
    kv, run, exp = gc.case(name)
    ax = run["data_axes"]
    p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    q = O.params_from_kv(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    for f, _ in O.KmoParams._fields_:
        mine = getattr(p, "diff" if f == "isdiff" else f)
        assert np.float32(mine) == np.float32(getattr(q, f)), f
    if p.cig:
        n3, o3, d3 = exp["axes"][2]
        assert (p.ng, np.float32(p.g0), np.float32(p.dg)) == (n3, np.float32(o3), np.float32(d3))


def test_defaults_are_the_reference_defaults():
    p = kmig.parse_params({}, 10, 0.0, 1.0)
    assert (p.adj, p.cig, p.cmp, p.angle_interp, p.angle_grad_radius, p.angle_smooth) == (
        1, 0, 1, 1, 2, 2)
    assert (p.aperture, p.aperture_taper, p.antialias, p.tau) == (90.0, 10.0, 1.0, 0.0)
    assert (p.oazmin, p.oazmax, p.oaztaper) == (80.0, 50.0, 10.0)
    assert p.aperture_trace == -1 and p.interp_type == ord("h") and p.ng == 1
    p = kmig.parse_params({"mode": "ANGLE", "half_angle": 1}, 10, 0.0, 1.0)
    assert (p.cig, p.ng, p.g0, p.dg, p.oazmin, p.oazmax) == (1, 91, 0.0, 1.0, 40.0, 25.0)
    p = kmig.parse_params(["mode=angle", "angle_step=2", "angle_step=0.5", "angle_max=10"],
                          10, 0.0, 1.0)
    assert (p.ng, p.dg) == (21, 0.5)  # later keys win
    p = kmig.parse_params({"angle_grad_radius": 0, "angle_smooth": -3,
                           "aperture_trace_taper": -1, "cmp": "0.9"}, 3, 0, 1)
    assert (p.angle_grad_radius, p.angle_smooth, p.aperture_trace_taper, p.cmp) == (1, 0, 0, 0)
    for name, code in kmig.PRECISION.items():
        assert kmig.parse_params({"precision": name.upper()}, 3, 0, 1).precision == code
    with pytest.raises(kmig.KmigError, match="Supported are fp32, fp16, bf16, u16"):
        kmig.parse_params({"precision": "fp8"}, 3, 0, 1)


def test_parameter_errors_carry_reference_messages():
    with pytest.raises(kmig.KmigError, match="Unknown CIG mode: depth"):
        kmig.parse_params({"mode": "depth"}, 3, 0, 1)
    with pytest.raises(kmig.KmigError, match="Unknown interpolation type: cubic"):
        kmig.parse_params({"type": "cubic"}, 3, 0, 1)
    with pytest.raises(kmig.KmigError, match="CIG axis step must be positive"):
        kmig.parse_params({"mode": "angle", "angle_step": 0}, 3, 0, 1)
    with pytest.raises(kmig.KmigError, match="CIG axis sample number must be positive"):
        kmig.parse_params({"mode": "angle", "angle_n": 0}, 3, 0, 1)
    base = {"direct_mute": 1}
    for extra, msg in (({"direct_mute_slope": -1}, "direct_mute_slope must be finite"),
                       ({"direct_mute_taper": 0}, "direct_mute_taper must be finite and positive"),
                       ({"direct_mute_invert": 2}, "direct_mute_invert must be 0 or 1"),
                       ({"direct_mute_qc_shot": -2}, "direct_mute_qc_shot must be -1"),
                       ({"direct_mute_qc_shot": 5}, r"outside the input shot range \[0,4\]")):
        p = kmig.parse_params(dict(base, **extra), 3, 0, 1)
        with pytest.raises(kmig.KmigError, match=msg):
            kmig.check_params(p, 5)


def _write_case(tmp, name="angle_base"):
    kv, run, exp = gc.case(name)
    ax = run["data_axes"]
    rsf.write(str(tmp / "d.rsf"), run["data"], o=[a[1] for a in ax], d=[a[2] for a in ax])
    sax = run["s_axes"]
    for nm, arr in (("t", run["stab"]), ("s", run["sder"])):
        rsf.write(str(tmp / f"{nm}.rsf"), arr, o=[a[1] for a in sax], d=[a[2] for a in sax])
    rsf.write(str(tmp / "bad.rsf"), run["sder"][:-1], o=[a[1] for a in sax],
              d=[a[2] for a in sax])
    return kv


def _cli(tmp, *args):
    return subprocess.run([CLI, *args], cwd=tmp, capture_output=True, text=True)


def test_cli_argument_errors(tmp_path):
    _write_case(tmp_path)
    r = _cli(tmp_path, "migration=m.rsf")
    assert r.returncode == 1 and "ERROR: Need seismic_data=" in r.stderr
    base = ["seismic_data=d.rsf", "migration=m.rsf", "stable=t.rsf", "rtable=t.rsf",
            "rderiv=s.rsf"]
    r = _cli(tmp_path, *base)
    assert r.returncode == 1 and "Need sderiv=" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=bad.rsf")
    assert r.returncode == 1 and "Dimension mismatch between stable= and sderiv=" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "mode=angle", "sgradx=t.rsf")
    assert r.returncode == 1 and "provide all of sgradx=" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "direct_mute=1", "direct_mute_qc_shot=1")
    assert r.returncode == 1 and "requires direct_mute_qc_before=" in r.stderr
    # adj=0 (modelling) takes its data axes from keys, with the reference's messages
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "adj=0")
    assert r.returncode == 1 and "ERROR: Need nt=" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "adj=0", "nt=100")
    assert r.returncode == 1 and "ERROR: Need dt=" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "adj=0", "nt=100", "dt=0.004", "dh=0.1",
             "ds=0.1", "direct_mute=1", "direct_mute_slope=1")
    assert r.returncode == 1 and "direct_mute=1 is an input-data preprocessing option" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=s.rsf", "adj=2")
    assert r.returncode == 1 and "adj=2" in r.stderr
    r = _cli(tmp_path, *base, "sderiv=missing.rsf")
    assert r.returncode == 1 and "Cannot open" in r.stderr


def test_rsf_roundtrip_and_xdr(tmp_path):
    a = np.random.default_rng(0).normal(size=(3, 4, 5)).astype(np.float32)
    rsf.write(str(tmp_path / "a.rsf"), a, o=[0.5, 1.0, 2.0], d=[0.1, 0.2, 0.3],
              labels=["Time", "Offset", "Shot"])
    b, h = rsf.read(str(tmp_path / "a.rsf"))
    assert np.array_equal(a, b) and [x.n for x in h.axes[:3]] == [5, 4, 3]
    assert h.extra["label2"] == "Offset"
    # a header without data_format is XDR (big-endian)
    a.astype(">f4").tofile(tmp_path / "x.rsf@")
    (tmp_path / "x.rsf").write_text('in="x.rsf@"\nn1=5 n2=4 n3=3\n')
    c, _ = rsf.read(str(tmp_path / "x.rsf"))
    assert np.array_equal(a, c)

"""The resident kernel with receivers that move with the shot: cmp=1, the
reference's default geometry (src/kirchmig2d_cig.cpp:576-579; receiver
position s + h, :1061).

The kernel keeps receiver columns instead of receiver indices: a column holds
one receiver tuple slot (one receiver position) for every shot, and a chunk of
columns sweeps only the shots that have a trace in it, each trace's receiver
index and trace-aperture weight coming from a per-(shot, column) table.  It is
held to the FP64 oracle and the reference's own CIGs at relL2 <= 1e-4 (as the
golden cases of every variant), to the stream kernel (the reference's
per-point order) at 1e-4, and to itself: culls on and off bitwise, shot groups
to float rounding."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu

CMP1 = ["angle_cmp1", "mute_invert", "hard_cone"]


def _args(run):
    return (run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"],
            run["rder"], run["r_axes"])


def _kmig():
    from kirch_datuming_amd import kmig
    return kmig


@pytest.mark.parametrize("tile", [9, 10, 11, 13, 15])
@pytest.mark.parametrize("name", CMP1)
def test_golden_cmp1_cases_on_the_resident_kernel(name, tile):
    kv, run, exp = gc.case(name)
    assert int(kv.get("cmp", 1)) == 1
    print(f"cmp=1 case {name}, resident preset {tile}", flush=True)
    kmig = _kmig()
    ax = run["data_axes"]
    p = kmig.parse_params(dict(kv, variant="resident", tile=tile), int(ax[1][0]), ax[1][1],
                          ax[1][2])
    with kmig.Migrator(p) as m:
        m.set_tables(0, run["stab"], run["sder"], run["s_axes"])
        m.set_tables(1, run["rtab"], run["rder"], run["r_axes"])
        m.set_data(run["data"], ax)
        m.run()
        m.sync()
        cig = m.cig()
        assert m.stats()["variant"] == 2
    ref64 = O.migrate(kv, *_args(run), precision="f64")[0]
    r64, rref = O.rel_l2(cig, ref64), O.rel_l2(cig, exp["cig"])
    print(f"{name} preset {tile}: rel vs f64 {r64:.2e}, vs reference f32 {rref:.2e}")
    assert r64 <= 1e-4 and rref <= 1e-4


def _vz_cmp1(aperture_trace=None):
    from tests.test_gpu_parity import _vz_tables
    rng = np.random.default_rng(23)
    nz, nx, d = 60, 120, 0.02
    pos = 0.02 * np.arange(nx)
    T, D = _vz_tables(nz, nx, d, pos)
    nt, dt, ns, nh = 500, 0.004, 30, 24
    t = np.arange(nt) * dt
    data = np.zeros((ns, nh, nt), np.float32)
    for k in range(40):  # band-limited events plus noise
        t0 = rng.uniform(0.2, 1.8)
        a = (np.pi * 15.0 * (t - t0)) ** 2
        data += (rng.normal() * (1 - 2 * a) * np.exp(-a)).astype(np.float32)[None, None, :]
    data += 0.05 * rng.normal(size=data.shape).astype(np.float32)
    dax = [[nt, 0.0, dt], [nh, -0.46, 0.04], [ns, 0.5, 0.04]]
    tax = [[nz, 0.0, d], [nx, 0.0, d], [len(pos), 0.0, 0.02]]
    kv = dict(cmp=1, mode="angle", angle_min=0, angle_max=60, angle_step=1, tau=0.01)
    if aperture_trace:
        kv.update(aperture_trace=aperture_trace, aperture_trace_taper=3)
    return kv, (data, dax, T, D, tax, T, D, tax)


def _run(kv, args, **extra):
    kmig = _kmig()
    data, dax = args[0], args[1]
    p = kmig.parse_params(dict(kv, **extra), int(dax[1][0]), dax[1][1], dax[1][2])
    with kmig.Migrator(p) as m:
        m.set_tables(0, args[2], args[3], args[4])
        m.set_tables(1, args[5], args[6], args[7])
        m.set_data(data, dax)
        m.run()
        m.sync()
        return m.cig(), m.stats(), (m.count_updates() if m.stats()["variant"] == 2
                                    and extra.get("precision", "fp32") == "fp32" else None)


@pytest.mark.parametrize("aperture_trace", [None, 8])
def test_cmp1_vz_against_fp64_stream_and_itself(monkeypatch, aperture_trace):
    kv, args = _vz_cmp1(aperture_trace)
    monkeypatch.delenv("KMIG_CULL", raising=False)
    monkeypatch.delenv("KMIG_SPLIT", raising=False)
    print(f"cmp=1 v(z) 60x120, 30 shots x 24 offsets, aperture_trace={aperture_trace}",
          flush=True)
    res, st, cnt = _run(kv, args, variant="resident")
    assert st["variant"] == 2
    ref64 = O.migrate(kv, *args, precision="f64")[0]
    stream = _run(kv, args, variant="stream")[0]
    r64, rst = O.rel_l2(res, ref64), O.rel_l2(res, stream)
    print(f"resident vs f64 {r64:.2e}, vs stream {rst:.2e}; evaluated "
          f"{cnt['evaluated'] / cnt['nominal']:.3f}, accepted {cnt['accepted'] / cnt['nominal']:.3f}")
    assert np.abs(ref64).max() > 0
    assert r64 <= 1e-4 and rst <= 1e-4
    # columns without a trace in a shot still run in their chunk, so the
    # evaluated updates can exceed the nominal ones here (24 offsets in
    # chunks of 8 or 16 columns); the accepted ones cannot
    assert 0 < cnt["accepted"] <= min(cnt["evaluated"], cnt["nominal"])
    assert st["receiver_slots"] <= 120  # one slot per lattice position (0.04 km)
    monkeypatch.setenv("KMIG_CULL", "0")
    nocull = _run(kv, args, variant="resident")[0]
    monkeypatch.delenv("KMIG_CULL")
    assert np.array_equal(nocull.view(np.uint32), res.view(np.uint32))
    for split in ("1", "3"):
        monkeypatch.setenv("KMIG_SPLIT", split)
        sp = _run(kv, args, variant="resident")[0]
        assert O.rel_l2(sp, res) <= 1e-6
    monkeypatch.delenv("KMIG_SPLIT")
    fp16 = _run(kv, args, variant="resident", precision="fp16")[0]
    r16 = O.rel_l2(fp16, ref64)
    print(f"fp16 tuples vs f64 {r16:.2e}")
    assert r16 <= 5e-4

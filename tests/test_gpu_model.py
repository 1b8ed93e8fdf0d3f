"""adj=0 modelling on the GPU (kmig_model through the C ABI).

The GPU sprays every accepted update into the traces with float atomics (in
no fixed order) and then doubly integrates each trace in float, so it is held
to the FP64 oracle by relative L2 <= 3e-4: the oracle's own float build sits
at about 5e-5 from FP64 on these cases, the running sums of the integration
carrying the rounding of every earlier sample.  The transpose property
<L m, d> = <m, L^T d> is checked with both operators on the GPU.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu

TOL = 3e-4
REPEAT = 2e-4
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kirch_datuming_amd", "bin", "kirchmig2d_cig")
# every golden geometry but type_p, where the reference reads outside its
# table slice at ix < 2; the first group also runs on the resident kernel
# (angle gathers, two-bin interpolation, receivers fixed in space), the
# second (offset gathers, stacks, cmp=1, nearest bins) only on the stream one
RESIDENT = ["angle_base", "mute_qc", "aperture_trace", "type_l", "half", "gradfiles",
            "salt_like", "default_axis"]
STREAM_ONLY = ["angle_cmp1", "mute_invert", "nearest", "hard_cone", "offset_cig", "stack"]


def _kmig():
    from kirch_datuming_amd import kmig
    return kmig


def _inputs(name, seed):
    kv, run, _ = gc.case(name)
    kv = {k: v for k, v in kv.items() if not k.startswith(("direct_mute", "diff"))}
    rng = np.random.default_rng(seed)
    ax = run["data_axes"]
    p = O.params_from_kv(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    ng = p.ng if p.cig else 1
    nz, nx = run["s_axes"][0][0], run["s_axes"][1][0]
    m = rng.normal(size=(ng, nx, nz)).astype(np.float32)
    d = rng.normal(size=run["data"].shape).astype(np.float32)
    tabs = (run["stab"], run["sder"], run["s_axes"], run["rtab"], run["rder"], run["r_axes"])
    return kv, ax, m, d, tabs, run["grads"]


@pytest.mark.parametrize("name,variant", [(n, v) for n in RESIDENT for v in ("auto", "stream")]
                         + [(n, "auto") for n in STREAM_ONLY])
def test_model_against_fp64_oracle(name, variant):
    kv, ax, m, _, tabs, grads = _inputs(name, 5)
    print(f"model case {name} variant {variant}", flush=True)
    got = _kmig().model(dict(kv, variant=variant), m, ax, *tabs, grads=grads)
    ref = O.model(kv, m, ax, *tabs, grads=grads, precision="f64")
    assert got.shape == ref.shape
    r = O.rel_l2(got, ref)
    print(f"{name}: rel vs f64 {r:.2e}")
    assert np.abs(ref).max() > 0 and r < TOL


@pytest.mark.parametrize("name,variant", [("angle_base", "resident"), ("salt_like", "resident"),
                                          ("gradfiles", "resident"), ("angle_base", "stream"),
                                          ("offset_cig", "stream"), ("stack", "stream"),
                                          ("nearest", "stream")])
def test_gpu_model_is_the_transpose_of_gpu_migration(name, variant):
    kv, ax, m, d, tabs, grads = _inputs(name, 11)
    kv = dict(kv, variant=variant)
    print(f"dot-product case {name} variant {variant}", flush=True)
    kmig = _kmig()
    Lm = kmig.model(kv, m, ax, *tabs, grads=grads).astype(np.float64)
    LTd = kmig.migrate(kv, d, ax, *tabs, grads=grads)[0].astype(np.float64)
    lhs = float(np.dot(Lm.ravel(), d.astype(np.float64).ravel()))
    rhs = float(np.dot(m.astype(np.float64).ravel(), LTd.ravel()))
    # each side is a float sum of ~1e5 terms of both signs: a few 1e-5 of
    # the larger term norm bounds their rounding
    scale = np.linalg.norm(Lm) * np.linalg.norm(d)
    assert abs(lhs - rhs) <= 1e-4 * scale, (lhs, rhs, scale)


def test_repeat_model_runs_and_context_reuse():
    kmig = _kmig()
    kv, ax, m, _, tabs, _ = _inputs("angle_base", 7)
    kv = dict(kv, adj=0)
    p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    with kmig.Migrator(p) as mg:
        mg.set_tables(0, tabs[0], tabs[1], tabs[2])
        mg.set_tables(1, tabs[3], tabs[4], tabs[5])
        a = mg.model(m, ax)
        b = mg.model(m, ax)
        # atomics add in no fixed order, and the double integration carries
        # each sample's rounding into every later one: runs agree to the
        # float build's distance from FP64
        assert O.rel_l2(a, b) < REPEAT
        # a second image through the same context: linear in the image
        c = mg.model(2.0 * m, ax)
        assert O.rel_l2(c, 2.0 * a) < REPEAT
        st = mg.stats()
        assert st["ms_migrate"] > 0
        # data axes with fewer shots reuse the context's buffers
        ax2 = [ax[0], ax[1], [3, ax[2][1], ax[2][2]]]
        e = mg.model(m, ax2)
        assert e.shape == (3, a.shape[1], a.shape[2])
        assert O.rel_l2(e, a[:3]) < REPEAT


def test_model_refusals():
    kmig = _kmig()
    kv, ax, m, _, tabs, _ = _inputs("angle_base", 1)
    for prec in ("fp16", "bf16", "u16"):
        with pytest.raises(kmig.KmigError, match="precision=fp32"):
            kmig.model(dict(kv, precision=prec), m, ax, *tabs)
    with pytest.raises(kmig.KmigError, match="adj=0"):
        kmig.Migrator(kmig.parse_params(dict(kv, adj=1), int(ax[1][0]), ax[1][1],
                                        ax[1][2])).model(m, ax)


def test_cli_adj0_writes_the_modelled_data(tmp_path):
    from kirch_datuming_amd import rsf
    kv, ax, m, _, tabs, _ = _inputs("angle_base", 3)
    sax = tabs[2]
    for nm, arr in (("t", tabs[0]), ("s", tabs[1])):
        rsf.write(str(tmp_path / f"{nm}.rsf"), arr, o=[a[1] for a in sax],
                  d=[a[2] for a in sax])
    p = O.params_from_kv(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    rsf.write(str(tmp_path / "img.rsf"), m, o=[sax[0][1], sax[1][1], p.g0],
              d=[sax[0][2], sax[1][2], p.dg])
    (nt, t0, dt), (nh, h0, dh), (ns, s0, ds) = ax
    args = [CLI, "adj=0", "seismic_data=dat.rsf", "migration=img.rsf", "stable=t.rsf",
            "sderiv=s.rsf", "rtable=t.rsf", "rderiv=s.rsf", f"nt={nt}", f"t0={t0}",
            f"dt={dt}", f"nh={nh}", f"h0={h0}", f"dh={dh}", f"ns={ns}", f"s0={s0}",
            f"ds={ds}"] + [f"{k}={v}" for k, v in kv.items()]
    print("cli adj=0 angle_base", flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dat, h = rsf.read(str(tmp_path / "dat.rsf"))
    assert dat.shape == (ns, nh, nt)
    assert [h.axes[k].n for k in range(3)] == [nt, nh, ns]
    ref = _kmig().model(kv, m, ax, *tabs)
    assert O.rel_l2(dat, ref) < REPEAT

"""Offset gathers and the stacked image (cig=0, the program's default output)
on the resident kernel: every update goes with all of its weight to its
trace's own bin (the offset bin of its receiver index, or the single stack
bin; src/kirchmig2d_cig.cpp:1100-1102), with no directions and no angle
weights; the time-window culls still apply.  Held to the FP64 oracle and the
reference's golden CIGs at relL2 <= 1e-4, to the stream kernel at 1e-4, and to
itself: culls on and off bitwise, shot groups to float rounding."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu


def _args(run):
    return (run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"],
            run["rder"], run["r_axes"])


def _run(kv, args, **extra):
    from kirch_datuming_amd import kmig
    data, dax = args[0], args[1]
    p = kmig.parse_params(dict(kv, **extra), int(dax[1][0]), dax[1][1], dax[1][2])
    with kmig.Migrator(p) as m:
        m.set_tables(0, args[2], args[3], args[4])
        m.set_tables(1, args[5], args[6], args[7])
        m.set_data(data, dax)
        m.run()
        m.sync()
        return m.cig(), m.stats()


@pytest.mark.parametrize("tile", [9, 11, 15])
@pytest.mark.parametrize("name", ["offset_cig", "stack"])
def test_golden_offset_and_stack_on_the_resident_kernel(name, tile):
    kv, run, exp = gc.case(name)
    print(f"{name}: resident preset {tile}", flush=True)
    cig, st = _run(kv, _args(run), variant="resident", tile=tile)
    assert st["variant"] == 2
    ref64 = O.migrate(kv, *_args(run), precision="f64")[0]
    r64, rref = O.rel_l2(cig, ref64), O.rel_l2(cig, exp["cig"])
    print(f"{name} preset {tile}: rel vs f64 {r64:.2e}, vs reference f32 {rref:.2e}")
    assert r64 <= 1e-4 and rref <= 1e-4


def _vz(cmp, mode):
    from tests.test_gpu_cmp1 import _vz_cmp1
    kv, args = _vz_cmp1(aperture_trace=8)
    kv = {k: v for k, v in kv.items() if not k.startswith(("mode", "angle_"))}
    kv["cmp"] = cmp
    if mode == "offset":
        kv.update(mode="offset", cig=1)
    else:
        kv["cig"] = 0
    if not cmp:  # receivers fixed: an absolute receiver axis inside the tables
        data, dax = args[0], [list(args[1][0]), [24, 0.3, 0.06], list(args[1][2])]
        args = (data, dax) + tuple(args[2:])
    return kv, args


@pytest.mark.parametrize("cmp", [0, 1])
@pytest.mark.parametrize("mode", ["offset", "stack"])
def test_vz_offset_and_stack_against_fp64_stream_and_itself(monkeypatch, cmp, mode):
    kv, args = _vz(cmp, mode)
    monkeypatch.delenv("KMIG_CULL", raising=False)
    monkeypatch.delenv("KMIG_SPLIT", raising=False)
    print(f"v(z) {mode}, cmp={cmp}", flush=True)
    res, st = _run(kv, args, variant="resident")
    assert st["variant"] == 2
    ref64 = O.migrate(kv, *args, precision="f64")[0]
    stream = _run(kv, args, variant="stream")[0]
    r64, rst = O.rel_l2(res, ref64), O.rel_l2(res, stream)
    print(f"{mode} cmp={cmp}: resident vs f64 {r64:.2e}, vs stream {rst:.2e}")
    assert np.abs(ref64).max() > 0 and r64 <= 1e-4 and rst <= 1e-4
    monkeypatch.setenv("KMIG_CULL", "0")
    assert np.array_equal(_run(kv, args, variant="resident")[0].view(np.uint32), res.view(np.uint32))
    monkeypatch.setenv("KMIG_CULL", "1")
    monkeypatch.setenv("KMIG_SPLIT", "3")
    assert O.rel_l2(_run(kv, args, variant="resident")[0], res) <= 1e-6


def test_salt_stack_full_run_against_the_stream_kernel():
    """The program's default output (stacked image, cmp=1) on the salt
    workload at full size: resident kernel against the stream kernel."""
    from kirch_datuming_amd import configs
    w = configs.salt(threads=16, cmp=1, stack=True)
    print("salt stack: resident", flush=True)
    res, st = _run(w.kv, w.args())
    assert st["variant"] == 2 and np.all(np.isfinite(res))
    print(f"salt stack: resident migrate {st['ms_migrate']:.1f} ms (preset {st['tile']}); "
          "stream kernel", flush=True)
    stream, st2 = _run(w.kv, w.args(), variant="stream")
    r = O.rel_l2(res, stream)
    print(f"salt stack: stream migrate {st2['ms_migrate']:.1f} ms; relL2 {r:.2e}")
    assert r <= 1e-4

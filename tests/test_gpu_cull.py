"""Group culling of the resident migrate kernel changes nothing.

The kernel skips a (shot, receiver chunk) group of a wave when per-cell bounds
of the packed directions show that every opening angle of the group lies
above the largest angle that receives weight, or when the per-cell smallest
integer samples put every update's last tap past the end of the trace; each
skipped update would have added an exact zero.  The CIG with culling (default) is therefore bitwise the
CIG without it (KMIG_CULL=0) at the same preset, which these tests check on
workloads where many groups are culled and many are partly rejected: the v(z)
workload at its full geometry (61 shots x 201 receivers, 201 x 301 points),
in FP32 over four tile shapes and in FP16, the batched five-layer graph, and
a mix of options (half angles, no depth-mute taper, a bin range that starts
above 0 degrees, the angle taper).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(w, kv, monkeypatch, cull):
    from kirch_datuming_amd import kmig
    monkeypatch.setenv("KMIG_CULL", "1" if cull else "0")
    return kmig.migrate(kv, *w.args())[0]


@pytest.fixture(scope="module")
def small():
    from kirch_datuming_amd import configs
    return configs.small(threads=8)


@pytest.mark.parametrize("tile,precision", [(9, "fp32"), (11, "fp32"), (13, "fp32"),
                                            (16, "fp32"), (11, "fp16")])
def test_culling_is_bitwise_neutral(small, monkeypatch, tile, precision):
    kv = dict(small.kv, variant="resident", tile=tile, precision=precision)
    print(f"cull: small workload, preset {tile}, {precision}", flush=True)
    off = _run(small, kv, monkeypatch, False)
    on = _run(small, kv, monkeypatch, True)
    assert np.abs(off).max() > 0
    assert np.array_equal(off.view(np.uint32), on.view(np.uint32))


@pytest.mark.parametrize("nt,tau", [(300, 0.0), (450, 0.1), (450, -0.04)])
def test_time_window_culling_is_bitwise_neutral(small, monkeypatch, nt, tau):
    """Short traces (and a time shift tau) put the deep points' taps past the
    trace end, where whole groups are culled by the per-cell smallest
    integer samples; the CIG is still bitwise the unculled one."""
    from kirch_datuming_amd import kmig
    args = list(small.args())
    args[0] = np.ascontiguousarray(args[0][..., :nt])
    args[1] = [[nt] + list(args[1][0][1:])] + [list(a) for a in args[1][1:]]
    kv = dict(small.kv, variant="resident", tile=11, tau=tau)
    print(f"cull: small workload cut to nt={nt}, tau={tau}", flush=True)
    out = {}
    for cull in (False, True):
        monkeypatch.setenv("KMIG_CULL", "1" if cull else "0")
        out[cull] = kmig.migrate(kv, *args)[0]
    assert np.abs(out[False]).max() > 0
    assert np.array_equal(out[False].view(np.uint32), out[True].view(np.uint32))


def test_culling_with_other_angle_options(small, monkeypatch):
    kv = dict(small.kv, variant="resident", tile=9, half_angle=1, oaztaper=0, oazmin=35,
              oazmax=20, angle_min=5, angle_max=45, angle_step=1, angle_n=41,
              angle_taper_start=30, angle_taper_end=40)
    print("cull: small workload, half angles, bins 5..45, angle taper", flush=True)
    off = _run(small, kv, monkeypatch, False)
    on = _run(small, kv, monkeypatch, True)
    assert np.abs(off).max() > 0
    assert np.array_equal(off.view(np.uint32), on.view(np.uint32))


def test_culling_in_the_batch_graph(monkeypatch):
    from kirch_datuming_amd import configs, kmig
    ws = configs.layers(threads=8, scale=0.05)
    out = {}
    for cull in (False, True):
        monkeypatch.setenv("KMIG_CULL", "1" if cull else "0")
        print(f"cull: five-layer batch graph, cull={cull}", flush=True)
        ms = []
        for w in ws:
            ax = w.data_axes
            p = kmig.parse_params(dict(w.kv, tile=9), int(ax[1][0]), ax[1][1], ax[1][2])
            m = kmig.Migrator(p)
            m.set_tables(0, w.stab, w.sder, w.s_axes)
            m.set_tables(1, w.rtab, w.rder, w.r_axes)
            m.set_data(w.data, ax)
            ms.append(m)
        with kmig.Batch(ms) as b:
            b.run()
            b.sync()
            out[cull] = [m.cig() for m in ms]
        for m in ms:
            m.close()
    for k in range(len(ws)):
        assert np.array_equal(out[False][k].view(np.uint32), out[True][k].view(np.uint32)), k

"""GPU depth-layer stitch (kmig_batch_stitch) against the stitch oracle, bit
for bit, on five overlapping salt layers migrated as one batch."""
import numpy as np
import pytest

from oracle import stitch as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    from kirch_datuming_amd import configs, kmig
    ws = configs.layers(threads=8, scale=0.04, overlap=12)
    ms = []
    for w in ws:
        ax = w.data_axes
        p = kmig.parse_params(dict(w.kv, tile=9), int(ax[1][0]), ax[1][1], ax[1][2])
        m = kmig.Migrator(p)
        m.set_tables(0, w.stab, w.sder, w.s_axes)
        m.set_tables(1, w.rtab, w.rder, w.r_axes)
        m.set_data(w.data, ax)
        ms.append(m)
    b = kmig.Batch(ms)
    b.run()
    b.sync()
    cigs = [m.cig() for m in ms]
    yield b, ws, cigs
    b.close()
    for m in ms:
        m.close()


@pytest.mark.parametrize("opts", [dict(), dict(stitch_blend=-1), dict(stitch_blend=0),
                                  dict(stitch_blend=5, scale=[1.0, 0.8, 1.2, 1.0, 0.9],
                                       scalex_aper=40, scalex_scale=1.7, angles=(3, 40))])
def test_stitch_matches_oracle_bitwise(batch, opts):
    b, ws, cigs = batch
    start = [w.z_range[0] for w in ws]
    end = [w.z_range[1] for w in ws]
    print(f"stitch {opts}", flush=True)
    gpu = b.stitch(start, end, **opts)
    ref, bad = S.stitch(cigs, start, end, **opts)
    assert bad == 0 and gpu.shape == ref.shape == (ref.shape[0], 1000, 400)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_stitch_errors_carry_the_reference_messages(batch):
    from kirch_datuming_amd import kmig
    b, ws, _ = batch
    start = [w.z_range[0] for w in ws]
    end = [w.z_range[1] for w in ws]
    bad = list(start)
    bad[2] = end[1] + 1
    with pytest.raises(kmig.KmigError, match="Layer 3 does not overlap previous layer"):
        b.stitch(bad, end)
    with pytest.raises(kmig.KmigError, match="First layer must start at depth sample 0"):
        b.stitch([1] + start[1:], end)
    with pytest.raises(kmig.KmigError, match="stitch_blend must be -1 or nonnegative"):
        b.stitch(start, end, stitch_blend=-2)
    with pytest.raises(kmig.KmigError, match="dimensions are"):
        b.stitch(start, [e + 1 for e in end])

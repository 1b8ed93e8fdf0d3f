"""CPU checks of the depth-layer stitch oracle (oracle/stitch.py) against what
the reference's loop fixes (src/layercom_cig.cpp:484-545): constants survive
the complementary blend, the blend runs from the upper layer's value (w = 0)
to the lower one's (w = 1) over the last stitch_blend overlap samples,
stitch_blend = 0 keeps the upper layer, -1 blends the whole overlap, and the
lateral ramp scales only the layers below the first."""
import math

import numpy as np

from oracle import stitch as S


def _layers(vals, start, end, ng=3, nx=5):
    return [np.full((ng, nx, e - s + 1), v, np.float32) for v, s, e in zip(vals, start, end)]


def test_constant_layers_stay_constant():
    start, end = [0, 8, 15], [11, 19, 29]
    out, bad = S.stitch(_layers([2.5, 2.5, 2.5], start, end), start, end, stitch_blend=-1)
    assert bad == 0 and out.shape == (3, 5, 30)
    np.testing.assert_allclose(out, 2.5, rtol=2e-7)


def test_blend_weights_and_keep_upper():
    start, end = [0, 10], [19, 29]           # overlap 10 samples (10..19)
    out, _ = S.stitch(_layers([1.0, 3.0], start, end), start, end, stitch_blend=4)
    col = out[0, 0]
    assert np.all(col[:16] == 1.0)           # 10..15 keep the upper layer (keep_upper = 6)
    w = [0.5 * (1 - math.cos(math.pi * b / 3)) for b in range(4)]
    for b in range(4):                        # 16..19 blend, w from 0 to 1
        assert abs(col[16 + b] - (1.0 * (1 - w[b]) + 3.0 * w[b])) < 1e-6
    assert col[16] == 1.0 and col[19] == 3.0
    assert np.all(col[20:] == 3.0)


def test_blend_zero_and_whole_overlap():
    start, end = [0, 10], [19, 29]
    out0, _ = S.stitch(_layers([1.0, 3.0], start, end), start, end, stitch_blend=0)
    assert np.all(out0[0, 0, :20] == 1.0) and np.all(out0[0, 0, 20:] == 3.0)
    outw, _ = S.stitch(_layers([1.0, 3.0], start, end), start, end, stitch_blend=-1)
    col = outw[0, 0]
    assert col[10] == 1.0 and col[19] == 3.0 and 1.0 < col[15] < 3.0
    one, _ = S.stitch(_layers([1.0, 3.0], start, end), start, end, stitch_blend=1)
    assert np.all(one[0, 0, :19] == 1.0) and one[0, 0, 19] == 3.0   # u = 1: w = 1


def test_lateral_ramp_and_scales():
    start, end = [0, 5], [9, 14]
    lat = S.lateral_scale(7, 3, 2.0)
    np.testing.assert_allclose(lat, [2.0, 1.5, 1.0, 1.0, 1.0, 1.5, 2.0])
    out, _ = S.stitch(_layers([1.0, 1.0], start, end, nx=7), start, end, scale=[1.0, 0.5],
                      stitch_blend=0, scalex_aper=3, scalex_scale=2.0)
    assert np.all(out[:, :, :10] == 1.0)                       # first layer: no ramp
    np.testing.assert_allclose(out[0, :, 12], 0.5 * lat)       # below: scale x ramp
    win, _ = S.stitch(_layers([1.0, 1.0], start, end), start, end, angles=(1, 2))
    assert win.shape == (2, 5, 15)


def test_non_finite_samples_are_counted_and_zeroed():
    start, end = [0, 3], [5, 8]
    lay = _layers([1.0, 1.0], start, end)
    lay[1][0, 0, 4] = np.nan
    out, bad = S.stitch(lay, start, end)
    assert bad == 1 and out[0, 0, 7] == 0.0

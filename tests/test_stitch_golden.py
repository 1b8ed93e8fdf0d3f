"""The stitch oracle and the gather extraction against the reference's own
outputs: layercom_cig and extract_adcig, compiled from the reference sources,
run on seeded layer gathers by tests/golden/make_stitch_golden.py (eight
option cases: default, partial, zero and whole-overlap blends, per-layer and
single scales, lateral ramps of aperture 3 and 1, one-sample overlaps,
START/END tokens).  Both are float arithmetic in the reference's order, so
they are held bit for bit."""
import json
import os

import numpy as np
import pytest

from oracle import stitch as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(HERE, "stitch_cases.json")))
DATA = np.load(os.path.join(HERE, "stitch.npz"))


def _scales(opt, n):
    s = opt.get("scale")
    if s is None:
        return None
    vals = [np.float32(float(v)) for v in s.split(",")]
    return vals * n if len(vals) == 1 else vals


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_stitch_is_the_reference_output(name):
    c = CASES[name]
    rng = c["ranges"]
    layers = [DATA[f"{name}_L{k}"] for k in range(1, len(rng) + 1)]
    opt = c["options"]
    out, nonfinite = S.stitch(layers, [r[0] for r in rng], [r[1] for r in rng],
                              scale=_scales(opt, len(rng)),
                              stitch_blend=opt.get("stitch_blend", 15),
                              scalex_aper=opt.get("scalex_aper", 0),
                              scalex_scale=opt.get("scalex_scale", 1.0))
    assert nonfinite == 0
    ref = DATA[f"{name}_out"]
    assert out.shape == ref.shape
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_extract_is_the_reference_output():
    from kirch_datuming_amd import kmig
    n = 0
    for name, c in CASES.items():
        full = DATA[f"{name}_out"]
        for j, (key, val) in enumerate(c["extract"]):
            kw = {key: val, "o2": c["o2"], "d2": c["d2"]}
            g = kmig.extract_adcig(full, **kw)
            assert np.array_equal(g, DATA[f"{name}_extract{j}"]), (name, key, val)
            n += 1
    assert n == 4


def test_extract_rejects_positions_outside_the_axis():
    from kirch_datuming_amd import kmig
    full = DATA["base_out"]
    with pytest.raises(kmig.KmigError, match=r"ix must be in \[0,8\]"):
        kmig.extract_adcig(full, ix=9)
    with pytest.raises(kmig.KmigError, match="is outside lateral axis"):
        kmig.extract_adcig(full, x=0.6, o2=0.1, d2=0.05)

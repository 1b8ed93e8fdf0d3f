"""Shot groups of the resident migrate launch.

A launch runs several blocks per tile, each over a contiguous share of the
batch's shots, and sums their partial CIGs in group order (the launch picks
the count; KMIG_SPLIT forces one).  The sums per image point then run group by
group instead of over all shots at once, so a split CIG equals the unsplit
one to float rounding, and repeated runs are bitwise equal (no atomics).
(The resident kernel runs only as one shot batch, so the CIG of earlier
launches enters a launch through its group 0 only on repeated runs.)"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from kirch_datuming_amd import configs
    return configs.small(threads=8)


def _run(w, kv, monkeypatch, split):
    from kirch_datuming_amd import kmig
    if split is None:
        monkeypatch.delenv("KMIG_SPLIT", raising=False)
    else:
        monkeypatch.setenv("KMIG_SPLIT", str(split))
    return kmig.migrate(kv, *w.args())[0]


@pytest.mark.parametrize("tile,precision", [(9, "fp32"), (15, "fp32"), (11, "fp16")])
def test_split_matches_unsplit(small, monkeypatch, tile, precision):
    kv = dict(small.kv, variant="resident", tile=tile, precision=precision)
    print(f"split: small workload, preset {tile}, {precision}", flush=True)
    one = _run(small, kv, monkeypatch, 1)
    auto = _run(small, kv, monkeypatch, None)
    eight = _run(small, kv, monkeypatch, 8)
    assert np.abs(one).max() > 0
    assert O.rel_l2(auto, one) < 1e-6 and O.rel_l2(eight, one) < 1e-6
    again = _run(small, kv, monkeypatch, None)
    assert np.array_equal(auto.view(np.uint32), again.view(np.uint32))


def test_split_on_a_shot_range(small, monkeypatch):
    """A run over shots [5, 45): the launch covers the range's 40 shots in
    groups and matches the unsplit launch of the same range."""
    from kirch_datuming_amd import kmig
    kv = dict(small.kv, variant="resident", tile=9)
    ax = small.data_axes
    p = kmig.parse_params(kv, int(ax[1][0]), ax[1][1], ax[1][2])
    out = {}
    for split in ("1", "5"):
        monkeypatch.setenv("KMIG_SPLIT", split)
        print(f"split: shots 5..45, KMIG_SPLIT={split}", flush=True)
        with kmig.Migrator(p) as m:
            m.set_tables(0, small.stab, small.sder, small.s_axes)
            m.set_tables(1, small.rtab, small.rder, small.r_axes)
            m.set_data(small.data, ax)
            m.run((5, 45))
            m.sync()
            out[split] = m.cig()
    assert np.abs(out["1"]).max() > 0
    assert O.rel_l2(out["5"], out["1"]) < 1e-6

"""The headline workload through the production path.

Salt model 400 x 1000 (dz = dx = 10 m), 768 shots x 401 receivers (cmp=0),
nt = 2001, 61 angle bins, direct mute (slope 50 s/km, taper 0.05 s): the
path the benchmark times, i.e. the autotuned resident preset, shot groups
whose partial CIGs k_sum_parts adds, and the direction and time-window culls.
Its result is held

* against the FP64 oracle on four evenly spaced shots at full image size,
  with two shot groups forced (KMIG_SPLIT=2) so the partial sums are in the
  checked path: relative L2 <= 1e-4 (rsf_l2cmp's metric,
  src/rsf_l2cmp.cpp:9-10), the same bound as the golden cases;
* on the full 768-shot run, by properties that hold at any size: finite and
  nonzero; the sum of the runs over two shot ranges equals the full run to
  1e-6 (float rounding of a regrouped sum); the run with the culls off is
  bitwise the same (a culled group adds exact zeros); and the stream kernel,
  which keeps the reference's per-point order (shot-major, receiver-minor,
  src/kirchmig2d_cig.cpp:1203-1339) and is pinned on the golden cases, agrees
  to 1e-4.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def salt():
    from kirch_datuming_amd import configs
    print("salt: building the workload (eikods2d tables, scatterer data)", flush=True)
    return configs.salt(threads=16)


def _migrator(w, **kv):
    from kirch_datuming_amd import kmig
    ax = w.data_axes
    p = kmig.parse_params(dict(w.kv, **kv), int(ax[1][0]), ax[1][1], ax[1][2])
    m = kmig.Migrator(p)
    m.set_tables(0, w.stab, w.sder, w.s_axes)
    m.set_tables(1, w.rtab, w.rder, w.r_axes)
    m.set_data(w.data, ax)
    return m


def test_salt_shot_sample_against_fp64(salt, monkeypatch):
    shots = [96, 288, 480, 672]
    sub = salt.shot_subset(shots)
    monkeypatch.setenv("KMIG_SPLIT", "2")
    print(f"salt: GPU on shots {shots}, two shot groups forced", flush=True)
    with _migrator(sub) as m:
        m.run()
        m.sync()
        cig = m.cig()
        st = m.stats()
    assert st["variant"] == 2 and st["migrate_launches"] == 1
    print(f"salt: FP64 oracle on shots {shots}", flush=True)
    ref = O.migrate(salt.kv, *sub.args(), precision="f64")[0]
    r = O.rel_l2(cig, ref)
    print(f"salt sample (preset {st['tile']}): relL2 vs FP64 {r:.2e}")
    assert np.abs(ref).max() > 0
    assert r <= 1e-4


def test_salt_full_run_properties(salt, monkeypatch):
    monkeypatch.delenv("KMIG_SPLIT", raising=False)
    monkeypatch.delenv("KMIG_CULL", raising=False)
    ns = salt.data.shape[0]
    with _migrator(salt) as m:
        print("salt: full production run", flush=True)
        m.run()
        m.sync()
        full = m.cig()
        st = m.stats()
        cnt = m.count_updates()
        assert st["variant"] == 2 and st["shot_batches"] == 1
        assert np.all(np.isfinite(full)) and np.abs(full).max() > 0
        print(f"salt: preset {st['tile']}, evaluated {cnt['evaluated'] / cnt['nominal']:.3f}, "
              f"accepted {cnt['accepted'] / cnt['nominal']:.3f} of the nominal updates")
        assert cnt["nominal"] == st["updates"]
        assert 0 < cnt["accepted"] <= cnt["evaluated"] <= cnt["nominal"]
        # two shot ranges add up to the full run
        parts = []
        for rng in ((0, ns // 2), (ns // 2, ns)):
            print(f"salt: shots {rng}", flush=True)
            m.run(rng)
            m.sync()
            parts.append(m.cig())
        r_sum = O.rel_l2(parts[0] + parts[1], full)
        # culls off: bitwise the same
        monkeypatch.setenv("KMIG_CULL", "0")
        print("salt: full run with the culls off", flush=True)
        m.run()
        m.sync()
        nocull = m.cig()
        monkeypatch.delenv("KMIG_CULL")
    print(f"salt: sum of shot ranges vs full {r_sum:.2e}")
    assert r_sum <= 1e-6
    assert np.array_equal(nocull.view(np.uint32), full.view(np.uint32))
    print("salt: full run on the stream kernel (reference order)", flush=True)
    with _migrator(salt, variant="stream") as m:
        m.run()
        m.sync()
        stream = m.cig()
        assert m.stats()["variant"] == 1
    r_str = O.rel_l2(full, stream)
    print(f"salt: resident vs stream kernel, full run: relL2 {r_str:.2e}")
    assert r_str <= 1e-4


@pytest.fixture(scope="module")
def salt_cmp1():
    from kirch_datuming_amd import configs
    print("salt cmp=1: building the workload (800 receiver tables)", flush=True)
    return configs.salt(threads=16, cmp=1)


def test_salt_cmp1_on_the_resident_kernel(salt_cmp1, monkeypatch):
    """The reference's default geometry (cmp=1) at full size: the resident
    kernel's receiver columns against the FP64 oracle on four shots (two shot
    groups forced), and the full run against the stream kernel and against
    itself with the culls off."""
    w = salt_cmp1
    shots = [96, 288, 480, 672]
    sub = w.shot_subset(shots)
    monkeypatch.setenv("KMIG_SPLIT", "2")
    print(f"salt cmp=1: GPU on shots {shots}, two shot groups forced", flush=True)
    with _migrator(sub) as m:
        m.run()
        m.sync()
        cig = m.cig()
        assert m.stats()["variant"] == 2
    monkeypatch.delenv("KMIG_SPLIT")
    print(f"salt cmp=1: FP64 oracle on shots {shots}", flush=True)
    ref = O.migrate(w.kv, *sub.args(), precision="f64")[0]
    r = O.rel_l2(cig, ref)
    print(f"salt cmp=1 sample: relL2 vs FP64 {r:.2e}")
    assert np.abs(ref).max() > 0 and r <= 1e-4
    with _migrator(w) as m:
        print("salt cmp=1: full production run", flush=True)
        m.run()
        m.sync()
        full = m.cig()
        st = m.stats()
        cnt = m.count_updates()
        assert st["variant"] == 2 and np.all(np.isfinite(full))
        print(f"salt cmp=1: preset {st['tile']}, {st['receiver_slots']} receiver slots, "
              f"migrate {st['ms_migrate']:.1f} ms, tuples {st['ms_tuple_prep']:.1f} ms; evaluated "
              f"{cnt['evaluated'] / cnt['nominal']:.3f}, accepted "
              f"{cnt['accepted'] / cnt['nominal']:.3f}")
        monkeypatch.setenv("KMIG_CULL", "0")
        m.run()
        m.sync()
        nocull = m.cig()
        monkeypatch.delenv("KMIG_CULL")
    assert np.array_equal(nocull.view(np.uint32), full.view(np.uint32))
    print("salt cmp=1: full run on the stream kernel", flush=True)
    with _migrator(w, variant="stream") as m:
        m.run()
        m.sync()
        stream = m.cig()
        ms = m.stats()["ms_total"]
    r_str = O.rel_l2(full, stream)
    print(f"salt cmp=1: resident vs stream kernel ({ms:.0f} ms), full run: relL2 {r_str:.2e}")
    assert r_str <= 1e-4

"""16-bit tuple storage against the FP64 oracle: FP16 fields (fractional
sample, anti-alias half-width and impulse weight as IEEE half) in the 16-byte
layout, BF16 and UNORM16 fields in a 12-byte tuple with per-slot scaling of
the anti-alias term; the gradient direction stays FP32 in every form.

The precision target of the 16-bit forms is relL2 <= 1e-3.  Each bound below
is the measured error with margin, so a regression shows:
  u16   measured 2.0e-5 .. 2.3e-5   bound 5e-5   (target met)
  fp16  measured 3.0e-4 .. 3.1e-4   bound 5e-4   (target met)
  bf16  measured 1.8e-3 .. 2.0e-3   bound 4e-3   (target NOT met: bfloat16 keeps
        8 significant bits, so the fractional sample is stored to ~2^-9 of a
        sample, 2x the error the 1e-3 target allows)
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu

BOUND = {"fp32": 1e-4, "u16": 5e-5, "fp16": 5e-4, "bf16": 4e-3}
TARGET = 1e-3  # the 16-bit forms' precision goal (bf16 misses it)


def _vz_case():
    from tests.test_gpu_parity import _vz_tables
    rng = np.random.default_rng(5)
    nz, nx, d = 50, 70, 0.02
    pos = 0.04 * np.arange(36)
    T, D = _vz_tables(nz, nx, d, pos)
    nt, dt, ns, nh = 600, 0.004, 6, 20
    t = np.arange(nt) * dt
    data = np.zeros((ns, nh, nt), np.float32)
    for k in range(40):  # band-limited events, so the angle gathers are smooth
        t0 = rng.uniform(0.2, 2.0)
        a = (np.pi * 15.0 * (t - t0)) ** 2
        data += (rng.normal() * (1 - 2 * a) * np.exp(-a)).astype(np.float32)[None, None, :]
    data += 0.05 * rng.normal(size=data.shape).astype(np.float32)
    dax = [[nt, 0.0, dt], [nh, 0.05, 0.07], [ns, 0.2, 0.22]]
    tax = [[nz, 0.0, d], [nx, 0.0, d], [len(pos), 0.0, 0.04]]
    kv = dict(cmp=0, mode="angle", angle_min=0, angle_max=60, angle_step=1)
    return kv, (data, dax, T, D, tax, T, D, tax)


@pytest.mark.parametrize("prec", ["fp32", "u16", "fp16", "bf16"])
def test_precision_modes_against_fp64(prec):
    from kirch_datuming_amd import kmig
    for label, (kv, args) in (("golden angle_base", (gc.case("angle_base")[0],
                                                       _golden_args("angle_base"))),
                              ("v(z) 50x70", _vz_case())):
        ref = O.migrate(kv, *args, precision="f64")[0]
        cig = kmig.migrate(dict(kv, precision=prec, variant="resident"), *args)[0]
        r = O.rel_l2(cig, ref)
        print(f"{prec} {label}: relL2 vs FP64 {r:.2e}")
        assert r < BOUND[prec]


def _golden_args(name):
    _, run, _ = gc.case(name)
    return (run["data"], run["data_axes"], run["stab"], run["sder"], run["s_axes"], run["rtab"],
            run["rder"], run["r_axes"])


def test_16bit_modes_need_the_resident_variant():
    """Nearest-bin gathers run on the stream kernel only, which keeps FP32
    tables: a 16-bit precision there is refused."""
    from kirch_datuming_amd import kmig
    kv, run, _ = gc.case("nearest")
    with pytest.raises(kmig.KmigError, match="applies to the resident kernel"):
        kmig.migrate(dict(kv, precision="u16"), *_golden_args("nearest"))


@pytest.mark.parametrize("prec", ["u16", "fp16", "bf16"])
def test_fused_16bit_pass_matches_the_separate_passes(monkeypatch, prec):
    """The fused tuple pass takes the 16-bit anti-alias scale over all points
    of a slot (the separate passes over the points with W > 0): the two CIGs
    agree to well within the mode's own error."""
    from kirch_datuming_amd import kmig
    kv, args = _vz_case()
    kv = dict(kv, precision=prec, variant="resident")
    monkeypatch.setenv("KMIG_TUPLE_FUSED", "0")
    sep = kmig.migrate(kv, *args)[0]
    monkeypatch.setenv("KMIG_TUPLE_FUSED", "1")
    fus = kmig.migrate(kv, *args)[0]
    ref = O.migrate(dict(kv, precision="fp32"), *args, precision="f64")[0]
    d, r = O.rel_l2(fus, sep), O.rel_l2(fus, ref)
    print(f"{prec}: fused vs separate {d:.2e}, fused vs FP64 {r:.2e}")
    assert d < BOUND[prec]
    assert r < BOUND[prec]

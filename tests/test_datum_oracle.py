"""The kirchdat2d_freq oracle (direct frequency-domain datuming) pinned.

The reference's frequency program needs FFTW, which is not built here.  Its
time-domain counterpart kirchdat2d_auto_cs (built from the reference's own
sources, tests/golden/make_datum_golden.py) applies the same operator sample
by sample; with the options only the frequency program has switched off
(antialias=0, data_taper=0, taper=0) the two agree up to the FFT's rounding
and the float rounding of the time program's own sums: relative L2 <= 1e-5
on five geometries (cmp=0 narrow and wide aperture, cmp=1 with ds = dh, with
jump = 2, and a split spread), for the output and for the receiver-side
(interm) cube.  The options the time program lacks are pinned by closed
forms and invariants below.
"""
import json
import os

import numpy as np
import pytest

from oracle import datum as D
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
INFO = json.load(open(os.path.join(HERE, "golden", "datum_cases.json")))
GOLD = np.load(os.path.join(HERE, "golden", "datum.npz"))
M = INFO["model"]
SDATUM = M["should_datum"] * M["d"]
POS = (0.0, M["d"])
TOL = 1e-5


def _case(name):
    """A golden geometry on the direct path (use_bf=0); the butterfly
    operators are pinned in test_bf_oracle.py."""
    c = INFO["cases"][name]
    return dict(c["params"], use_bf=0), GOLD[name + "/data"], c["data_axes"]


def test_greensol_is_bitwise_the_reference():
    out = D.greensol(GOLD["time"], [[M["nz"], 0.0, M["d"]], [M["nx"], 0.0, M["d"]]], POS, SDATUM)
    assert np.array_equal(out.view(np.uint32), GOLD["green"].view(np.uint32))


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(INFO["cases"]))
def test_against_the_time_domain_reference(name, precision):
    kv, data, ax = _case(name)
    kv = dict(kv, antialias=0, data_taper=0)
    g = GOLD["green"]
    out, mid = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision=precision, interm=True)
    assert O.rel_l2(mid, GOLD[name + "/interm"]) < TOL
    assert O.rel_l2(out, GOLD[name + "/out"]) < TOL
    assert np.abs(GOLD[name + "/out"]).max() > 0.1


def test_fft_against_numpy():
    rng = np.random.default_rng(2)
    for n in (1, 2, 8, 64, 512):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        assert np.allclose(D.fft(x), np.fft.fft(x), rtol=1e-12, atol=1e-12 * n)
        assert np.allclose(D.fft(x, inverse=True), np.fft.ifft(x) * n, rtol=1e-12,
                           atol=1e-12 * n)


@pytest.mark.parametrize("tau", [0.0137, 0.2, 0.31234, 1.0])
def test_H_is_the_spectrum_of_the_time_filter(tau):
    """H(tau, w) = DFT of the causal time filter
    h[n] = sum_k F[k]/dt ((1-delta) [n = m+k] + delta [n = m+k-1]),
    m = ceil(tau/dt), delta = m - tau/dt, with F the second difference of
    sqrt(((tau + k dt)/tau)^2 - 1) (first-principles, in double)."""
    dt, nsam, nfft = 0.004, 14, 512
    k = np.arange(nsam, dtype=np.float64)
    f = np.sqrt(((tau + k * dt) / tau) ** 2 - 1.0)
    f[0] = 0.0
    F = np.empty(nsam - 1)
    F[0] = f[1]
    F[1:] = f[2:] - 2 * f[1:-1] + f[:-2]
    q = tau / dt
    m = int(np.ceil(q))
    delta = m - q
    h = np.zeros(nfft)
    for kk in range(nsam - 1):
        h[m + kk] += (1 - delta) * F[kk] / dt
        h[m + kk - 1] += delta * F[kk] / dt
    spec = np.fft.rfft(h)
    for iw in (0, 1, 17, 100, 256):
        got = D.H(tau, dt, nsam, nfft, iw, precision="f64")
        assert abs(got - spec[iw]) <= 1e-9 * max(1.0, abs(spec[iw])), (iw, got, spec[iw])
    # the float build follows within float rounding
    got32 = D.H(tau, dt, nsam, nfft, 17)
    assert abs(got32 - spec[17]) <= 2e-4 * max(1.0, abs(spec[17]))
    assert D.H(0.0, dt, nsam, nfft, 5) == 0 and D.H(-1.0, dt, nsam, nfft, 5) == 0


def test_filter_taps_closed_form():
    tau, dt, nsam = 0.25, 0.002, 27
    f = np.sqrt(((tau + np.arange(nsam) * dt) / tau) ** 2 - 1.0)
    f[0] = 0.0
    ref = np.concatenate([[f[1]], f[2:] - 2 * f[1:-1] + f[:-2]])
    assert np.allclose(D.filt(tau, dt, nsam), ref, rtol=2e-4, atol=1e-6)


def test_taper_weights_closed_form():
    L = D.lib()
    assert L.kmdo_taper_weight(2, 1, 9, 3, 1) == 0.0       # outside the aperture
    assert L.kmdo_taper_weight(2, 10, 9, 3, 1) == 0.0
    assert L.kmdo_taper_weight(0, 4, 9, 0, 1) == 1.0       # taper=0: no taper
    assert L.kmdo_taper_weight(0, 0, 20, 4, 1) == 0.0      # the edges are zero
    assert L.kmdo_taper_weight(0, 10, 20, 4, 1) == 1.0     # the interior is one
    # half way into a ramp: 0.5 (1 - cos(pi/2)) = 0.5, times the other side's 1
    assert abs(L.kmdo_taper_weight(0, 2, 20, 4, 1) - 0.5) < 1e-7
    # stride widens the ramp: 2 samples of stride 2 cover half of tap 2
    assert abs(L.kmdo_taper_weight(0, 2, 20, 2, 2) - 0.5) < 1e-7
    for c in range(0, 21):                                  # symmetric
        assert L.kmdo_taper_weight(0, c, 20, 5, 1) == L.kmdo_taper_weight(0, 20 - c, 20, 5, 1)


def test_edge_taper_closed_form():
    L = D.lib()
    n, tap = 30, 6
    w = np.array([L.kmdo_edge_taper(i, n, tap) for i in range(n)])
    assert w[0] == 0.0 and w[-1] == 0.0
    assert np.all(w[tap:n - tap] == 1.0)
    assert abs(w[3] - 0.5) < 1e-7                           # 0.5 (1 - cos(pi/2))
    assert np.allclose(w, w[::-1])
    assert L.kmdo_edge_taper(0, n, 0) == 1.0
    # the taper shrinks to (n - 1) / 2 samples on short axes
    w5 = [L.kmdo_edge_taper(i, 5, 10) for i in range(5)]
    assert w5[0] == 0.0 and w5[2] == 1.0 and abs(w5[1] - 0.5) < 1e-7


def test_antialias_weight_closed_form():
    L = D.lib()
    nfft, dt = 1024, 0.004                 # df = 0.244 Hz, fnyq = 125 Hz
    dtau = 0.01                            # fmax = 50 Hz, fpass = 40 Hz
    f = lambda iw: iw / (nfft * dt)
    iw_pass = int(40.0 / f(1))
    iw_stop = int(np.ceil(50.0 / f(1)))
    assert L.kmdo_antialias_weight(1.0, nfft, dt, iw_pass, dtau) == 1.0
    assert L.kmdo_antialias_weight(1.0, nfft, dt, iw_stop, dtau) == 0.0
    mid = int(round(45.0 / f(1)))          # 45 Hz: 0.5 (1 + cos(pi/2)) = 0.5
    assert abs(L.kmdo_antialias_weight(1.0, nfft, dt, mid, dtau) -
               0.5 * (1 + np.cos(np.pi * (f(mid) - 40.0) / 10.0))) < 1e-6
    assert L.kmdo_antialias_weight(0.0, nfft, dt, 300, dtau) == 1.0   # off
    assert L.kmdo_antialias_weight(1.0, nfft, dt, 0, dtau) == 1.0     # DC
    assert L.kmdo_antialias_weight(1.0, nfft, dt, 300, 1e-4) == 1.0   # fmax > fnyq
    # antialias scales the slope: twice the factor halves fmax
    assert L.kmdo_antialias_weight(2.0, nfft, dt, int(30.0 / f(1)), dtau) == 0.0


def test_data_taper_is_an_input_scaling():
    """data_taper=N equals tapering the traces beforehand with the same
    half-cosine weights (both in float: bitwise)."""
    L = D.lib()
    kv, data, ax = _case("cmp0")
    kv = dict(kv, antialias=0)
    ns, nh, _ = data.shape
    ws = np.array([L.kmdo_edge_taper(i, ns, 3) for i in range(ns)], np.float32)
    wh = np.array([L.kmdo_edge_taper(i, nh, 3) for i in range(nh)], np.float32)
    pre = (ws[:, None] * wh[None, :])[:, :, None] * data
    g = GOLD["green"]
    a = D.datum(dict(kv, data_taper=3), data, ax, g, POS, g, POS, SDATUM)
    b = D.datum(dict(kv, data_taper=0), pre.astype(np.float32), ax, g, POS, g, POS, SDATUM)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, D.datum(dict(kv, data_taper=0), data, ax, g, POS, g, POS,
                                          SDATUM))


def test_linear_and_precisions_agree_with_options_on():
    kv, data, ax = _case("cmp1_split")
    kv = dict(kv, antialias=1.5, data_taper=2, taper=3)
    g = GOLD["green"]
    a = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64")
    b = D.datum(kv, 2.0 * data, ax, g, POS, g, POS, SDATUM, precision="f64")
    assert O.rel_l2(b, 2.0 * a) < 1e-7      # float32 outputs of the FP64 build
    a32 = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f32")
    assert O.rel_l2(a32, a) < 1e-5
    # antialias and the cosine taper both act (the output changes)
    plain = D.datum(dict(kv, antialias=0, taper=0), data, ax, g, POS, g, POS, SDATUM,
                    precision="f64")
    assert O.rel_l2(plain, a) > 1e-3


def test_frequency_range_splits_the_sum():
    kv, data, ax = _case("cmp0")
    g = GOLD["green"]
    full = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64")
    nw = D.nfft(ax[0][0], int(kv.get("length", 0.025) / ax[0][2]) + 2, ax[0][2], g, g) // 2 + 1
    lo = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64", iw_range=(0, nw // 3))
    hi = D.datum(kv, data, ax, g, POS, g, POS, SDATUM, precision="f64", iw_range=(nw // 3, nw))
    assert O.rel_l2(lo + hi, full) < 1e-6     # float32 outputs of the FP64 build


def test_tie_geometry_follows_the_frequency_program():
    """cmp=1 with ds = 2 dh: half the receiver gathers start on an exact .5
    tie of (r - h)/ds + 0.5.  The frequency program rounds it in float
    (+ 0.5f), which the float build follows; the result stays finite and in
    the cube (the time program, rounding in double, differs there)."""
    g = GOLD["green"]
    rng = np.random.default_rng(4)
    data = rng.normal(size=(8, 14, 64)).astype(np.float32)
    ax = [[64, 0.0, 0.004], [14, 0.0, 0.05], [8, 0.05, 0.1]]
    out = D.datum({"cmp": 1, "aperture": 6, "use_bf": 0}, data, ax, g, POS, g, POS, SDATUM)
    assert np.isfinite(out).all() and np.abs(out).max() > 0


def test_table_range_errors_carry_the_reference_messages():
    g = GOLD["green"]
    data = np.zeros((4, 6, 64), np.float32)
    with pytest.raises(ValueError, match="Receiver table too small."):
        D.datum({"cmp": 0, "use_bf": 0}, data, [[64, 0.0, 0.004], [6, 1.5, 0.05], [4, 0.1, 0.05]],
                g, POS, g, POS, SDATUM)
    with pytest.raises(ValueError, match="Source table too small."):
        D.datum({"cmp": 0, "use_bf": 0}, data, [[64, 0.0, 0.004], [6, 0.1, 0.05], [4, 1.5, 0.05]],
                g, POS, g, POS, SDATUM)

"""The GPU butterfly's host planner (kirch_datuming_amd/csrc/kmdat_bfplan.cpp)
against the oracle's per-frequency panel decisions (no GPU).

tests/bf_plan_probe.cpp, built with g++ from the planner's own source, plans
an operator; the planner lists every panel that can become a butterfly with
its last admissible frequency and its candidate ancestors, and the GPU's mask
kernel adds, per frequency, that every entry is active and that no ancestor
is a butterfly.  That rule applied to the planner's list must give, at every
frequency, exactly the butterfly panels of the oracle's restatement of the
reference's segmented factor (oracle/kbf_oracle.c, pinned to bf1d.cpp), for
  * a banded (cmp=0) operator on the 64-position golden table, panel levels
    1-3 with leaf 8, and the salt datuming operator (128 positions) at the
    workflow's settings (leaf 16, p 12, one level);
  * pair-list (cmp=1-like) operators: three matrices whose rows skip their
    first column (a zero-weight pair) and carry antialias cut-offs.
The plan's groups must also keep every output trace of a chunk in one group
(the GPU updates traces without atomics)."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import bf
from oracle import datum as D

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DINFO = json.load(open(os.path.join(HERE, "golden", "datum_cases.json")))
DGOLD = np.load(os.path.join(HERE, "golden", "datum.npz"))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("bfplan")
    exe = str(d / "bf_plan_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "kirch_datuming_amd", "csrc"),
                    os.path.join(HERE, "bf_plan_probe.cpp"),
                    os.path.join(ROOT, "kirch_datuming_amd", "csrc", "kmdat_bfplan.cpp"),
                    "-o", exe], check=True)

    def run(hdr, fl, blobs):
        path = str(d / "in.bin")
        with open(path, "wb") as f:
            f.write(struct.pack("14i", *hdr))
            f.write(struct.pack("3f", *fl))
            for b in blobs:
                f.write(np.ascontiguousarray(b).tobytes())
        out = subprocess.run([exe, path], capture_output=True, text=True, check=True).stdout
        nodes, outs, groups = [], {}, []
        for line in out.split("\n"):
            t = line.split()
            if not t:
                continue
            if t[0] == "error":
                raise AssertionError(line)
            if t[0] == "node":
                v = list(map(int, t[1:]))
                nodes.append(dict(inst=v[0], row0=v[1], col0=v[2], L=v[3], ires=v[4],
                                  anc=v[6:6 + v[5]]))
            elif t[0] == "out":
                outs[int(t[1])] = list(map(int, t[2:]))
            elif t[0] == "group":
                groups.append((int(t[1]), list(map(int, t[2:]))))
        return nodes, outs, groups
    return run


def _aa(antialias, nfft, dt, iw, aarg):
    L = D.lib()
    return np.array([L.kmdo_antialias_weight(antialias, nfft, dt, iw, float(x))
                     for x in aarg.ravel()], np.float32).reshape(aarg.shape)


def _planner_bf(nodes, amps, iw, leaf):
    """The GPU's rule: a node is the butterfly at iw when iw <= its last
    admissible bin, every entry is active and no candidate ancestor passes
    the same two tests."""
    def raw(nd):
        nl = leaf << nd["L"]
        a = amps[nd["inst"]][nd["row0"]:nd["row0"] + nl, nd["col0"]:nd["col0"] + nl]
        return iw <= nd["ires"] and bool((a != 0).all())
    return {(nd["inst"], nd["row0"], nd["col0"]) for nd in nodes
            if raw(nd) and not any(raw(nodes[k]) for k in nd["anc"])}


def _oracle_bf(taus, amps, omega, p, leaf, levels):
    got = set()
    for k, (t, a) in enumerate(zip(taus, amps)):
        _, pan = bf.seg_apply(t, a.astype(np.complex64), omega, np.ones((1, t.shape[0])),
                              p=p, leaf=leaf, levels=levels)
        got |= {(k, int(r), int(c)) for r, c, _, _, kind in pan if kind == bf.BUTTERFLY}
    return got


def _banded(G, dpos, datum, aper):
    """cmp=0 receiver-side operator on table G: row ih, column ic, one key
    per entry (tau = G[ic, ih], the reference's pair weight and slope)."""
    n = G.shape[0]
    nb = 2 * aper + 1
    pkey = -np.ones((n, nb), np.int32)
    pw = np.zeros((n, nb), np.float32)
    tau = np.zeros((n, n), np.float32)
    aarg = np.zeros((n, n), np.float32)
    w = np.zeros((n, n), np.float32)
    d = np.float32(dpos)
    for ih in range(n):
        for ic in range(max(0, ih - aper), min(n - 1, ih + aper) + 1):
            t = G[ic, ih]
            dist = np.float32(np.float32(datum) * np.float32(datum) +
                              np.float32((ic - ih) ** 2) * d * d)
            wv = np.float32(1.0 / np.pi * float(d) * float(datum) * float(t) / float(dist))
            if ic == 0:
                sl = (G[1, ih] - G[0, ih]) / d
            elif ic == n - 1:
                sl = (G[n - 1, ih] - G[n - 2, ih]) / d
            else:
                sl = np.float32(0.5) * (G[ic + 1, ih] - G[ic - 1, ih]) / d
            pkey[ih, ic - ih + aper] = ih * n + ic
            pw[ih, ic - ih + aper] = wv
            tau[ih, ic], aarg[ih, ic], w[ih, ic] = t, np.float32(sl * abs(d)), wv
    return pkey, pw, tau, aarg, w


def _check_banded(probe, G, dpos, datum, aper, nt, dt, p, leaf, levels, step):
    n = G.shape[0]
    pkey, pw, tau, aarg, w = _banded(G, dpos, datum, aper)
    nsam = int(np.float32(0.05) / np.float32(dt)) + 2
    nfft = D.nfft(nt, nsam, dt, G, G)
    nw = nfft // 2 + 1
    nodes, _, groups = probe([0, n, 2 * aper + 1, aper, n * n, nw, nfft, p, leaf, levels, 1, 1,
                              0, 0], [dt, 1.0, 2.0],
                             [pkey, pw, tau.ravel(), aarg.ravel()])
    assert nodes, "no butterfly candidates"
    seen = 0
    for iw in range(0, nw, step):
        amp = np.where(tau > 0, w * _aa(1.0, nfft, dt, iw, aarg), 0).astype(np.float32)
        amp[np.abs(amp) < 1e-20] = 0
        om = np.float32(2.0) * np.float32(np.pi) * np.float32(iw) / (np.float32(nfft) * np.float32(dt))
        want = _oracle_bf([tau], [amp], om, p, leaf, levels)
        got = _planner_bf(nodes, [amp], iw, leaf)
        assert got == want, (iw, sorted(want ^ got))
        seen += len(want)
    assert seen > 0
    _check_groups(nodes, {}, groups, leaf)


def _check_groups(nodes, outs, groups, leaf):
    """Per chunk, no output trace (banded: row) in two groups."""
    owner = {}
    for g, (chunk, items) in enumerate(groups):
        for k in items:
            nd = nodes[k]
            rows = outs.get(k, range(nd["row0"], nd["row0"] + (leaf << nd["L"])))
            for r in rows:
                key = (chunk, nd["inst"] if k not in outs else -1, r)
                assert owner.setdefault(key, g) == g, key


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_banded_golden_table(probe, levels):
    g = DGOLD["green"]
    M = DINFO["model"]
    _check_banded(probe, g, M["d"], M["should_datum"] * M["d"], 63, 256, 0.004, 5, 8, levels,
                  step=3)


def test_banded_salt_workflow_settings(probe):
    from kirch_datuming_amd import configs
    w = configs.datum(threads=8, npos=128, nt=1001)
    _check_banded(probe, w.green, w.pos_axis[1], w.datum, 127, 1001, 0.002, 12, 16, 1, step=5)


def test_pair_lists(probe):
    """Three matrices of 40 rows; row r holds columns max(1, r - 30) ..
    min(39, r + 30) in order (its column 0 is a dropped zero-weight pair),
    traveltimes that differ per matrix, antialias cut-offs growing with |c - r|."""
    ninst, n, A, dx, dt, nfft = 3, 40, 30, 0.01, 0.004, 256
    nw = nfft // 2 + 1
    pairs, ro, rq0, rq1, tau_k, aarg_k = [], [], [], [], [], []
    taus = [np.zeros((n, n), np.float32) for _ in range(ninst)]
    argm = [np.zeros((n, n), np.float32) for _ in range(ninst)]
    for k in range(ninst):
        z = 0.3 + 0.05 * k
        for r in range(n):
            ro.append(k * n + r)
            rq0.append(len(pairs))
            for c in range(max(1, r - A), min(n - 1, r + A) + 1):
                t = np.float32(np.sqrt(z * z + ((c - r) * dx) ** 2) / 2.0 + 0.001 * c)
                ag = np.float32(abs(c - r) * 2.5e-4)
                pairs.append((k * n + c, len(tau_k), 1.0, 0))
                tau_k.append(t)
                aarg_k.append(ag)
                taus[k][r, c], argm[k][r, c] = t, ag
            rq1.append(len(pairs))
    pr = np.array(pairs, dtype=[("tr", "<i4"), ("key", "<i4"), ("w", "<f4"), ("pad", "<i4")])
    for leaf, p, levels in ((8, 5, 1), (8, 6, 2)):
        nodes, outs, groups = probe(
            [1, n, 0, 0, len(tau_k), nw, nfft, p, leaf, levels, ninst, n, 0, len(pairs)],
            [dt, 1.0, 2.0],
            [pr, np.array(ro, np.int32), np.array(rq0, np.int32), np.array(rq1, np.int32),
             np.array(tau_k, np.float32), np.array(aarg_k, np.float32)])
        assert nodes and all(o == [nodes[k]["inst"] * n + nodes[k]["row0"] + r
                                   for r in range(leaf << nodes[k]["L"])]
                             for k, o in outs.items())
        seen = 0
        for iw in range(0, nw, 2):
            amps = []
            for k in range(ninst):
                a = np.where(taus[k] > 0, _aa(1.0, nfft, dt, iw, argm[k]), 0).astype(np.float32)
                a[:, 0] = 0
                amps.append(a)
            om = np.float32(2.0) * np.float32(np.pi) * np.float32(iw) / (np.float32(nfft) * np.float32(dt))
            want = _oracle_bf(taus, amps, om, p, leaf, levels)
            got = _planner_bf(nodes, amps, iw, leaf)
            assert got == want, (leaf, p, levels, iw, sorted(want ^ got))
            seen += len(want)
        assert seen > 0
        _check_groups(nodes, outs, groups, leaf)

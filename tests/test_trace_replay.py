"""CPU replay of the coalesced trace preparation (k_trace_fwd, k_trace_bwd in
kirch_datuming_amd/csrc/kmig_prep.hip; no GPU).

The replay walks each block as the kernels do -- NTR traces per block, chunks
of CS samples loaded as 16-byte quads, rows RPI i + h (RPI = 256 / CS) at
offsets (RPI i nt + c0) + (h nt + 4 q) with no clamp, the swizzled tile
(quad (s / 4) ^ j of row j), the per-lane scans, the chunk-boundary sample of
the backward pass -- for both shapes the library has (16 x 256, 32 x 128),
and checks (1) every load
against the buffer length the library allocates (trace_alloc: 64 traces and
64 samples past the last trace), (2) that the swizzle is a permutation whose
row-wise and per-lane accesses are free of bank conflicts, and (3) the
prepared samples, the pairs and the QC copies against a plain per-trace
statement of the same float32 arithmetic (src/kirchmig2d_cig.cpp:110-129,
SERECKIRCH/src/kirmig.cpp:8-24).  Geometries: trace counts around the block
size, nt below, at and around the chunk size, a shot range that starts
inside the volume, skipped traces (wap < 0) and the QC shot."""
import numpy as np
import pytest

TR, PAD = 64, 64
f32 = np.float32


def trace_alloc(ntraces, nt):
    return (ntraces + PAD) * nt + 64


def chunk_at(j, s, CS):
    QPR = CS // 4
    return j * CS + ((((s >> 2) ^ j) & (QPR - 1)) << 2) + (s & 3)


def mute_w(t0, dt, boundary, taper, it):
    time = f32(f32(t0) + f32(f32(it) * f32(dt)))
    if time >= f32(boundary + taper):
        return f32(1.0)
    if time > boundary:
        u = f32((time - boundary) / f32(taper))
        return f32(f32(0.5) * f32(f32(1.0) - f32(np.cos(f32(np.pi) * u))))
    return f32(0.0)


class Geo:
    def __init__(self, ns, nh, nt, sb, se, mute, qc_shot, seed):
        rng = np.random.default_rng(seed)
        self.ns, self.nh, self.nt, self.sb, self.se = ns, nh, nt, sb, se
        ntot = ns * nh
        self.alloc = trace_alloc(ntot, nt)
        self.raw = np.full(self.alloc, np.nan, f32)
        self.raw[:ntot * nt] = rng.standard_normal(ntot * nt).astype(f32)
        self.wap = np.where(rng.random(ntot) < 0.2, -1.0, 1.0).astype(f32)
        self.off = rng.random(ntot).astype(f32)
        self.mute, self.qc_shot = mute, qc_shot
        self.t0, self.dt, self.slope, self.intercept, self.taper = 0.0, 0.004, 0.5, 0.02, 0.03

    def boundary(self, g):
        return f32(f32(self.intercept) + f32(f32(self.slope) * self.off[g]))


def replay(G, NTR, CS):
    QPR = CS // 4
    RPI = 64 // QPR
    NL = NTR // RPI
    nt, ntot = G.nt, G.ns * G.nh
    prep = np.full(G.alloc, np.nan, f32)
    pairs = np.full((G.alloc, 2), np.nan, f32)
    qcb = np.full(G.nh * nt, np.nan, f32)
    qca = np.full(G.nh * nt, np.nan, f32)
    ntr = (G.se - G.sb) * G.nh
    nblk = (ntr + NTR - 1) // NTR
    loads = []

    def at(j, s):
        return chunk_at(j, s, CS)

    def lane(t):
        return t // QPR, t % QPR

    def load_chunk(buf, g0, c0):
        pre = np.empty((NL, 64, 4), f32)  # [i][lane][sample of the quad]
        for t in range(64):
            h, q = lane(t)
            for i in range(NL):
                for e in range(4):
                    x = g0 * nt + c0 + (RPI * i * nt) + (h * nt + 4 * q) + e
                    loads.append(x)
                    assert 0 <= x < G.alloc, (g0, c0, t, i, x, G.alloc)
                    pre[i, t, e] = buf[x]
        return pre

    def stage(tile, pre):
        for t in range(64):
            h, q = lane(t)
            for i in range(NL):
                for e in range(4):
                    tile[at(RPI * i + h, 4 * q + e)] = pre[i, t, e]

    for b in range(nblk):
        g0 = G.sb * G.nh + b * NTR
        nloc = min(NTR, ntr - b * NTR)
        # forward
        tile = np.full(NTR * CS, np.nan, f32)
        qoff = G.qc_shot * G.nh - g0
        any_qc = G.mute and G.qc_shot >= 0 and g0 < G.qc_shot * G.nh + G.nh and \
            g0 + nloc > G.qc_shot * G.nh
        qlo, qhi = max(0, qoff), min(nloc, G.qc_shot * G.nh + G.nh - g0)
        tt = np.zeros(NTR, f32)
        pre = load_chunk(G.raw, g0, 0)
        for c0 in range(0, nt, CS):
            nc = min(CS, nt - c0)
            stage(tile, pre)
            if c0 + CS < nt:
                pre = load_chunk(G.raw, g0, c0 + CS)

            def qc_copy(dst):
                for t in range(64):
                    h, q = lane(t)
                    for j in range(qlo + h, qhi, RPI):
                        for e in range(4):
                            if 4 * q + e < nc:
                                dst[(j - qoff) * nt + c0 + 4 * q + e] = tile[at(j, 4 * q + e)]

            def steps(t, do_mute, do_sum):
                g = g0 + t
                for k in range(nc):
                    v = tile[at(t, k)]
                    if do_mute:
                        v = f32(v * mute_w(G.t0, G.dt, G.boundary(g), G.taper, c0 + k))
                    if do_sum:
                        tt[t] = v if c0 + k == 0 else f32(tt[t] + v)
                        v = tt[t]
                    tile[at(t, k)] = v

            if any_qc:
                qc_copy(qcb)
                for t in range(nloc):
                    if G.mute:
                        steps(t, True, False)
                qc_copy(qca)
                for t in range(nloc):
                    if G.wap[g0 + t] >= 0:
                        steps(t, False, True)
            else:
                for t in range(nloc):
                    keep = G.wap[g0 + t] >= 0
                    if G.mute or keep:
                        steps(t, bool(G.mute), keep)
            for t in range(64):
                h, q = lane(t)
                for i in range(NL):
                    j = RPI * i + h
                    if j >= nloc or 4 * q >= nc:
                        continue
                    for e in range(4):
                        if 4 * q + e < nc:
                            prep[(g0 + j) * nt + c0 + 4 * q + e] = tile[at(j, 4 * q + e)]
    for b in range(nblk):
        g0 = G.sb * G.nh + b * NTR
        nloc = min(NTR, ntr - b * NTR)
        tile = np.full(NTR * CS, np.nan, f32)
        bnd = np.full(NTR, np.nan, f32)
        tt = np.zeros(NTR, f32)
        last = ((nt - 1) // CS) * CS
        pre = load_chunk(prep, g0, last)
        out = prep.copy()
        for c0 in range(last, -1, -CS):
            nc = min(CS, nt - c0)
            stage(tile, pre)
            if c0 > 0:
                pre = load_chunk(prep, g0, c0 - CS)
            for t in range(nloc):
                if G.wap[g0 + t] < 0:
                    continue
                for k in range(nc - 1, -1, -1):
                    if c0 + k == nt - 1:
                        tt[t] = tile[at(t, k)]
                    else:
                        tt[t] = f32(tt[t] + tile[at(t, k)])
                        tile[at(t, k)] = tt[t]
            for t in range(64):
                h, q = lane(t)
                for i in range(NL):
                    j = RPI * i + h
                    if j >= nloc or 4 * q >= nc:
                        continue
                    v = [tile[at(j, 4 * q + e)] if 4 * q + e < CS else np.nan for e in range(4)]
                    sn = 4 * q + 4
                    after = tile[at(j, sn)] if sn < nc else (bnd[j] if c0 + sn < nt else f32(0))
                    for e in range(4):
                        if 4 * q + e >= nc:
                            break
                        o = (g0 + j) * nt + c0 + 4 * q + e
                        out[o] = v[e]
                        nx = v[e + 1] if e < 3 else after
                        pairs[o] = (v[e], f32(nx - v[e])) if c0 + 4 * q + e + 1 < nt else (0, 0)
            for t in range(nloc):
                bnd[t] = tile[at(t, 0)]
        prep[g0 * nt:(g0 + nloc) * nt] = out[g0 * nt:(g0 + nloc) * nt]
    return prep, pairs, qcb, qca, loads


def direct(G):
    """Per trace, in sample order: the reference's mute, QC copies, doubint."""
    nt = G.nt
    prep = np.full(G.alloc, np.nan, f32)
    pairs = np.full((G.alloc, 2), np.nan, f32)
    qcb = np.full(G.nh * nt, np.nan, f32)
    qca = np.full(G.nh * nt, np.nan, f32)
    for g in range(G.sb * G.nh, G.se * G.nh):
        x = G.raw[g * nt:(g + 1) * nt].copy()
        qc = G.mute and g // G.nh == G.qc_shot
        if qc:
            qcb[(g % G.nh) * nt:(g % G.nh + 1) * nt] = x
        if G.mute:
            for it in range(nt):
                x[it] = f32(x[it] * mute_w(G.t0, G.dt, G.boundary(g), G.taper, it))
        if qc:
            qca[(g % G.nh) * nt:(g % G.nh + 1) * nt] = x
        if G.wap[g] >= 0:
            tt = f32(0)
            for it in range(nt):
                tt = x[it] if it == 0 else f32(tt + x[it])
                x[it] = tt
            tt = x[nt - 1]
            for it in range(nt - 2, -1, -1):
                tt = f32(tt + x[it])
                x[it] = tt
        prep[g * nt:(g + 1) * nt] = x
        for it in range(nt):
            pairs[g * nt + it] = (x[it], f32(x[it + 1] - x[it])) if it + 1 < nt else (0, 0)
    return prep, pairs, qcb, qca


SHAPES = [(16, 256), (32, 128)]


@pytest.mark.parametrize("ntr_cs", SHAPES)
@pytest.mark.parametrize("ns,nh,nt,sb,se,mute,qc", [
    (1, 1, 16, 0, 1, 0, -1),       # one trace of the shortest length the chunk kernels take
    (2, 33, 17, 0, 2, 1, 1),       # 66 traces (the last block partial), nt < chunk
    (3, 12, 32, 1, 3, 1, 2),       # shot range from the second shot
    (2, 20, 257, 0, 2, 0, -1),     # nt one past a 256-sample chunk
    (1, 40, 300, 0, 1, 1, 0),      # QC shot, partial last chunk of both shapes
])
def test_trace_prep_replay(ns, nh, nt, sb, se, mute, qc, ntr_cs):
    NTR, CS = ntr_cs
    G = Geo(ns, nh, nt, sb, se, mute, qc, seed=ns * 1000 + nh * 10 + nt)
    prep, pairs, qcb, qca, loads = replay(G, NTR, CS)
    rp, rpairs, rqcb, rqca = direct(G)
    lo, hi = G.sb * G.nh * nt, G.se * G.nh * nt
    assert np.array_equal(prep[lo:hi].view(np.uint32), rp[lo:hi].view(np.uint32))
    assert np.array_equal(pairs[lo:hi].view(np.uint32), rpairs[lo:hi].view(np.uint32))
    if mute and qc >= 0 and sb <= qc < se:
        assert np.array_equal(qcb.view(np.uint32), rqcb.view(np.uint32))
        assert np.array_equal(qca.view(np.uint32), rqca.view(np.uint32))
    # the loads reach past the valid traces (into the allocation's pad) only
    assert max(loads) < G.alloc


@pytest.mark.parametrize("ntr_cs", SHAPES)
def test_chunk_bound_for_every_trace_length(ntr_cs):
    """A block's loads reach NTR nt + CS - 2 samples past its first trace
    (rows up to NTR - 1, the last chunk starting at most nt - 1, a quad up to
    CS - 1); from the last valid trace that stays inside trace_alloc() for
    every nt >= 16, the shortest trace the chunk kernels take."""
    NTR, CS = ntr_cs
    for nt in range(16, 5000):
        last = ((nt - 1) // CS) * CS
        reach = (NTR - 1) * nt + last + CS - 1   # from the block's first trace
        assert reach < trace_alloc(1, nt)  # a block starting at the last valid trace


@pytest.mark.parametrize("ntr_cs", SHAPES)
def test_swizzle_is_conflict_free(ntr_cs):
    NTR, CS = ntr_cs
    QPR = CS // 4
    idx = [chunk_at(j, s, CS) for j in range(NTR) for s in range(CS)]
    assert sorted(idx) == list(range(NTR * CS))
    # quads stay whole and 16-byte aligned
    for j in range(NTR):
        for qd in range(QPR):
            base = chunk_at(j, 4 * qd, CS)
            assert base % 4 == 0
            assert [chunk_at(j, 4 * qd + e, CS) for e in range(4)] == [base + e for e in range(4)]
    # the scan: lane t at row t reads sample k; distinct banks for up to 16 lanes
    for k in range(CS):
        banks = [chunk_at(t, k, CS) % 64 for t in range(min(NTR, 16))]
        assert len(set(banks)) == len(banks)


def test_replay_catches_a_short_allocation():
    """A buffer without the pad (ntraces x nt) is caught by the bound check."""
    G = Geo(1, 65, 17, 0, 1, 0, -1, seed=3)
    G.alloc = 65 * 17
    with pytest.raises(AssertionError):
        replay(G, 16, 256)

"""The synthetic workloads of the benchmark and the parity tests.

small   v(z) = 1.5 + 0.5 z km/s, nz=201, nx=301, dz=dx=0.01 km; 61 shots every
        0.05 km at z=0; 201 receivers at 0.015 km (cmp=0, absolute receiver
        axis); eikods2d tables (fast marching, box 2) at the 61 shot and 201
        receiver positions; nt=1001, dt=2 ms; 15 Hz Ricker at the two-way times of
        flat reflectors z = 0.6 / 1.2 / 1.7 km (r = +0.1 / -0.1 / +0.1) plus
        Gaussian noise at 20 dB SNR; angle axis 0..60 deg, 61 bins.
layers  the salt model cut into five depth layers migrated as separate
        problems (nz 80/70/90/100/60 from the top; shots 768/640/512/384/256 and
        receivers 401/361/321/281/241 spread over the same 0..9.59 / 0..10 km;
        per-layer mutes, slopes 50/40/30/20/10 s/km, tapers 0.05..0.10 s);
        eikods2d tables inside each layer from its top (the datum the
        recursive workflow moves the acquisition to), Born data from scatterers
        inside each layer; angle axis 0..60 deg for every layer, so the five
        CIGs stack into one (400, 1000, 61) angle gather volume.
salt    five-layer model 1.5 / 2.0 / 2.6 / 4.5 (salt, Gaussian-bump top) /
        3.2 km/s smoothed with a 5-point Gaussian, nz=400, nx=1000,
        dz=dx=0.01 km; 768 shots every 0.0125 km; 401 receivers every
        0.025 km (cmp=0); nt=2001, dt=2 ms; 20 Hz Ricker point-scatterer
        (Born-style) events plus a direct arrival; direct_mute=1 with slope
        50 s/km, intercept 0, taper 0.05 s; angle axis 0..60 deg, 61 bins.

Velocity models and data come from libkmsynth (own C++, seeded RNG) and the
traveltime tables from libkmeik (kmig_eikods2d, bitwise the reference's
eikods2d), so the inputs are identical on every machine.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_SYN = None


def synth():
    global _SYN
    if _SYN is None:
        path = os.path.join(HERE, "lib", "libkmsynth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        fp = C.POINTER(C.c_float)
        f, i = C.c_float, C.c_int
        L.kms_velocity_vz.argtypes = [fp, i, i, f, f, f]
        L.kms_velocity_salt.argtypes = [fp, i, i, f, f]
        L.kms_time_vz.restype = C.c_double
        L.kms_time_vz.argtypes = [C.c_double] * 5
        L.kms_data_reflectors.argtypes = [fp, i, f, f, i, f, f, i, f, f, i, f, f, i, fp, fp, f,
                                          f, C.c_uint64]
        L.kms_data_scatterers.argtypes = [fp, i, f, f, i, f, f, i, f, f, i, fp, i, f, f, fp, i,
                                          f, f, i, i, i, f, f, C.c_uint64]
        L.kms_velocity_grf.argtypes = [fp, i, i, f, f, f, f, f, C.c_uint64]
        _SYN = L
    return _SYN


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class Workload:
    name: str
    kv: dict
    data: np.ndarray            # (ns, nh, nt)
    data_axes: list             # [[nt,t0,dt],[nh,h0,dh],[ns,s0,ds]]
    stab: np.ndarray            # (npos, nx, nz)
    sder: np.ndarray
    s_axes: list                # [[nz,z0,dz],[nx,x0,dx],[npos,o,d]]
    rtab: np.ndarray
    rder: np.ndarray
    r_axes: list
    vel: np.ndarray = field(default=None)
    z_range: tuple = field(default=None)  # full-model depth samples [start, end] of a layer

    @property
    def updates(self) -> float:
        (nt, _, _), (nh, _, _), (ns, _, _) = self.data_axes
        nz, nx = self.s_axes[0][0], self.s_axes[1][0]
        return float(ns) * nh * nz * nx

    def args(self):
        return (self.data, self.data_axes, self.stab, self.sder, self.s_axes, self.rtab,
                self.rder, self.r_axes)

    def shot_subset(self, shots):
        """Same workload restricted to the given shot indices (evenly spaced)."""
        idx = np.asarray(shots)
        (nt, t0, dt), hax, (ns, s0, ds) = self.data_axes
        if len(idx) > 1:
            step = int(idx[1] - idx[0])
            assert np.all(np.diff(idx) == step)
        else:
            step = 1
        sub_axes = [[nt, t0, dt], hax, [len(idx), s0 + idx[0] * ds, ds * step]]
        return Workload(self.name + f"[{len(idx)} shots]", self.kv,
                        np.ascontiguousarray(self.data[idx]), sub_axes, self.stab, self.sder,
                        self.s_axes, self.rtab, self.rder, self.r_axes, self.vel)


def tables(vel, dz, dx, npos, o, d, box=2, threads=0):
    """Traveltime tables and their source derivatives at npos surface
    positions, made as the reference's workflow makes them
    (run_salt_adcig.sh:108-119): shotgen positions o + k d in float, then
    eikods2d with a box of `box` samples (kmig_eikods2d, bitwise the
    program's time= and tds1= outputs).  vel: (nx, nz), z fastest."""
    from . import eik
    shots = np.zeros((npos, 3), np.float32)
    shots[:, 1] = np.float32(o) + np.arange(npos, dtype=np.float32) * np.float32(d)
    r = eik.eikods2d(vel, np.float32(dz), np.float32(dx), shots, b1=box, b2=box,
                     threads=threads, grads=False, tdl1=False)
    return r["time"], r["tds1"]


def small(threads=0, seed=20260101) -> Workload:
    L = synth()
    nz, nx, d = 201, 301, 0.01
    vel = np.empty((nx, nz), np.float32)
    L.kms_velocity_vz(_p(vel), nz, nx, d, 1.5, 0.5)
    ns, s0, ds = 61, 0.0, 0.05
    nh, h0, dh = 201, 0.0, 0.015
    nt, dt = 1001, 0.002
    ST, SD = tables(vel, d, d, ns, s0, ds, 2, threads)
    RT, RD = tables(vel, d, d, nh, h0, dh, 2, threads)
    data = np.empty((ns, nh, nt), np.float32)
    zr = np.array([0.6, 1.2, 1.7], np.float32)
    amp = np.array([0.1, -0.1, 0.1], np.float32)
    L.kms_data_reflectors(_p(data), nt, 0.0, dt, nh, h0, dh, ns, s0, ds, 0, 1.5, 0.5, 3, _p(zr),
                          _p(amp), 15.0, 20.0, seed)
    kv = dict(cmp=0, mode="angle", angle_min=0, angle_max=60, angle_step=1, angle_n=61)
    tax = [[nz, 0.0, d], [nx, 0.0, d]]
    return Workload("small v(z) 201x301, 61 shots x 201 receivers, nt=1001", kv, data,
                    [[nt, 0.0, dt], [nh, h0, dh], [ns, s0, ds]], ST, SD, tax + [[ns, s0, ds]],
                    RT, RD, tax + [[nh, h0, dh]], vel)


def salt(threads=0, seed=20260202, nshots=768, cmp=0, stack=False) -> Workload:
    """The salt workload.  cmp=0 (the workflow's geometry): 401 fixed
    receivers every 0.025 km.  cmp=1 (the program's default geometry): 401
    split-spread offsets -2.5 .. 2.5 km every 0.0125 km moving with the shot,
    receiver tables at the 800 positions 0 .. 9.9875 km every 0.0125 km
    (receivers beyond the model edge take the edge table, as the reference's
    clamped table lookup does).  stack=True: the stacked image (cig=0, the
    program's default output) instead of 61 angle bins."""
    L = synth()
    nz, nx, d = 400, 1000, 0.01
    vel = np.empty((nx, nz), np.float32)
    L.kms_velocity_salt(_p(vel), nz, nx, d, d)
    ns, s0, ds = nshots, 0.0, 0.0125
    nt, dt = 2001, 0.002
    if cmp:
        nh, h0, dh = 401, -2.5, 0.0125
        nr, r0, dr = 800, 0.0, 0.0125
    else:
        nh, h0, dh = 401, 0.0, 0.025
        nr, r0, dr = nh, h0, dh
    ST, SD = tables(vel, d, d, ns, s0, ds, 2, threads)
    RT, RD = tables(vel, d, d, nr, r0, dr, 2, threads)
    data = np.empty((ns, nh, nt), np.float32)
    L.kms_data_scatterers(_p(data), nt, 0.0, dt, nh, h0, dh, ns, s0, ds, int(cmp), _p(ST), ns,
                          s0, ds, _p(RT), nr, r0, dr, nz, nx, 60, 20.0, 1.5, seed)
    kv = dict(cmp=int(cmp), mode="angle", angle_min=0, angle_max=60, angle_step=1, angle_n=61,
              direct_mute=1, direct_mute_slope=50, direct_mute_intercept=0,
              direct_mute_taper=0.05)
    tax = [[nz, 0.0, d], [nx, 0.0, d]]
    geo = "401 offsets -2.5..2.5 km (cmp=1)" if cmp else "401 receivers"
    if stack:
        kv = {k: v for k, v in kv.items() if not k.startswith(("mode", "angle_"))}
        kv["cig"] = 0
        geo += ", stacked image"
    return Workload(f"salt 400x1000, {ns} shots x {geo}, nt=2001", kv, data,
                    [[nt, 0.0, dt], [nh, h0, dh], [ns, s0, ds]], ST, SD, tax + [[ns, s0, ds]],
                    RT, RD, tax + [[nr, r0, dr]], vel)


@dataclass
class DatumWorkload:
    name: str
    kv: dict
    data: np.ndarray            # (ns, nh, nt)
    data_axes: list             # [[nt,t0,dt],[nh,h0,dh],[ns,s0,ds]]
    green: np.ndarray           # (npos, npos) datum table (sources and receivers)
    pos_axis: tuple             # (o, d) of the table positions
    datum: float                # datum depth, km

    def corner(self, n):
        """The first n shots x n receivers of the cube (same tables, datum
        and keys): a bounded sample of the same operator."""
        (nt, t0, dt), (_, h0, dh), (_, s0, ds) = self.data_axes
        return DatumWorkload(self.name + f"[{n}x{n} corner]", self.kv,
                             np.ascontiguousarray(self.data[:n, :n]),
                             [[nt, t0, dt], [n, h0, dh], [n, s0, ds]], self.green,
                             self.pos_axis, self.datum)


def datum(threads=0, seed=20260404, npos=768, nt=2001, should_datum=80) -> DatumWorkload:
    """The datuming step of the recursive workflow (run_salt_adcig.sh:155-166)
    on the salt model: 768 shots and receivers at the model's 10 m sampling
    (shotgen nshot=nx=768, dy=dx=0.01, :18-20, :111-112), absolute receiver
    coordinates (cmp=0, :20), the datum at the base of the first layer
    (80 samples, 0.8 km), the workflow's keys (aperture=200 data_taper=20
    taper=0 length=0.05 antialias=1 use_bf=1); datum tables cut from
    fast-marching traveltimes as greensol_auto does, synthetic diffraction
    data of 2001 samples at 2 ms."""
    from . import dat
    L = synth()
    nz, nx, d = 400, 1000, 0.01
    vel = np.empty((nx, nz), np.float32)
    L.kms_velocity_salt(_p(vel), nz, nx, d, d)
    T, _ = tables(vel, d, d, npos, 0.0, d, 2, threads)
    green = dat.greensol(T, [[nz, 0.0, d], [nx, 0.0, d]], (0.0, d), should_datum * d)
    dt = 0.002
    data = np.empty((npos, npos, nt), np.float32)
    L.kms_data_scatterers(_p(data), nt, 0.0, dt, npos, 0.0, d, npos, 0.0, d, 0, _p(T), npos,
                          0.0, d, _p(T), npos, 0.0, d, nz, nx, 60, 20.0, 1.5, seed)
    kv = dict(cmp=0, aperture=200, data_taper=20, taper=0, length=0.05, antialias=1, use_bf=1)
    return DatumWorkload(f"datum salt {npos} shots x {npos} receivers, nt={nt}, datum "
                         f"{should_datum * d:.2f} km", kv, data,
                         [[nt, 0.0, dt], [npos, 0.0, d], [npos, 0.0, d]], green, (0.0, d),
                         should_datum * d)


LAYERS_NZ = (80, 70, 90, 100, 60)
LAYERS_NS = (768, 640, 512, 384, 256)
LAYERS_NH = (401, 361, 321, 281, 241)
LAYERS_SLOPE = (50.0, 40.0, 30.0, 20.0, 10.0)
LAYERS_TAPER = (0.05, 0.0625, 0.075, 0.0875, 0.10)


def layers(threads=0, seed=20260303, scale=1.0, overlap=0):
    """The five depth layers as separate Workloads (see module docstring).
    scale < 1 shrinks shot and receiver counts (for tests).  overlap > 0
    extends every layer after the first upward by that many samples (the
    overlapping layers the depth-layer stitch blends); each Workload's
    z_range is its [start, end] in full-model depth samples."""
    L = synth()
    nx, d = 1000, 0.01
    full = np.empty((nx, 400), np.float32)
    L.kms_velocity_salt(_p(full), 400, nx, d, d)
    out, zb = [], 0
    for k, nz0 in enumerate(LAYERS_NZ):
        z0 = zb - (overlap if k > 0 else 0)
        nz = nz0 + (zb - z0)
        vel = np.ascontiguousarray(full[:, z0:z0 + nz])
        ns = max(2, int(round(LAYERS_NS[k] * scale)))
        nh = max(2, int(round(LAYERS_NH[k] * scale)))
        ds, dh = 9.5875 / (ns - 1), 10.0 / (nh - 1)
        nt, dt = 2001, 0.002
        ST, SD = tables(vel, d, d, ns, 0.0, ds, 2, threads)
        RT, RD = tables(vel, d, d, nh, 0.0, dh, 2, threads)
        data = np.empty((ns, nh, nt), np.float32)
        L.kms_data_scatterers(_p(data), nt, 0.0, dt, nh, 0.0, dh, ns, 0.0, ds, 0, _p(ST), ns,
                              0.0, ds, _p(RT), nh, 0.0, dh, nz, nx, 30, 20.0, 1.5, seed + k)
        kv = dict(cmp=0, mode="angle", angle_min=0, angle_max=60, angle_step=1, angle_n=61,
                  direct_mute=1, direct_mute_slope=LAYERS_SLOPE[k], direct_mute_intercept=0,
                  direct_mute_taper=LAYERS_TAPER[k])
        tax = [[nz, 0.0, d], [nx, 0.0, d]]
        out.append(Workload(f"layer {k + 1}: {nz}x{nx}, {ns} shots x {nh} receivers", kv, data,
                            [[nt, 0.0, dt], [nh, 0.0, dh], [ns, 0.0, ds]], ST, SD,
                            tax + [[ns, 0.0, ds]], RT, RD, tax + [[nh, 0.0, dh]], vel,
                            (z0, z0 + nz - 1)))
        zb += nz0
    return out


def stack_layers(cigs):
    """Concatenate per-layer CIGs (ng, nx, nz_k) along depth into one volume."""
    return np.concatenate(cigs, axis=2)


# ----------------------------------------------------------- surrogate --
# The learned traveltime surrogate of the small workload: inputs scaled over
# its 3 x 2 km section (positions 0..3 km), reference slowness 1/2 s/km.
SURROGATE = dict(s_min=0.0, s_max=3.0, z_min=0.0, z_max=2.0, x_min=0.0, x_max=3.0, z_src=0.0,
                 slowness=0.5, sigma_t=0.001, sigma_g=0.3, batch=16384, lr=2e-3, lr_final=2e-5)
# The same network over the salt section (10 x 4 km, positions 0..10 km; the
# salt model's slowness runs 1/4.5 .. 1/1.5 s/km).
SURROGATE_SALT = dict(s_min=0.0, s_max=10.0, z_min=0.0, z_max=4.0, x_min=0.0, x_max=10.0,
                      z_src=0.0, slowness=0.4, sigma_t=0.001, sigma_g=0.3, batch=16384,
                      lr=2e-3, lr_final=2e-5)


def table_records(T, D, pos, nper, seed, dz=0.01, dx=0.01):
    """Training records (s, z, x, T, dT/dx, dT/dz, dT/ds) drawn from position
    tables T, D (npos, nx, nz): nper random grid nodes per position; dT/dx and
    dT/dz are centred differences of the table (one-sided at the edges)."""
    rng = np.random.default_rng(seed)
    _, nx, nz = T.shape

    def diff(Tk, i, j, axis, n, h):
        # numpy.gradient's stencil at the sampled nodes only: centred inside,
        # one-sided first order at the edges
        lo, hi = np.maximum(i - 1, 0), np.minimum(i + 1, n - 1)
        a = Tk[hi, j] if axis == 0 else Tk[j, hi]
        b = Tk[lo, j] if axis == 0 else Tk[j, lo]
        return (a - b) / ((hi - lo) * h)

    out = []
    for k, s in enumerate(pos):
        Tk = T[k]
        ix = rng.integers(0, nx, nper)
        iz = rng.integers(0, nz, nper)
        gx = diff(Tk.astype(np.float64, copy=False), ix, iz, 0, nx, dx)
        gz = diff(Tk.astype(np.float64, copy=False), iz, ix, 1, nz, dz)
        out.append(np.stack([np.full(nper, s), iz * dz, ix * dx, Tk[ix, iz].astype(np.float64),
                             gx, gz, D[k][ix, iz]], axis=1))
    return np.concatenate(out).astype(np.float32)


def grf_records(nmodels=256, npos=8, nper=2048, seed=20260505, threads=0, corr=0.3,
                window=(201, 301), s_hi=3.0):
    """Records from eikods2d tables (kmig_eikods2d, box 2) of nmodels seeded
    Gaussian-random-field velocity models (nz=400, nx=1000, dz=dx=0.01 km, v in
    1.5..4.5 km/s, correlation length corr km): npos random surface sources in
    [0, s_hi] km per model, nper records per source inside the top-left window
    (nz, nx) of the small workload; dT/ds is the table's source derivative."""
    from . import eik
    L = synth()
    nz, nx = 400, 1000
    v = np.empty((nx, nz), np.float32)
    rng = np.random.default_rng(seed)
    out = []
    for m in range(nmodels):
        L.kms_velocity_grf(_p(v), nz, nx, 0.01, 0.01, corr, 1.5, 4.5, seed + 7919 * m)
        pos = np.sort(rng.uniform(0.0, s_hi, npos)).astype(np.float32)
        shots = np.zeros((npos, 3), np.float32)
        shots[:, 1] = pos
        r = eik.eikods2d(v, np.float32(0.01), np.float32(0.01), shots, b1=2, b2=2,
                         threads=threads, grads=False, tdl1=False)
        T = r["time"][:, :window[1], :window[0]]
        D = r["tds1"][:, :window[1], :window[0]]
        out.append(table_records(T, D, pos, nper, seed + 104729 * (m + 1)))
    return np.concatenate(out)

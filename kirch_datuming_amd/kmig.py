"""ctypes binding of libkmig.so, the C ABI declared in include/kmig.h.

Mirrors the reference's interface for the adj=1 path: the same key=value
parameters as kirchmig2d_cig (src/kirchmig2d_cig.cpp:548-875), the same array
layouts as its RSF files, and an exception carrying the reference's ERROR
message wherever the reference would print it and exit(1).

There is no fallback: if lib/libkmig.so is missing or no GPU is present,
the calls raise.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KMIG_LIB") or os.path.join(HERE, "lib", "libkmig.so")

KMIG_OK = 0
ERRORS = {-1: "KMIG_E_PARAM", -2: "KMIG_E_DIM", -3: "KMIG_E_IO", -4: "KMIG_E_DEVICE",
          -5: "KMIG_E_STATE", -6: "KMIG_E_NOMEM"}
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2, "u16": 3}
VARIANT = {"auto": 0, "stream": 1, "resident": 2}

# every exported symbol of include/kmig.h
EXPORTS = ("kmig_params_default", "kmig_params_parse", "kmig_params_check", "kmig_create",
           "kmig_destroy", "kmig_last_error", "kmig_set_tables", "kmig_set_data", "kmig_run",
           "kmig_sync", "kmig_get_cig", "kmig_get_qc", "kmig_get_stats", "kmig_migrate",
           "kmig_stream", "kmig_version", "kmig_batch_create", "kmig_batch_run",
           "kmig_batch_sync", "kmig_batch_stats", "kmig_batch_last_error", "kmig_batch_destroy",
           "kmig_set_surrogate", "kmig_get_surrogate_stats", "kmig_batch_stitch",
           "kmig_model", "kmig_count_updates", "kmig_batch_stitch_noalloc", "kmig_set_graph",
           "kmig_set_surrogate_fallback", "kmig_probe_gather")


class KmigError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


class KmigParams(C.Structure):
    _fields_ = [
        ("adj", C.c_int), ("cig", C.c_int), ("cig_is_angle", C.c_int), ("cmp", C.c_int),
        ("half_angle", C.c_int),
        ("oazmin", C.c_float), ("oazmax", C.c_float), ("oaztaper", C.c_float),
        ("diff", C.c_int), ("aperture_trace", C.c_int), ("aperture_trace_taper", C.c_int),
        ("direct_mute", C.c_int), ("direct_mute_invert", C.c_int),
        ("direct_mute_qc_shot", C.c_int),
        ("direct_mute_slope", C.c_float), ("direct_mute_intercept", C.c_float),
        ("direct_mute_taper", C.c_float),
        ("tau", C.c_float), ("aperture", C.c_float), ("aperture_taper", C.c_float),
        ("aperture_rmin", C.c_float), ("aperture_rtaper", C.c_float), ("antialias", C.c_float),
        ("angle_interp", C.c_int), ("angle_taper_start", C.c_float),
        ("angle_taper_end", C.c_float),
        ("angle_grad_radius", C.c_int), ("angle_smooth", C.c_int), ("interp_type", C.c_int),
        ("g0", C.c_float), ("dg", C.c_float), ("ng", C.c_int),
        ("precision", C.c_int), ("tile", C.c_int), ("variant", C.c_int),
    ]


class KmigStats(C.Structure):
    _fields_ = [("updates", C.c_double), ("ms_trace_prep", C.c_double),
                ("ms_tuple_prep", C.c_double), ("ms_migrate", C.c_double),
                ("ms_total", C.c_double), ("migrate_launches", C.c_int),
                ("receiver_slots", C.c_int), ("shot_batches", C.c_int),
                ("variant", C.c_int), ("tile", C.c_int), ("graph", C.c_int),
                ("tuple_direct", C.c_int)]


_LIB = None


def lib():
    """Load libkmig.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        vp = C.c_void_p
        L.kmig_params_default.argtypes = [C.POINTER(KmigParams)]
        L.kmig_params_default.restype = None
        L.kmig_params_parse.argtypes = [C.POINTER(KmigParams), C.c_int, C.POINTER(C.c_char_p),
                                        C.c_int, C.c_float, C.c_float, C.c_char_p, C.c_int]
        L.kmig_params_check.argtypes = [C.POINTER(KmigParams), C.c_int, C.c_char_p, C.c_int]
        L.kmig_create.argtypes = [C.POINTER(vp), C.POINTER(KmigParams)]
        L.kmig_destroy.argtypes = [vp]
        L.kmig_destroy.restype = None
        L.kmig_last_error.argtypes = [vp]
        L.kmig_last_error.restype = C.c_char_p
        L.kmig_set_tables.argtypes = [vp, C.c_int, fp, fp, fp, fp, C.c_int, C.c_int, C.c_int,
                                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float]
        L.kmig_set_data.argtypes = [vp, fp, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                    C.c_float, C.c_int, C.c_float, C.c_float]
        L.kmig_run.argtypes = [vp, C.c_int, C.c_int]
        L.kmig_model.argtypes = [vp, fp, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                 C.c_float, C.c_int, C.c_float, C.c_float, fp]
        L.kmig_sync.argtypes = [vp]
        L.kmig_get_cig.argtypes = [vp, fp]
        L.kmig_get_qc.argtypes = [vp, fp, fp]
        L.kmig_get_stats.argtypes = [vp, C.POINTER(KmigStats)]
        dp = C.POINTER(C.c_double)
        L.kmig_count_updates.argtypes = [vp, dp, dp, dp]
        L.kmig_set_graph.argtypes = [vp, C.c_int]
        L.kmig_stream.argtypes = [vp]
        L.kmig_stream.restype = vp
        L.kmig_version.restype = C.c_char_p
        L.kmig_probe_gather.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.kmig_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int]
        L.kmig_batch_run.argtypes = [vp]
        L.kmig_batch_sync.argtypes = [vp]
        L.kmig_batch_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
        L.kmig_batch_last_error.argtypes = [vp]
        L.kmig_batch_last_error.restype = C.c_char_p
        L.kmig_batch_destroy.argtypes = [vp]
        L.kmig_batch_destroy.restype = None
        L.kmig_batch_stitch.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_float), C.c_int, C.c_int, C.c_double,
                                        C.c_int, C.c_int, fp]
        L.kmig_batch_stitch_noalloc.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_float), C.c_int, C.c_int,
                                                C.c_double, C.c_int, C.c_int, fp]
        L.kmig_set_surrogate.argtypes = [vp, C.c_int, vp, C.c_float]
        L.kmig_set_surrogate_fallback.argtypes = [vp, C.c_int, C.c_int]
        L.kmig_get_surrogate_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_float)]
        _LIB = L
    return _LIB


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def parse_params(kv, nh: int, h0: float, dh: float) -> KmigParams:
    """key=value dict or list -> KmigParams with the reference's semantics."""
    L = lib()
    if isinstance(kv, dict):
        items = [f"{k}={v}" for k, v in kv.items()]
    else:
        items = list(kv)
    arr = (C.c_char_p * max(1, len(items)))(*[s.encode() for s in items])
    p = KmigParams()
    err = C.create_string_buffer(512)
    rc = L.kmig_params_parse(C.byref(p), len(items), arr, int(nh), float(h0), float(dh), err,
                             512)
    if rc != KMIG_OK:
        raise KmigError(rc, err.value.decode())
    return p


def check_params(p: KmigParams, ns: int):
    err = C.create_string_buffer(512)
    rc = lib().kmig_params_check(C.byref(p), int(ns), err, 512)
    if rc != KMIG_OK:
        raise KmigError(rc, err.value.decode())


class Migrator:
    """One kmig context: resident tables and data on the current GPU."""

    def __init__(self, params: KmigParams):
        self.L = lib()
        self.p = params
        self.ctx = C.c_void_p()
        rc = self.L.kmig_create(C.byref(self.ctx), C.byref(params))
        if rc != KMIG_OK:
            msg = self._err()
            self.close()
            raise KmigError(rc, msg)
        self.nzx = None
        self.shape = None

    def _err(self):
        return self.L.kmig_last_error(self.ctx).decode() if self.ctx else ""

    def _chk(self, rc):
        if rc != KMIG_OK:
            raise KmigError(rc, self._err())

    def set_tables(self, side, T, dTds, axes, gx=None, gz=None):
        """axes = [[nz,z0,dz],[nx,x0,dx],[npos,o,d]] (RSF order)."""
        T, dTds, gx, gz = _f32(T), _f32(dTds), _f32(gx), _f32(gz)
        (nz, z0, dz), (nx, x0, dx), (npos, o, d) = axes[:3]
        self._hold = (T, dTds, gx, gz)
        self._chk(self.L.kmig_set_tables(self.ctx, side, _fp(T), _fp(dTds), _fp(gx), _fp(gz),
                                         int(nz), int(nx), int(npos), z0, dz, x0, dx, o, d))
        self.nz, self.nx = int(nz), int(nx)

    def set_data(self, data, axes):
        data = _f32(data)
        (nt, t0, dt), (nh, h0, dh), (ns, s0, ds) = axes[:3]
        self._chk(self.L.kmig_set_data(self.ctx, _fp(data), int(nt), t0, dt, int(nh), h0, dh,
                                       int(ns), s0, ds))
        self.nt, self.nh, self.ns = int(nt), int(nh), int(ns)

    def run(self, shots=None):
        sb, se = (0, self.ns) if shots is None else shots
        self._chk(self.L.kmig_run(self.ctx, int(sb), int(se)))

    def model(self, image, axes):
        """adj=0: the image (ng, nx, nz) modelled into data (ns, nh, nt) on the
        data axes [[nt, t0, dt], [nh, h0, dh], [ns, s0, ds]] (synchronous)."""
        img = _f32(image)
        (nt, t0, dt), (nh, h0, dh), (ns, s0, ds) = axes[:3]
        out = np.empty((int(ns), int(nh), int(nt)), np.float32)
        self._chk(self.L.kmig_model(self.ctx, _fp(img), int(nt), t0, dt, int(nh), h0, dh,
                                    int(ns), s0, ds, _fp(out)))
        self.nt, self.nh, self.ns = int(nt), int(nh), int(ns)
        return out

    def sync(self):
        self._chk(self.L.kmig_sync(self.ctx))

    def cig(self):
        ng = self.p.ng if self.p.cig else 1
        out = np.empty((ng, self.nx, self.nz), np.float32)
        self._chk(self.L.kmig_get_cig(self.ctx, _fp(out)))
        return out

    def qc(self):
        b = np.empty((self.nh, self.nt), np.float32)
        a = np.empty((self.nh, self.nt), np.float32)
        self._chk(self.L.kmig_get_qc(self.ctx, _fp(b), _fp(a)))
        return b, a

    def stats(self) -> dict:
        st = KmigStats()
        self._chk(self.L.kmig_get_stats(self.ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in KmigStats._fields_}

    def stream(self):
        return self.L.kmig_stream(self.ctx)

    def set_graph(self, enable=True):
        """Capture the step as a hipGraph after its first eager run and replay
        it on later runs with the same plan (kmig_set_graph)."""
        self._chk(self.L.kmig_set_graph(self.ctx, int(bool(enable))))

    def count_updates(self) -> dict:
        """Nominal, evaluated (groups surviving the culls) and accepted
        (taps inside the trace, nonzero weight) updates of the last run's
        resident launch, from a counting re-run (synchronous)."""
        n, e, a = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.kmig_count_updates(self.ctx, C.byref(n), C.byref(e), C.byref(a)))
        return {"nominal": n.value, "evaluated": e.value, "accepted": a.value}

    def set_surrogate(self, side, net, max_dt=1.0e-3):
        """Take traveltimes of `side` from a trained nn.Surrogate wherever they
        are within max_dt seconds of the table (net=None removes it)."""
        h = net.h if net is not None else None
        if not hasattr(self, "_nets"):
            self._nets = {}
        self._nets[side] = net  # keep the network alive while the context uses it
        self._chk(self.L.kmig_set_surrogate(self.ctx, int(side), h, float(max_dt)))

    def set_surrogate_fallback(self, side, per_slot=True):
        """Per-slot (default: a shot or receiver position takes the network
        everywhere or nowhere) or per-point fallback to the tables."""
        self._chk(self.L.kmig_set_surrogate_fallback(self.ctx, int(side), 0 if per_slot else 1))

    def surrogate_stats(self, side) -> dict:
        pts, used, mx = C.c_double(), C.c_double(), C.c_float()
        self._chk(self.L.kmig_get_surrogate_stats(self.ctx, int(side), C.byref(pts),
                                                  C.byref(used), C.byref(mx)))
        return {"points": pts.value, "used": used.value, "max_residual": mx.value}

    def close(self):
        if self.ctx:
            self.L.kmig_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several Migrators run as one graph-captured step (one migrate launch)."""

    def __init__(self, migrators):
        self.L = lib()
        self.m = list(migrators)
        arr = (C.c_void_p * len(self.m))(*[mm.ctx.value for mm in self.m])
        self.b = C.c_void_p()
        rc = self.L.kmig_batch_create(C.byref(self.b), arr, len(self.m))
        if rc != KMIG_OK:
            msg = self.L.kmig_batch_last_error(self.b).decode() if self.b else ""
            self.close()
            raise KmigError(rc, msg)

    def _chk(self, rc):
        if rc != KMIG_OK:
            raise KmigError(rc, self.L.kmig_batch_last_error(self.b).decode())

    def run(self):
        self._chk(self.L.kmig_batch_run(self.b))

    def sync(self):
        self._chk(self.L.kmig_batch_sync(self.b))

    def stats(self):
        u, ms, nodes = C.c_double(), C.c_double(), C.c_int()
        self._chk(self.L.kmig_batch_stats(self.b, C.byref(u), C.byref(ms), C.byref(nodes)))
        return {"updates": u.value, "ms": ms.value, "graph_nodes": nodes.value}

    def stitch(self, start, end, scale=None, stitch_blend=15, scalex_aper=0, scalex_scale=1.0,
               angles=None):
        """Depth-layer stitch of the batch's gathers (layercom_cig): layer k
        covers full depth samples [start[k], end[k]]; angles = (first, count)
        windows the angle axis.  Returns (count, nx, end[-1] + 1) float32."""
        n = len(self.m)
        st = (C.c_int * n)(*[int(v) for v in start])
        en = (C.c_int * n)(*[int(v) for v in end])
        sc = (C.c_float * n)(*[float(v) for v in scale]) if scale is not None else None
        m0 = self.m[0]
        ng = m0.p.ng if m0.p.cig else 1
        g0, gn = angles if angles is not None else (0, ng)
        out = np.empty((gn, m0.nx, int(end[-1]) + 1), np.float32)
        self._chk(self.L.kmig_batch_stitch(self.b, st, en, sc, int(stitch_blend),
                                           int(scalex_aper), float(scalex_scale), int(g0),
                                           int(gn), _fp(out)))
        return out

    def stitch_into(self, out_ptr, start, end, scale=None, stitch_blend=15, scalex_aper=0,
                    scalex_scale=1.0, angles=None):
        """The same stitch written to caller-owned DEVICE memory at address
        out_ptr (an int; at least count * nx * (end[-1] + 1) floats), through
        kmig_batch_stitch_noalloc: no device allocation in the call."""
        n = len(self.m)
        st = (C.c_int * n)(*[int(v) for v in start])
        en = (C.c_int * n)(*[int(v) for v in end])
        sc = (C.c_float * n)(*[float(v) for v in scale]) if scale is not None else None
        m0 = self.m[0]
        ng = m0.p.ng if m0.p.cig else 1
        g0, gn = angles if angles is not None else (0, ng)
        self._chk(self.L.kmig_batch_stitch_noalloc(self.b, st, en, sc, int(stitch_blend),
                                                   int(scalex_aper), float(scalex_scale),
                                                   int(g0), int(gn),
                                                   C.cast(C.c_void_p(int(out_ptr)),
                                                          C.POINTER(C.c_float))))

    def close(self):
        if self.b:
            self.L.kmig_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def extract_adcig(adcig, ix=None, x=None, o2=0.0, d2=1.0):
    """One lateral position of a (angle, x, z) gather volume as an (angle, z)
    gather (src/extract_adcig.cpp:126-281): by index ix, or by position x on
    the lateral axis o2 + i * d2, rounded to the nearest sample as there
    (:163-177); out-of-range requests raise with the reference's messages."""
    ng, nx, nz = adcig.shape
    if ix is None:
        if x is None:
            raise KmigError(-1, "ix= or x= is required")
        if d2 == 0.0:
            raise KmigError(-1, "Cannot use x= because d2=0")
        if not np.isfinite(x):
            raise KmigError(-1, "Invalid x= value")
        q = (float(x) - o2) / d2
        nearest = int(np.floor(abs(q) + 0.5)) * (1 if q >= 0 else -1)  # llround
        if not 0 <= nearest < nx:
            raise KmigError(-1, f"x={x:g} is outside lateral axis [{o2:g}, {o2 + (nx - 1) * d2:g}]")
        ix = nearest
    if not 0 <= ix < nx:
        raise KmigError(-1, f"ix must be in [0,{nx - 1}]")
    return np.ascontiguousarray(adcig[:, ix, :])


def model(kv, image, data_axes, stab, sder, s_axes, rtab, rder, r_axes, grads=None):
    """Full adj=0 run on the GPU: the image (ng, nx, nz) modelled into data
    (ns, nh, nt) on data_axes."""
    kv = dict(kv, adj=0)
    p = parse_params(kv, int(data_axes[1][0]), float(data_axes[1][1]), float(data_axes[1][2]))
    check_params(p, int(data_axes[2][0]))
    with Migrator(p) as m:
        g = grads or [None] * 4
        m.set_tables(0, stab, sder, s_axes, g[0], g[1])
        m.set_tables(1, rtab, rder, r_axes, g[2], g[3])
        return m.model(image, data_axes)


def migrate(kv, data, data_axes, stab, sder, s_axes, rtab, rder, r_axes, grads=None,
            shots=None, extra=None):
    """Full adj=1 run on the GPU; returns (cig (ng,nx,nz), qc_before, qc_after)."""
    kv = dict(kv)
    if extra:
        kv.update(extra)
    p = parse_params(kv, int(data_axes[1][0]), float(data_axes[1][1]), float(data_axes[1][2]))
    check_params(p, int(data_axes[2][0]))
    with Migrator(p) as m:
        g = grads or [None] * 4
        m.set_tables(0, stab, sder, s_axes, g[0], g[1])
        m.set_tables(1, rtab, rder, r_axes, g[2], g[3])
        m.set_data(data, data_axes)
        m.run(shots)
        m.sync()
        cig = m.cig()
        qcb = qca = None
        if p.direct_mute and p.direct_mute_qc_shot >= 0:
            sb, se = (0, m.ns) if shots is None else shots
            if sb <= p.direct_mute_qc_shot < se:
                qcb, qca = m.qc()
        return cig, qcb, qca


def probe_gather(lanes_per_line=1, iters=2000):
    """Time the L1 gather-rate probe (kmig_probe_gather): returns (ms of the
    timed launch, its wave-level load instructions)."""
    L = lib()
    ms, n = C.c_double(), C.c_double()
    rc = L.kmig_probe_gather(int(lanes_per_line), int(iters), C.byref(ms), C.byref(n))
    if rc:
        raise KmigError(rc, "kmig_probe_gather failed")
    return ms.value, n.value

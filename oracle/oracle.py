"""ctypes binding of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module, and only to check or to time the CPU baseline.

``params_from_kv`` restates the reference's key=value handling for the adj=1
path (src/kirchmig2d_cig.cpp:548-875 and parse_output_axis :323-368), with
se_get_par_int = (int)double and se_get_par_float = (float)double
(SEFILESYSTEM/src/se_par_sep.cpp:432-446).
"""

from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class KmoParams(C.Structure):
    _fields_ = [
        ("adj", C.c_int), ("cig", C.c_int), ("cig_is_angle", C.c_int), ("cmp", C.c_int),
        ("half_angle", C.c_int),
        ("oazmin", C.c_float), ("oazmax", C.c_float), ("oaztaper", C.c_float),
        ("isdiff", C.c_int), ("aperture_trace", C.c_int), ("aperture_trace_taper", C.c_int),
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
    ]


class KmoGeom(C.Structure):
    _fields_ = [
        ("nt", C.c_int), ("nh", C.c_int), ("ns", C.c_int),
        ("t0", C.c_float), ("dt", C.c_float), ("h0", C.c_float), ("dh", C.c_float),
        ("s0", C.c_float), ("ds", C.c_float),
        ("nz", C.c_int), ("nx", C.c_int),
        ("z0", C.c_float), ("dz", C.c_float), ("x0", C.c_float), ("dx", C.c_float),
        ("sny", C.c_int), ("sy0", C.c_float), ("sdy", C.c_float),
        ("rny", C.c_int), ("ry0", C.c_float), ("rdy", C.c_float),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        for nm, outp in (("kmo_migrate_f32", fp), ("kmo_migrate_f64", dp)):
            f = getattr(L, nm)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(KmoParams), C.POINTER(KmoGeom), fp, fp, fp, fp, fp,
                          fp, fp, fp, fp, C.c_int, C.c_int, outp, fp, fp]
        for nm, outp in (("kmo_model_f32", fp), ("kmo_model_f64", dp)):
            f = getattr(L, nm)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(KmoParams), C.POINTER(KmoGeom), fp, fp, fp, fp, fp,
                          fp, fp, fp, fp, C.c_int, C.c_int, outp]
        L.kmo_pick_f32.restype = C.c_float
        L.kmo_pick_f32.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp]
        L.kmo_pick0_f64.restype = None
        L.kmo_pick0_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, dp]
        L.kmo_pick_f64.restype = C.c_double
        L.kmo_pick_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.kmo_doubint_f32.argtypes = [C.c_int, fp]
        L.kmo_doubint_f64.argtypes = [C.c_int, dp]
        L.kmo_interp_side_f32.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                          C.c_float, C.c_float, C.c_float, fp, fp, fp, fp,
                                          C.POINTER(C.c_int)]
        L.kmo_grad_plane_f32.argtypes = [fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                         fp, fp]
        L.kmo_smooth_f32.argtypes = [fp, C.c_int, C.c_int, C.c_int]
        L.kmo_angle_f32.restype = C.c_float
        L.kmo_angle_f32.argtypes = [C.c_float] * 4 + [C.c_int]
        L.kmo_linear_bins_f32.restype = C.c_int
        L.kmo_linear_bins_f32.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int,
                                          C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.kmo_mute_f32.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_int]
        for nm in ("kmo_leg_weight_f32", "kmo_floor_weight_f32", "kmo_angle_taper_f32"):
            getattr(L, nm).restype = C.c_float
            getattr(L, nm).argtypes = [C.c_float] * 3
        L.kmo_depth_mute_f32.restype = C.c_float
        L.kmo_depth_mute_f32.argtypes = [C.c_float, C.c_int, C.c_int, C.c_float, C.c_float,
                                         C.c_float]
        _LIB = L
    return _LIB


def _f(v) -> float:
    """se_get_par_float: parse as double, cast to float."""
    return float(np.float32(float(v)))


def _i(v) -> int:
    """se_get_par_int: parse as double, cast (truncate) to int32."""
    return int(float(v))


def output_axis(kv: dict, is_angle: bool, nh: int, h0: float, dh: float, half: bool):
    """parse_output_axis, src/kirchmig2d_cig.cpp:323-368 (float arithmetic)."""
    f32 = np.float32
    hend = f32(h0) + f32(nh - 1) * f32(dh)
    vmin = f32(0) if is_angle else min(f32(h0), hend)
    vmax = (f32(90) if half else f32(180)) if is_angle else max(f32(h0), hend)
    vstep = f32(1) if is_angle else (abs(f32(dh)) if f32(dh) != 0 else f32(1))
    nval = (91 if half else 181) if is_angle else nh
    given = False
    pre = ("angle" if is_angle else "offset")
    for p in ("cig", pre):
        if p + "_min" in kv:
            vmin = f32(_f(kv[p + "_min"]))
        if p + "_max" in kv:
            vmax = f32(_f(kv[p + "_max"]))
        if p + "_step" in kv:
            vstep = f32(_f(kv[p + "_step"]))
        if p + "_n" in kv:
            nval = _i(kv[p + "_n"])
            given = True
    if vstep <= 0:
        raise ValueError("CIG axis step must be positive.")
    if not given:
        nval = int(math.floor(f32((vmax - vmin) / vstep) + f32(0.5))) + 1
    if nval <= 0:
        raise ValueError("CIG axis sample number must be positive.")
    return float(vmin), float(vstep), int(nval)


def params_from_kv(kv: dict, nh: int, h0: float, dh: float) -> KmoParams:
    p = KmoParams()
    mode_given = "mode" in kv or "cig_mode" in kv
    mode = str(kv.get("mode", kv.get("cig_mode", "offset"))).lower()
    if mode not in ("offset", "angle"):
        raise ValueError(f"Unknown CIG mode: {mode}")
    p.cig_is_angle = int(mode == "angle")
    p.adj = _i(kv.get("adj", 1))
    p.cig = _i(kv["cig"]) if "cig" in kv else int(mode_given)
    p.cmp = _i(kv.get("cmp", 1))
    p.half_angle = _i(kv.get("half_angle", 0))
    p.oazmin, p.oazmax = (40.0, 25.0) if p.half_angle else (80.0, 50.0)
    if "oazmin" in kv:
        p.oazmin = _f(kv["oazmin"])
    if "oazmax" in kv:
        p.oazmax = _f(kv["oazmax"])
    p.oaztaper = _f(kv.get("oaztaper", 10.0))
    p.isdiff = _i(kv.get("diff", 0))
    p.aperture_trace = _i(kv.get("aperture_trace", -1))
    p.aperture_trace_taper = max(0, _i(kv.get("aperture_trace_taper", 20)))
    p.direct_mute = _i(kv.get("direct_mute", 0))
    p.direct_mute_slope = _f(kv.get("direct_mute_slope", 0.0))
    p.direct_mute_invert = _i(kv.get("direct_mute_invert", 0))
    p.direct_mute_intercept = _f(kv.get("direct_mute_intercept", 0.0))
    p.direct_mute_taper = _f(kv.get("direct_mute_taper", 0.05))
    p.direct_mute_qc_shot = _i(kv.get("direct_mute_qc_shot", -1))
    p.tau = _f(kv.get("tau", 0.0))
    p.aperture = _f(kv.get("aperture", 90.0))
    p.aperture_taper = _f(kv.get("aperture_taper", 10.0))
    p.aperture_rmin = _f(kv.get("aperture_rmin", 0.0))
    p.aperture_rtaper = _f(kv.get("aperture_rtaper", 0.05))
    p.antialias = _f(kv.get("antialias", 1.0))
    p.angle_interp = _i(kv.get("angle_interp", 1))
    p.angle_taper_start = _f(kv.get("angle_taper_start", -1.0))
    p.angle_taper_end = _f(kv.get("angle_taper_end", -1.0))
    p.angle_grad_radius = max(1, _i(kv.get("angle_grad_radius", 2)))
    p.angle_smooth = max(0, _i(kv.get("angle_smooth", 2)))
    t = str(kv.get("type", "hermit"))
    if t[:1] not in ("l", "p", "h"):
        raise ValueError(f"Unknown interpolation type: {t}")
    p.interp_type = ord(t[0])
    if p.cig:
        dh_eff = 0.0 if nh == 1 else dh
        p.g0, p.dg, p.ng = output_axis(kv, bool(p.cig_is_angle), nh, h0, dh_eff,
                                       bool(p.half_angle))
    else:
        p.g0, p.dg, p.ng = 0.0, 1.0, 1
    return p


def geom(data_axes, s_axes, r_axes) -> KmoGeom:
    """Axes are [[n,o,d], ...] in file order: data (t,h,s), tables (z,x,pos)."""
    g = KmoGeom()
    (g.nt, g.t0, g.dt), (g.nh, g.h0, g.dh), (g.ns, g.s0, g.ds) = [
        (int(a[0]), _f(a[1]), _f(a[2])) for a in data_axes[:3]]
    (g.nz, g.z0, g.dz), (g.nx, g.x0, g.dx), (g.sny, g.sy0, g.sdy) = [
        (int(a[0]), _f(a[1]), _f(a[2])) for a in s_axes[:3]]
    g.rny, g.ry0, g.rdy = int(r_axes[2][0]), _f(r_axes[2][1]), _f(r_axes[2][2])
    return g


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def migrate(kv, data, data_axes, stab, sder, s_axes, rtab, rder, r_axes, grads=None,
            shots=None, precision="f32", threads=None):
    """Run the oracle; returns (cig (ng,nx,nz), qc_before, qc_after)."""
    L = lib()
    g = geom(data_axes, s_axes, r_axes)
    p = params_from_kv(kv, g.nh, g.h0, g.dh)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (data, stab, sder, rtab, rder)]
    gr = [np.ascontiguousarray(a, dtype=np.float32) for a in grads] if grads else [None] * 4
    ng = p.ng if p.cig else 1
    dtype = np.float32 if precision == "f32" else np.float64
    out = np.zeros((ng, g.nx, g.nz), dtype=dtype)
    qcb = np.zeros((g.nh, g.nt), np.float32)
    qca = np.zeros((g.nh, g.nt), np.float32)
    sb, se = (0, g.ns) if shots is None else shots
    if threads is not None:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    fn = L.kmo_migrate_f32 if precision == "f32" else L.kmo_migrate_f64
    outp = out.ctypes.data_as(C.POINTER(C.c_float if precision == "f32" else C.c_double))
    rc = fn(C.byref(p), C.byref(g), *[_fp(a) for a in arrs], *[_fp(a) for a in gr], sb, se,
            outp, _fp(qcb), _fp(qca))
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return out, qcb, qca


def model(kv, image, data_axes, stab, sder, s_axes, rtab, rder, r_axes, grads=None,
          shots=None, precision="f32"):
    """adj=0 modelling of the image (ng, nx, nz) into data (ns, nh, nt), the
    reference's run on one thread (its point order)."""
    L = lib()
    g = geom(data_axes, s_axes, r_axes)
    p = params_from_kv(dict(kv, adj=0), g.nh, g.h0, g.dh)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (image, stab, sder, rtab, rder)]
    gr = [np.ascontiguousarray(a, dtype=np.float32) for a in grads] if grads else [None] * 4
    dtype = np.float32 if precision == "f32" else np.float64
    out = np.zeros((g.ns, g.nh, g.nt), dtype=dtype)
    sb, se = (0, g.ns) if shots is None else shots
    fn = L.kmo_model_f32 if precision == "f32" else L.kmo_model_f64
    outp = out.ctypes.data_as(C.POINTER(C.c_float if precision == "f32" else C.c_double))
    rc = fn(C.byref(p), C.byref(g), *[_fp(a) for a in arrs], *[_fp(a) for a in gr], sb, se,
            outp)
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return out


def rel_l2(a, b) -> float:
    """rel_l2 = sqrt(sum((a-b)^2) / sum(b^2)) in double, src/rsf_l2cmp.cpp:9-10."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float(np.sum(b * b))
    num = float(np.sum((a - b) ** 2))
    return math.sqrt(num / den) if den > 0 else math.sqrt(num)

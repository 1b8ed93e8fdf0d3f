"""Python side of the kirchdat2d_freq oracle (kmdat_oracle.c).  TEST
INFRASTRUCTURE ONLY: tests/ and bench.py's CPU baseline use it to check the
datuming library, never the product.

Arrays follow the reference's RSF layouts: data (ns, nh, nt) with time
fastest; datum tables (n, n) as written by greensol_auto, row = the
summation position's table index, column = the output position's; the
traveltime cube (npos, nx, nz).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import oracle as _o

ERRORS = {-1: "bad arguments", -2: "Receiver table too small.", -3: "Source table too small.",
          -4: "cmp=1 regrouping reaches traces outside the cube"}


class KmdoCfg(C.Structure):
    _fields_ = [("cmp", C.c_int), ("aper", C.c_int), ("tap", C.c_int), ("data_tap", C.c_int),
                ("antialias", C.c_float), ("length", C.c_float), ("sdatum", C.c_float),
                ("rdatum", C.c_float), ("use_bf", C.c_int), ("bf_p", C.c_int),
                ("bf_leaf", C.c_int), ("bf_levels", C.c_int), ("bf_amp_eps", C.c_float),
                ("bf_phase_tol", C.c_float)]


_BOUND = False


def lib():
    global _BOUND
    L = _o.lib()
    if not _BOUND:
        fp = C.POINTER(C.c_float)
        i, f = C.c_int, C.c_float
        for nm in ("kmdo_datum_f32", "kmdo_datum_f64"):
            fn = getattr(L, nm)
            fn.restype = i
            fn.argtypes = [C.POINTER(KmdoCfg), fp, i, i, i, f, f, f, f, f, fp, i, f, f, fp, i,
                           f, f, fp, fp, i, i]
        L.kmdo_nfft.restype = i
        L.kmdo_nfft.argtypes = [i, i, f, i, i, fp, fp]
        L.kmdo_filter_f32.argtypes = [f, f, i, fp]
        L.kmdo_H_f32.argtypes = [f, f, i, i, i, fp, fp]
        L.kmdo_H_f64.argtypes = [C.c_double, C.c_double, i, i, i, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.kmdo_taper_weight.restype = f
        L.kmdo_taper_weight.argtypes = [i, i, i, i, i]
        L.kmdo_edge_taper.restype = f
        L.kmdo_edge_taper.argtypes = [i, i, i]
        L.kmdo_antialias_weight.restype = f
        L.kmdo_antialias_weight.argtypes = [f, i, f, i, f]
        L.kmdo_fft_f64.argtypes = [i, C.POINTER(C.c_double), C.POINTER(C.c_double), i]
        L.kmdo_greensol.restype = i
        L.kmdo_greensol.argtypes = [fp, i, i, i, f, f, f, f, f, f, f, fp]
        _BOUND = True
    return L


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def config(kv, sdatum, rdatum=None):
    """The reference's keys and defaults (src/kirchdat2d_frequency_cs.cpp:1216-1270),
    use_bf=1 and the bf_* keys included (:1244-1249)."""
    c = KmdoCfg()
    c.cmp = int(kv.get("cmp", 1))
    c.aper = int(kv.get("aperture", 50))
    c.tap = int(kv.get("taper", 10))
    c.data_tap = int(kv.get("data_taper", 0))
    c.antialias = float(kv.get("antialias", 1.0))
    c.length = float(kv.get("length", 0.025))
    c.sdatum = float(sdatum)
    c.rdatum = float(sdatum if rdatum is None else rdatum)
    c.use_bf = int(kv.get("use_bf", 1))
    c.bf_p = int(kv.get("bf_p", 12))
    c.bf_leaf = int(kv.get("bf_n_leaf", 16))
    c.bf_levels = int(kv.get("bf_panel_levels", 1))
    c.bf_amp_eps = float(kv.get("bf_amp_eps", 1e-20))
    c.bf_phase_tol = float(kv.get("bf_phase_tol", 2.0))
    return c


def datum(kv, data, data_axes, sgreen, sg_axis, rgreen, rg_axis, sdatum, rdatum=None,
          precision="f32", interm=False, iw_range=None):
    """Datuming: returns out (ns, nh, nt) (and the receiver-side cube when
    interm).  kv's use_bf (the reference's default 1) selects the butterfly
    operators, which the oracle restates in float only (precision="f32");
    use_bf=0 is the direct path in either precision.  data_axes [[nt, t0, dt],
    [nh, h0, dh], [ns, s0, ds]];
    sg_axis / rg_axis = (o, d) of the table positions."""
    L = lib()
    d = np.ascontiguousarray(data, np.float32)
    ns, nh, nt = d.shape
    (_, _, dt), (_, h0, dh), (_, s0, ds) = data_axes
    sg = np.ascontiguousarray(sgreen, np.float32)
    rg = np.ascontiguousarray(rgreen, np.float32)
    out = np.empty_like(d)
    mid = np.empty_like(d) if interm else None
    cfg = config(kv, sdatum, rdatum)
    b, e = iw_range if iw_range else (0, 0)
    fn = L.kmdo_datum_f32 if precision == "f32" else L.kmdo_datum_f64
    rc = fn(C.byref(cfg), _fp(d), nt, nh, ns, dt, h0, dh, s0, ds, _fp(sg), sg.shape[0],
            sg_axis[0], sg_axis[1], _fp(rg), rg.shape[0], rg_axis[0], rg_axis[1], _fp(out),
            _fp(mid), b, e)
    if rc:
        raise ValueError(f"kmdo_datum: {ERRORS.get(rc, rc)}")
    return (out, mid) if interm else out


def greensol(cube, cube_axes, pos_axis, datum_depth):
    """greensol_auto: cube (npos, nx, nz) with cube_axes [[nz, oz, dz],
    [nx, ox, dx]]; pos_axis (ogx, dgx); returns the (npos, npos) table."""
    L = lib()
    c = np.ascontiguousarray(cube, np.float32)
    npos, nx, nz = c.shape
    out = np.empty((npos, npos), np.float32)
    rc = L.kmdo_greensol(_fp(c), nz, nx, npos, cube_axes[0][1], cube_axes[0][2],
                         cube_axes[1][1], cube_axes[1][2], pos_axis[0], pos_axis[1],
                         float(datum_depth), _fp(out))
    if rc:
        raise ValueError({-1: "Datum out of range.", -2: "Receiver table too small."}[rc])
    return out


def nfft(nt, nsam, dt, rgreen, sgreen):
    L = lib()
    r = np.ascontiguousarray(rgreen, np.float32)
    s = np.ascontiguousarray(sgreen, np.float32)
    return L.kmdo_nfft(nt, nsam, dt, r.shape[0], s.shape[0], _fp(r), _fp(s))


def H(tau, dt, nsam, nfft_, iw, precision="f32"):
    L = lib()
    if precision == "f32":
        hr, hi = C.c_float(), C.c_float()
        L.kmdo_H_f32(tau, dt, nsam, nfft_, iw, C.byref(hr), C.byref(hi))
    else:
        hr, hi = C.c_double(), C.c_double()
        L.kmdo_H_f64(tau, dt, nsam, nfft_, iw, C.byref(hr), C.byref(hi))
    return complex(hr.value, hi.value)


def filt(tau, dt, nsam):
    out = np.zeros(max(nsam - 1, 1), np.float32)
    lib().kmdo_filter_f32(tau, dt, nsam, _fp(out))
    return out


def fft(x, inverse=False):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    dp = C.POINTER(C.c_double)
    lib().kmdo_fft_f64(len(x), re.ctypes.data_as(dp), im.ctypes.data_as(dp), int(inverse))
    return re + 1j * im

"""ctypes binding of the GPU datuming library (include/kmig_dat.h).

``datum`` runs the frequency-domain Kirchhoff datuming of kirchdat2d_freq on
the GPU: data (ns, nh, nt) with time fastest, datum tables (n, n) as
greensol_auto writes them, keys as the reference program takes them
(cmp, aperture, taper, data_taper, antialias, length, use_bf, bf_*).
``Datum`` is the resident context (tables and data stay on the GPU across
runs).  ``greensol`` cuts the datum table out of a traveltime cube.  The
binding loads kirch_datuming_amd/lib/libkmdat.so and raises if it is
missing; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkmdat.so")
EXPORTS = ("kmdat_config_default", "kmdat_config_check", "kmdat_create", "kmdat_destroy",
           "kmdat_last_error", "kmdat_set_green", "kmdat_set_data", "kmdat_run", "kmdat_sync",
           "kmdat_get_output", "kmdat_get_interm", "kmdat_get_stats", "kmdat_datum",
           "kmdat_greensol")


class DatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kmdat error {code}: {msg}")
        self.code = code
        self.msg = msg


class DatConfig(C.Structure):
    _fields_ = [("cmp", C.c_int), ("aperture", C.c_int), ("taper", C.c_int),
                ("data_taper", C.c_int), ("antialias", C.c_float), ("length", C.c_float),
                ("sdatum", C.c_float), ("rdatum", C.c_float), ("use_bf", C.c_int),
                ("bf_p", C.c_int), ("bf_n_leaf", C.c_int), ("bf_panel_levels", C.c_int),
                ("bf_amp_eps", C.c_float), ("bf_phase_tol", C.c_float), ("interm", C.c_int)]


class DatStats(C.Structure):
    _fields_ = [("nfft", C.c_int), ("nw", C.c_int), ("nsam", C.c_int),
                ("keys_receiver", C.c_longlong), ("keys_source", C.c_longlong),
                ("pairs_receiver", C.c_longlong), ("pairs_source", C.c_longlong),
                ("cmacs", C.c_double), ("ms_forward", C.c_float), ("ms_tables", C.c_float),
                ("ms_receiver", C.c_float), ("ms_source", C.c_float),
                ("ms_inverse", C.c_float), ("ms_total", C.c_float), ("banded", C.c_int),
                ("butterfly", C.c_int), ("bf_panels_receiver", C.c_longlong),
                ("bf_panels_source", C.c_longlong), ("bf_items_receiver", C.c_longlong),
                ("bf_items_source", C.c_longlong), ("ms_bf_receiver", C.c_float),
                ("ms_bf_source", C.c_float)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        ctx = C.c_void_p
        i, f = C.c_int, C.c_float
        L.kmdat_config_default.argtypes = [C.POINTER(DatConfig)]
        L.kmdat_config_default.restype = None
        L.kmdat_config_check.argtypes = [C.POINTER(DatConfig), C.c_char_p, C.c_size_t]
        L.kmdat_create.argtypes = [C.POINTER(ctx), C.POINTER(DatConfig)]
        L.kmdat_destroy.argtypes = [ctx]
        L.kmdat_destroy.restype = None
        L.kmdat_last_error.argtypes = [ctx]
        L.kmdat_last_error.restype = C.c_char_p
        L.kmdat_set_green.argtypes = [ctx, i, fp, i, f, f]
        L.kmdat_set_data.argtypes = [ctx, fp, i, f, i, f, f, i, f, f]
        L.kmdat_run.argtypes = [ctx]
        L.kmdat_sync.argtypes = [ctx]
        L.kmdat_get_output.argtypes = [ctx, fp]
        L.kmdat_get_interm.argtypes = [ctx, fp]
        L.kmdat_get_stats.argtypes = [ctx, C.POINTER(DatStats)]
        L.kmdat_datum.argtypes = [ctx, fp, i, f, i, f, f, i, f, f, fp, fp]
        L.kmdat_greensol.argtypes = [fp, i, i, i, f, f, f, f, f, f, f, fp]
        for nm in EXPORTS:
            if nm not in ("kmdat_config_default", "kmdat_destroy", "kmdat_last_error"):
                getattr(L, nm).restype = C.c_int
        _LIB = L
    return _LIB


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def config(kv, sdatum, rdatum=None, interm=False) -> DatConfig:
    """kirchdat2d_freq's keys over its defaults (src/kirchdat2d_frequency_cs.cpp:1216-1270)."""
    c = DatConfig()
    lib().kmdat_config_default(C.byref(c))
    for k, conv in (("cmp", int), ("aperture", int), ("taper", int), ("data_taper", int),
                    ("antialias", float), ("length", float), ("use_bf", int), ("bf_p", int),
                    ("bf_n_leaf", int), ("bf_panel_levels", int), ("bf_amp_eps", float),
                    ("bf_phase_tol", float)):
        if k in kv:
            setattr(c, k, conv(kv[k]))
    c.sdatum = float(sdatum)
    c.rdatum = float(sdatum if rdatum is None else rdatum)
    c.interm = 1 if interm else 0
    return c


def check(cfg: DatConfig):
    buf = C.create_string_buffer(512)
    rc = lib().kmdat_config_check(C.byref(cfg), buf, 512)
    if rc:
        raise DatError(rc, buf.value.decode())


class Datum:
    """Resident datuming context."""

    def __init__(self, cfg: DatConfig):
        self.L = lib()
        self.ctx = C.c_void_p()
        self.cfg = cfg
        rc = self.L.kmdat_create(C.byref(self.ctx), C.byref(cfg))
        if rc:
            msg = self.L.kmdat_last_error(self.ctx).decode() if self.ctx else "create failed"
            self.close()
            raise DatError(rc, msg)

    def _chk(self, rc):
        if rc:
            raise DatError(rc, self.L.kmdat_last_error(self.ctx).decode())

    def set_green(self, side, table, o, d):
        t = np.ascontiguousarray(table, np.float32)
        assert t.ndim == 2 and t.shape[0] == t.shape[1]
        self._chk(self.L.kmdat_set_green(self.ctx, int(side), _fp(t), t.shape[0], float(o),
                                         float(d)))

    def set_data(self, data, axes):
        d = np.ascontiguousarray(data, np.float32)
        (nt, _, dt), (nh, h0, dh), (ns, s0, ds) = axes
        assert d.shape == (int(ns), int(nh), int(nt))
        self.shape = d.shape
        self._chk(self.L.kmdat_set_data(self.ctx, _fp(d), int(nt), dt, int(nh), h0, dh, int(ns),
                                        s0, ds))

    def run(self):
        self._chk(self.L.kmdat_run(self.ctx))

    def sync(self):
        self._chk(self.L.kmdat_sync(self.ctx))

    def output(self):
        out = np.empty(self.shape, np.float32)
        self._chk(self.L.kmdat_get_output(self.ctx, _fp(out)))
        return out

    def interm(self):
        out = np.empty(self.shape, np.float32)
        self._chk(self.L.kmdat_get_interm(self.ctx, _fp(out)))
        return out

    def stats(self) -> dict:
        st = DatStats()
        self._chk(self.L.kmdat_get_stats(self.ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in DatStats._fields_}

    def close(self):
        if getattr(self, "ctx", None):
            self.L.kmdat_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def datum(kv, data, data_axes, sgreen, sg_axis, rgreen, rg_axis, sdatum, rdatum=None,
          interm=False, stats=False):
    """One datuming run on the GPU: returns the (ns, nh, nt) output (and the
    receiver-side cube when interm, the run's statistics when stats);
    sg_axis / rg_axis = (o, d) of the table positions."""
    cfg = config(kv, sdatum, rdatum, interm)
    check(cfg)
    with Datum(cfg) as dm:
        dm.set_green(0, sgreen, *sg_axis)
        dm.set_green(1, rgreen, *rg_axis)
        dm.set_data(data, data_axes)
        dm.run()
        dm.sync()
        res = [dm.output()]
        if interm:
            res.append(dm.interm())
        if stats:
            res.append(dm.stats())
        return res[0] if len(res) == 1 else tuple(res)


def greensol(cube, cube_axes, pos_axis, datum_depth):
    """greensol_auto on the host: cube (npos, nx, nz), cube_axes [[nz, oz, dz],
    [nx, ox, dx]], pos_axis (ogx, dgx); returns the (npos, npos) table."""
    c = np.ascontiguousarray(cube, np.float32)
    npos, nx, nz = c.shape
    out = np.empty((npos, npos), np.float32)
    rc = lib().kmdat_greensol(_fp(c), nz, nx, npos, cube_axes[0][1], cube_axes[0][2],
                              cube_axes[1][1], cube_axes[1][2], pos_axis[0], pos_axis[1],
                              float(datum_depth), _fp(out))
    if rc:
        raise DatError(rc, {-1: "Datum out of range.", -2: "Receiver table too small."}[rc])
    return out

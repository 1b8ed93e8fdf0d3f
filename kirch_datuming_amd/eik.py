"""ctypes binding of the fast-marching table source (include/kmig_eik.h).

``eikods2d`` computes, for every source position, the first-arrival
traveltime, its derivatives along the relative and source coordinates, and
the upwind spatial gradients, as the reference's eikods2d program writes them
(time=, tdl1=, tds1=, gradz=, gradx=), from a 2-D velocity model.  Host code
(OpenMP over positions); no GPU is needed.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkmeik.so")
EXPORTS = ("kmig_eik_config_default", "kmig_eikods2d")


class EikConfig(C.Structure):
    _fields_ = [("order", C.c_int), ("b1", C.c_int), ("b2", C.c_int), ("l", C.c_int),
                ("nthreads", C.c_int)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        L.kmig_eik_config_default.argtypes = [C.POINTER(EikConfig)]
        L.kmig_eik_config_default.restype = None
        L.kmig_eikods2d.argtypes = [C.POINTER(EikConfig), fp, C.c_int, C.c_int, C.c_float,
                                    C.c_float, C.c_float, C.c_float, fp, C.c_int, fp, fp, fp,
                                    fp, fp]
        L.kmig_eikods2d.restype = C.c_int
        _LIB = L
    return _LIB


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def eikods2d(vel, d1, d2, shots, o1=0.0, o2=0.0, order=2, b1=1, b2=1, l=1, threads=0,
             grads=True, tdl1=True):
    """vel: (n2, n1) float32 velocity, n1 = depth fastest; shots: (nshot, 3)
    (depth, lateral, cross-line) coordinates.  Returns a dict of
    (nshot, n2, n1) float32 arrays: time, tds1 (and tdl1, gradz, gradx), and
    'status' (0, or 1 when a front ran empty before every point was reached,
    which the reference only warns about)."""
    v = np.ascontiguousarray(vel, dtype=np.float32)
    n2, n1 = v.shape
    s = np.ascontiguousarray(shots, dtype=np.float32).reshape(-1, 3)
    ns = s.shape[0]
    out = {k: np.empty((ns, n2, n1), np.float32)
           for k in ("time", "tds1") + (("tdl1",) if tdl1 else ())}
    if grads:
        out["gradz"] = np.empty((ns, n2, n1), np.float32)
        out["gradx"] = np.empty((ns, n2, n1), np.float32)
    cfg = EikConfig()
    L = lib()
    L.kmig_eik_config_default(C.byref(cfg))
    cfg.order, cfg.b1, cfg.b2, cfg.l, cfg.nthreads = order, b1, b2, l, threads
    rc = L.kmig_eikods2d(C.byref(cfg), _fp(v), n1, n2, float(o1), float(o2), float(d1),
                         float(d2), _fp(s), ns, _fp(out["time"]), _fp(out.get("tdl1")),
                         _fp(out["tds1"]), _fp(out.get("gradz")), _fp(out.get("gradx")))
    if rc < 0:
        raise ValueError(f"kmig_eikods2d: invalid parameters (status {rc})")
    out["status"] = rc
    return out

"""Python side of the segmented-butterfly oracle (kbf_oracle.c) and of the
reference's own butterfly compiled into oracle/_ref/libbf1d_ref.so.  TEST
INFRASTRUCTURE ONLY: tests/ use it to check the datuming library, never the
product.

Matrices are row-major n x n: tau (float32) and amp (complex64), row = the
output index, column = the input index; the operator is
K = amp * exp(-1j * omega * tau).  Vectors are complex64 (nvec, n).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from oracle import oracle as _o

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(_HERE, "_ref", "libbf1d_ref.so")

# bf_p, bf_n_leaf, bf_panel_levels, bf_amp_eps, bf_phase_tol defaults
# (src/kirchdat2d_frequency_cs.cpp:1245-1249)
DEFAULTS = dict(p=12, leaf=16, levels=1, amp_eps=1e-20, phase_tol=2.0)

EXACT, BUTTERFLY = 0, 1


class KboCfg(C.Structure):
    _fields_ = [("p", C.c_int), ("leaf", C.c_int), ("levels", C.c_int), ("amp_eps", C.c_float),
                ("phase_tol", C.c_float)]


_BOUND = False
_REF = None


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def lib():
    global _BOUND
    L = _o.lib()
    if not _BOUND:
        fp, ip, i, f = C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_int, C.c_float
        L.kbo_seg_apply.restype = i
        L.kbo_seg_apply.argtypes = [i, fp, fp, f, C.POINTER(KboCfg), i, fp, fp, ip, i]
        L.kbo_strict_apply.restype = i
        L.kbo_strict_apply.argtypes = [i, fp, fp, f, i, i, i, fp, fp]
        L.kbo_lagrange.argtypes = [i, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
        _BOUND = True
    return L


def ref_lib():
    """The reference's bf1d.cpp compiled from source (oracle/ref.mk), or None
    when it has not been built."""
    global _REF
    if _REF is None and os.path.exists(REF_LIB):
        R = C.CDLL(REF_LIB)
        fp, ip, i, f = C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_int, C.c_float
        R.refbf_seg_apply.restype = i
        R.refbf_seg_apply.argtypes = [i, fp, fp, f, i, i, i, f, f, i, fp, fp, ip, i]
        R.refbf_strict_apply.restype = None
        R.refbf_strict_apply.argtypes = [i, fp, fp, f, i, i, i, fp, fp]
        _REF = R
    return _REF


def _prep(tau, amp, u):
    tau = np.ascontiguousarray(tau, np.float32)
    amp = np.ascontiguousarray(amp, np.complex64)
    u = np.ascontiguousarray(np.atleast_2d(u), np.complex64)
    n = tau.shape[0]
    assert tau.shape == (n, n) and amp.shape == (n, n) and u.shape[1] == n
    return tau, amp, u, n


def seg_apply(tau, amp, omega, u, p=12, leaf=16, levels=1, amp_eps=1e-20, phase_tol=2.0,
              ref=False, max_panels=1 << 16):
    """Segmented operator applied to the rows of u: returns (out (nvec, n),
    panels (npanel, 5): row0, col0, rows, cols, kind).  ref=True runs the
    compiled reference instead of the oracle."""
    tau, amp, u, n = _prep(tau, amp, u)
    out = np.zeros_like(u)
    pan = np.zeros((max_panels, 5), np.int32)
    if ref:
        R = ref_lib()
        if R is None:
            raise RuntimeError("oracle/_ref/libbf1d_ref.so is not built")
        k = R.refbf_seg_apply(n, _fp(tau), _fp(amp), float(omega), p, leaf, levels,
                              float(amp_eps), float(phase_tol), u.shape[0], _fp(u), _fp(out),
                              _ip(pan), max_panels)
    else:
        cfg = KboCfg(p, leaf, levels, amp_eps, phase_tol)
        k = lib().kbo_seg_apply(n, _fp(tau), _fp(amp), float(omega), C.byref(cfg), u.shape[0],
                                _fp(u), _fp(out), _ip(pan), max_panels)
        if k < 0:
            raise ValueError("kbo_seg_apply: invalid butterfly parameters")
    return out, pan[:min(k, max_panels)].copy()


def strict_apply(tau, amp, omega, u, p=12, leaf=16, ref=False):
    """One strict factor (n = leaf 2^L) applied to the rows of u."""
    tau, amp, u, n = _prep(tau, amp, u)
    out = np.zeros_like(u)
    if ref:
        R = ref_lib()
        if R is None:
            raise RuntimeError("oracle/_ref/libbf1d_ref.so is not built")
        R.refbf_strict_apply(n, _fp(tau), _fp(amp), float(omega), p, leaf, u.shape[0], _fp(u),
                             _fp(out))
    else:
        if lib().kbo_strict_apply(n, _fp(tau), _fp(amp), float(omega), p, leaf, u.shape[0],
                                  _fp(u), _fp(out)):
            raise ValueError("kbo_strict_apply: n must be leaf * 2^L and p < leaf")
    return out


def lagrange(nodes, x):
    nodes = np.ascontiguousarray(nodes, np.float64)
    out = np.zeros_like(nodes)
    dp = C.POINTER(C.c_double)
    lib().kbo_lagrange(len(nodes), nodes.ctypes.data_as(dp), float(x), out.ctypes.data_as(dp))
    return out


def exact_apply(tau, amp, omega, u):
    """K u in double, K = amp exp(-1j omega tau) with omega rounded to float."""
    tau, amp, u, n = _prep(tau, amp, u)
    K = amp.astype(np.complex128) * np.exp(-1j * np.float64(np.float32(omega)) *
                                           tau.astype(np.float64))
    return u.astype(np.complex128) @ K.T


def selftest_case(n=64):
    """The geometry of bf1d.cpp's own self-test (SERECKIRCH/src/bf1d.cpp:
    1073-1085): tau = 0.5 + 0.01 sqrt(1 + (r - c)^2), amp = a (1 + 0.1 i) with
    a = exp(-0.001 (r - c)^2), u = sin(0.1 r) + i cos(0.2 r), omega = 20."""
    r = np.arange(n, dtype=np.float32)
    d2 = (r[:, None] - r[None, :]) ** 2
    tau = (np.float32(0.5) + np.float32(0.01) * np.sqrt(np.float32(1.0) + d2)).astype(np.float32)
    a = np.exp(np.float32(-0.001) * d2).astype(np.float32)
    amp = (a + 1j * (np.float32(0.1) * a)).astype(np.complex64)
    u = (np.sin(np.float32(0.1) * r) + 1j * np.cos(np.float32(0.2) * r)).astype(np.complex64)
    return tau, amp, np.float32(20.0), u[None, :]

"""ctypes binding of the learned traveltime surrogate (include/kmig_nn.h).

``Surrogate`` owns one network on the GPU: BF16-MFMA training on
fast-marching table records, evaluation at arbitrary points, and the
parameter vector (FP32, the layout of include/kmig_nn.h).  There is no CPU
path: a missing library or GPU raises.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import kmig

PARAMS = 56324
EXPORTS = ("kmig_nn_config_default", "kmig_nn_create", "kmig_nn_set_data", "kmig_nn_train",
           "kmig_nn_eval", "kmig_nn_grad", "kmig_nn_get_params", "kmig_nn_set_params",
           "kmig_nn_stats", "kmig_nn_last_error", "kmig_nn_destroy")


class NnConfig(C.Structure):
    _fields_ = [("s_min", C.c_float), ("s_max", C.c_float), ("z_min", C.c_float),
                ("z_max", C.c_float), ("x_min", C.c_float), ("x_max", C.c_float),
                ("z_src", C.c_float), ("slowness", C.c_float), ("sigma_t", C.c_float),
                ("sigma_g", C.c_float), ("lr", C.c_float), ("lr_final", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("batch", C.c_int), ("seed", C.c_ulonglong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_BOUND = False


def lib():
    global _BOUND
    L = kmig.lib()
    if not _BOUND:
        fp, vp = C.POINTER(C.c_float), C.c_void_p
        L.kmig_nn_config_default.argtypes = [C.POINTER(NnConfig)]
        L.kmig_nn_create.argtypes = [C.POINTER(vp), C.POINTER(NnConfig)]
        L.kmig_nn_set_data.argtypes = [vp, fp, C.c_long]
        L.kmig_nn_train.argtypes = [vp, C.c_int, fp]
        L.kmig_nn_eval.argtypes = [vp, fp, C.c_long, fp]
        L.kmig_nn_grad.argtypes = [vp, fp, C.c_long, fp, fp]
        L.kmig_nn_get_params.argtypes = [vp, fp]
        L.kmig_nn_set_params.argtypes = [vp, fp]
        L.kmig_nn_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.kmig_nn_last_error.argtypes = [vp]
        L.kmig_nn_last_error.restype = C.c_char_p
        L.kmig_nn_destroy.argtypes = [vp]
        L.kmig_nn_destroy.restype = None
        _BOUND = True
    return L


def default_config(**kw) -> NnConfig:
    c = NnConfig()
    lib().kmig_nn_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Surrogate:
    def __init__(self, cfg: NnConfig | None = None, **kw):
        self.L = lib()
        self.cfg = cfg if cfg is not None else default_config(**kw)
        self.h = C.c_void_p()
        rc = self.L.kmig_nn_create(C.byref(self.h), C.byref(self.cfg))
        self._check(rc)

    def _check(self, rc):
        if rc != 0:
            msg = self.L.kmig_nn_last_error(self.h).decode() if self.h else ""
            raise kmig.KmigError(rc, msg)

    def set_data(self, records):
        r = np.ascontiguousarray(records, dtype=np.float32)
        assert r.ndim == 2 and r.shape[1] == 7
        self._check(self.L.kmig_nn_set_data(self.h, _fp(r), r.shape[0]))

    def train(self, steps):
        loss = np.zeros(max(steps, 1), np.float32)
        self._check(self.L.kmig_nn_train(self.h, int(steps), _fp(loss)))
        return loss[:steps]

    def eval(self, points):
        p = np.ascontiguousarray(points, dtype=np.float32)
        out = np.zeros((p.shape[0], 4), np.float32)
        self._check(self.L.kmig_nn_eval(self.h, _fp(p), p.shape[0], _fp(out)))
        return out

    def grad(self, records):
        r = np.ascontiguousarray(records, dtype=np.float32)
        g = np.zeros(PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.L.kmig_nn_grad(self.h, _fp(r), r.shape[0], _fp(g), _fp(loss)))
        return float(loss[0]), g

    @property
    def params(self):
        p = np.zeros(PARAMS, np.float32)
        self._check(self.L.kmig_nn_get_params(self.h, _fp(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == PARAMS
        self._check(self.L.kmig_nn_set_params(self.h, _fp(p)))

    def stats(self):
        ms, fl = C.c_double(), C.c_double()
        self._check(self.L.kmig_nn_stats(self.h, C.byref(ms), C.byref(fl)))
        return {"train_ms": ms.value, "train_flops": fl.value}

    def close(self):
        if self.h:
            self.L.kmig_nn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes binding of the surrogate MLP (include/sbz.h, sbz_mlp_*; kernels in csrc/sbz_mlp.hip)."""
import ctypes

import numpy as np

from . import _lib

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _declare(L):
    vp = ctypes.c_void_p
    L.sbz_mlp_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.sbz_mlp_destroy.argtypes = [vp]
    L.sbz_mlp_n_params.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.sbz_mlp_get_params.argtypes = [vp, _fp]
    L.sbz_mlp_set_params.argtypes = [vp, _fp]
    L.sbz_mlp_grad.argtypes = [vp, _fp, _fp, ctypes.c_int, _fp, _fp]
    L.sbz_mlp_set_data.argtypes = [vp, _fp, _fp, ctypes.c_longlong]
    L.sbz_mlp_train.argtypes = [vp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp, _fp]
    L.sbz_mlp_forward.argtypes = [vp, _fp, ctypes.c_longlong, _fp, _fp]


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """widths [w0, ..., wL]: GELU hidden layers, linear output, BF16 MFMA GEMMs, FP32 Adam."""

    def __init__(self, widths, max_batch, seed=0):
        self.L = _lib.lib()
        _declare(self.L)
        self.widths = [int(w) for w in widths]
        w = np.array(self.widths, dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.check(self.L.sbz_mlp_create(len(w) - 1, w.ctypes.data_as(_ip), int(max_batch), int(seed),
                                         ctypes.byref(h)), "sbz_mlp_create")
        self.h = h
        self.max_batch = int(max_batch)
        n = ctypes.c_longlong(0)
        _lib.check(self.L.sbz_mlp_n_params(h, ctypes.byref(n)), "sbz_mlp_n_params")
        self.n_params = n.value

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _lib.check(self.L.sbz_mlp_get_params(self.h, p.ctypes.data_as(_fp)), "sbz_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f(p)
        _lib.check(self.L.sbz_mlp_set_params(self.h, p.ctypes.data_as(_fp)), "sbz_mlp_set_params")

    def grad(self, x, t):
        x, t = _f(x), _f(t)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _lib.check(self.L.sbz_mlp_grad(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(x),
                                       g.ctypes.data_as(_fp), loss.ctypes.data_as(_fp)), "sbz_mlp_grad")
        return g, float(loss[0])

    def set_data(self, x, t):
        self._x, self._t = _f(x), _f(t)
        _lib.check(self.L.sbz_mlp_set_data(self.h, self._x.ctypes.data_as(_fp), self._t.ctypes.data_as(_fp),
                                           len(self._x)), "sbz_mlp_set_data")
        self.n_data = len(self._x)

    def train(self, idx, batch, lr):
        """Adam steps over consecutive batches of the row list idx; returns (losses, kernel ms)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        nb = len(idx) // batch
        loss = np.zeros(nb, dtype=np.float32)
        ms = ctypes.c_float(0.0)
        _lib.check(self.L.sbz_mlp_train(self.h, idx.ctypes.data_as(_ip), nb, int(batch), float(lr),
                                        loss.ctypes.data_as(_fp), ctypes.byref(ms)), "sbz_mlp_train")
        return loss, ms.value

    def forward(self, x):
        x = _f(x)
        y = np.zeros((len(x), self.widths[-1]), dtype=np.float32)
        ms = ctypes.c_float(0.0)
        _lib.check(self.L.sbz_mlp_forward(self.h, x.ctypes.data_as(_fp), len(x), y.ctypes.data_as(_fp),
                                          ctypes.byref(ms)), "sbz_mlp_forward")
        return y, ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.sbz_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

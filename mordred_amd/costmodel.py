"""ctypes binding of the learned cost model (include/mordred_costmodel.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

FEATURES = 32


class Stats(C.Structure):
    _fields_ = [("samples", C.c_longlong), ("scored", C.c_int), ("ewma_abs_log_error", C.c_double),
                ("ewma_abs_log_error_analytic", C.c_double), ("trusted", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_F = C.POINTER(C.c_float)
_SIGS = {
    "mordred_cost_model_create": ([C.c_int, C.c_ulonglong], C.c_void_p),
    "mordred_cost_model_destroy": ([C.c_void_p], None),
    "mordred_cost_model_last_error": ([C.c_void_p], C.c_char_p),
    "mordred_cost_model_num_params": ([], C.c_int),
    "mordred_cost_model_get_params": ([C.c_void_p, _F], C.c_int),
    "mordred_cost_model_set_params": ([C.c_void_p, _F], C.c_int),
    "mordred_cost_model_predict": ([C.c_void_p, _F, C.c_int, _F], C.c_int),
    "mordred_cost_model_gradients": ([C.c_void_p, _F, _F, C.c_int, _F, _F], C.c_int),
    "mordred_cost_model_step": ([C.c_void_p, _F, _F, C.c_int, C.c_float, _F], C.c_int),
    "mordred_cost_model_observe": ([C.c_void_p, _F, C.c_double, C.c_double], C.c_int),
    "mordred_cost_model_train": ([C.c_void_p, C.c_int, C.c_int, C.c_float, _F], C.c_int),
    "mordred_cost_model_get_stats": ([C.c_void_p, C.POINTER(Stats)], C.c_int),
}


def _lib():
    lib = abi.load()
    if not getattr(lib, "_cm_bound", False):
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        lib._cm_bound = True
    return lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


class CostModelError(RuntimeError):
    pass


class CostModel:
    """fp32 MLP 32 -> 64 -> 64 -> 1 on HIP device `device`."""

    def __init__(self, device: int = 0, seed: int = 1):
        self.lib = _lib()
        self.h = self.lib.mordred_cost_model_create(device, seed)
        if not self.h:
            raise CostModelError("mordred_cost_model_create failed")
        self.nparams = self.lib.mordred_cost_model_num_params()

    def close(self):
        if self.h:
            self.lib.mordred_cost_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise CostModelError(f"{what}: rc={rc} {self.lib.mordred_cost_model_last_error(self.h).decode()}")

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.lib.mordred_cost_model_get_params(self.h, _fp(p)), "get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self._check(self.lib.mordred_cost_model_set_params(self.h, _fp(p)), "set_params")

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        y = np.zeros(len(x), dtype=np.float32)
        self._check(self.lib.mordred_cost_model_predict(self.h, _fp(x), len(x), _fp(y)), "predict")
        return y

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.lib.mordred_cost_model_gradients(self.h, _fp(x), _fp(t), len(x), _fp(g), C.byref(loss)),
                    "gradients")
        return g, loss.value

    def step(self, x: np.ndarray, t: np.ndarray, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.lib.mordred_cost_model_step(self.h, _fp(x), _fp(t), len(x), lr, C.byref(loss)), "step")
        return loss.value

    def observe(self, x: np.ndarray, measured_ms: float, analytic_ms: float = -1.0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(FEATURES)
        self._check(self.lib.mordred_cost_model_observe(self.h, _fp(x), measured_ms, analytic_ms), "observe")

    def train(self, steps: int, batch: int = 256, lr: float = 1e-3) -> float:
        loss = C.c_float(0)
        self._check(self.lib.mordred_cost_model_train(self.h, steps, batch, lr, C.byref(loss)), "train")
        return loss.value

    def stats(self) -> dict:
        st = Stats()
        self._check(self.lib.mordred_cost_model_get_stats(self.h, C.byref(st)), "stats")
        return st.as_dict()


def split_params(p: np.ndarray):
    """(W1[32,64], b1, W2[64,64], b2, W3[64], b3) views of a parameter vector."""
    i = 0
    out = []
    for shape in ((FEATURES, 64), (64,), (64, 64), (64,), (64,), (1,)):
        n = int(np.prod(shape))
        out.append(p[i:i + n].reshape(shape))
        i += n
    return out

"""ctypes binding of the nexd_mlp_* entry points (include/nexd_gpu.h): a
multilayer perceptron trained and evaluated on the GPU by the library's own
kernels (BF16 matrix-core products with FP32 accumulation, GELU, MSE, Adam).
Nothing here computes on the CPU: without the library or a GPU, calls raise.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import capi

MLP_EXPORTS = ["nexd_mlp_create", "nexd_mlp_set_params", "nexd_mlp_get_params", "nexd_mlp_forward",
               "nexd_mlp_upload_dataset", "nexd_mlp_gradients", "nexd_mlp_train", "nexd_mlp_last_run_ms",
               "nexd_mlp_destroy", "nexd_mlp_last_error"]

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)


def _load():
    lib = capi.load()
    if not getattr(lib, "_mlp_ready", False):
        lib.nexd_mlp_create.argtypes = [ctypes.c_int, _I, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        lib.nexd_mlp_set_params.argtypes = [ctypes.c_void_p, ctypes.c_int, _F, _F]
        lib.nexd_mlp_get_params.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _F, _F]
        lib.nexd_mlp_forward.argtypes = [ctypes.c_void_p, _F, ctypes.c_int, _F]
        lib.nexd_mlp_upload_dataset.argtypes = [ctypes.c_void_p, _F, _F, ctypes.c_int]
        lib.nexd_mlp_gradients.argtypes = [ctypes.c_void_p, _I, ctypes.c_int, _F]
        lib.nexd_mlp_train.argtypes = [ctypes.c_void_p, _I, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, _F]
        lib.nexd_mlp_last_run_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        lib.nexd_mlp_destroy.argtypes = [ctypes.c_void_p]
        lib.nexd_mlp_destroy.restype = None
        lib.nexd_mlp_last_error.argtypes = [ctypes.c_void_p]
        lib.nexd_mlp_last_error.restype = ctypes.c_char_p
        for n in MLP_EXPORTS:
            if n not in ("nexd_mlp_destroy", "nexd_mlp_last_error"):
                getattr(lib, n).restype = ctypes.c_int
        lib._mlp_ready = True
    return lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_F)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def init_params(widths, seed: int = 0):
    """Weights N(0, 1/fan_in), zero biases (float32)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, o in zip(widths[:-1], widths[1:]):
        out.append(((rng.standard_normal((o, i)) / math.sqrt(i)).astype(np.float32), np.zeros(o, np.float32)))
    return out


class MLP:
    """widths = [in, hidden..., out]; batches are multiples of 64."""

    def __init__(self, widths, max_batch: int = 256, device: int = 0):
        self.lib = _load()
        self.widths = list(widths)
        w = (ctypes.c_int * len(widths))(*widths)
        h = ctypes.c_void_p()
        rc = self.lib.nexd_mlp_create(len(widths) - 1, w, max_batch, device, ctypes.byref(h))
        if rc != 0:
            raise capi.NexdError(f"nexd_mlp_create: {rc}: {self.lib.nexd_mlp_last_error(None).decode()}")
        self.h = h
        self.max_batch = max_batch
        self._keep = None

    def _chk(self, rc, what):
        if rc != 0:
            raise capi.NexdError(f"{what}: {rc}: {self.lib.nexd_mlp_last_error(self.h).decode()}")

    @property
    def nlayers(self):
        return len(self.widths) - 1

    def set_params(self, params):
        for l, (w, b) in enumerate(params):
            w, b = _f32(w), _f32(b)
            assert w.shape == (self.widths[l + 1], self.widths[l]) and b.shape == (self.widths[l + 1],)
            self._chk(self.lib.nexd_mlp_set_params(self.h, l, _f(w), _f(b)), "set_params")

    def _get(self, which):
        out = []
        for l in range(self.nlayers):
            w = np.zeros((self.widths[l + 1], self.widths[l]), np.float32)
            b = np.zeros(self.widths[l + 1], np.float32)
            self._chk(self.lib.nexd_mlp_get_params(self.h, l, which, _f(w), _f(b)), "get_params")
            out.append((w, b))
        return out

    def params(self):
        return self._get(0)

    def grads(self):
        return self._get(1)

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.zeros((x.shape[0], self.widths[-1]), np.float32)
        for s in range(0, x.shape[0], self.max_batch):
            xs = _f32(x[s:s + self.max_batch])
            ys = np.zeros((xs.shape[0], self.widths[-1]), np.float32)
            self._chk(self.lib.nexd_mlp_forward(self.h, _f(xs), xs.shape[0], _f(ys)), "forward")
            y[s:s + xs.shape[0]] = ys
        return y

    def upload_dataset(self, x, t):
        x, t = _f32(x), _f32(t)
        assert x.shape[0] == t.shape[0]
        self._chk(self.lib.nexd_mlp_upload_dataset(self.h, _f(x), _f(t), x.shape[0]), "upload_dataset")

    def gradients(self, idx) -> float:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ctypes.c_float()
        self._chk(self.lib.nexd_mlp_gradients(self.h, idx.ctypes.data_as(_I), idx.size, ctypes.byref(loss)),
                  "gradients")
        return float(loss.value)

    def train(self, idx, batch: int, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8) -> np.ndarray:
        """len(idx)//batch Adam steps over consecutive batches of idx; returns
        the loss of each step (before its update)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        nsteps = idx.size // batch
        losses = np.zeros(nsteps, np.float32)
        self._chk(self.lib.nexd_mlp_train(self.h, idx.ctypes.data_as(_I), batch, nsteps, lr, beta1, beta2, eps,
                                          _f(losses)), "train")
        return losses

    def last_run_ms(self) -> float:
        ms = ctypes.c_double()
        self._chk(self.lib.nexd_mlp_last_run_ms(self.h, ctypes.byref(ms)), "last_run_ms")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nexd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

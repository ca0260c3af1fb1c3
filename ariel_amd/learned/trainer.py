"""Python side of the library's network trainer (``ariel_trainer_*`` in
include/ariel.h): forward, backward and Adam run in the library's own HIP
kernels; this class only holds the handle and passes device pointers.

Data are torch tensors on cuda:0 (device memory and streams plumbing only),
shaped (batch, 1, N, N, N), float32, contiguous.
"""
import ctypes

import numpy as np

from .. import _abi
from . import layout, pack_flat, unpack_flat


class TrainerError(RuntimeError):
    pass


class Trainer:
    def __init__(self, init=None):
        """init: a weight blob (bytes) or a packed FP32 vector; None = zeros."""
        self._lib = _abi.load()
        self._h = ctypes.c_void_p()
        if self._lib.ariel_trainer_create(ctypes.byref(self._h)) != 0:
            raise TrainerError(self._lib.ariel_last_error(None).decode())
        self.count = layout()[1]
        if init is not None:
            self.set_weights(unpack_flat(init) if isinstance(init, (bytes, bytearray)) else init)

    def _check(self, rc):
        if rc != 0:
            raise TrainerError(self._lib.ariel_trainer_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.ariel_trainer_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        self._check(self._lib.ariel_trainer_set_weights(
            self._h, flat.ctypes.data_as(_abi._PF), flat.size))

    def get_weights(self):
        out = np.empty(self.count, np.float32)
        self._check(self._lib.ariel_trainer_get_weights(self._h, out.ctypes.data_as(_abi._PF), out.size))
        return out

    def blob(self):
        return pack_flat(self.get_weights())

    @staticmethod
    def _ptrs(*ts):
        out = []
        for t in ts:
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous()
            out.append(ctypes.c_void_p(t.data_ptr()))
        return out

    def step(self, r, e, diag, fluid, lr=1e-3, clip=1.0, update=True, want_grad=False):
        """One batch: returns the mean loss (and the gradient if want_grad)."""
        b, n = r.shape[0], r.shape[-1]
        loss = ctypes.c_double(0.0)
        grad = np.empty(self.count, np.float32) if want_grad else None
        self._check(self._lib.ariel_trainer_step(
            self._h, n, b, *self._ptrs(r, e, diag, fluid), float(lr), float(clip), int(bool(update)),
            ctypes.byref(loss), grad.ctypes.data_as(_abi._PF) if want_grad else None))
        return (loss.value, grad) if want_grad else loss.value

    def forward(self, r, diag, fluid, z):
        b, n = r.shape[0], r.shape[-1]
        self._check(self._lib.ariel_trainer_forward(self._h, n, b, *self._ptrs(r, diag, fluid, z)))
        return z

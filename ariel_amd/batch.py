"""Batched independent tanks over the C ABI (``ariel_batch_*`` in include/ariel.h).

Every stage of a batched step is one kernel launch over all tanks; each tank's
results equal those of stepping it alone. ``step_sequential`` is the baseline:
a loop of single-tank steps. The reference runs one FlipSim per process and has
no counterpart.
"""
import ctypes

from . import _abi
from .flipsim import ArielError, FlipSim


class FlipBatch:
    def __init__(self, descs):
        self._lib = _abi.load()
        self.descs = [_abi.SceneDesc.from_buffer_copy(d) for d in descs]
        arr = (_abi.SceneDesc * len(descs))(*self.descs)
        self._b = ctypes.c_void_p()
        rc = self._lib.ariel_batch_create(arr, len(descs), ctypes.byref(self._b))
        if rc != 0:
            raise ArielError(rc, self._lib.ariel_batch_last_error(None).decode())

    def _check(self, rc):
        if rc != 0:
            raise ArielError(rc, self._lib.ariel_batch_last_error(self._b).decode())

    def close(self):
        if self._b:
            self._lib.ariel_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = ctypes.c_int32()
        self._check(self._lib.ariel_batch_size(self._b, ctypes.byref(n)))
        return n.value

    def tank(self, i):
        """A FlipSim view of tank i (owned by the batch)."""
        h = ctypes.c_void_p()
        self._check(self._lib.ariel_batch_tank(self._b, i, ctypes.byref(h)))
        return FlipSim(self.descs[i], _handle=h.value)

    def Init(self):
        self._check(self._lib.ariel_batch_init(self._b))

    def autotune(self):
        self._check(self._lib.ariel_batch_autotune(self._b))

    def tuning(self, i):
        t = ctypes.c_int32()
        self._check(self._lib.ariel_batch_get_tuning(self._b, i, ctypes.byref(t)))
        return t.value

    def set_graphs(self, enable=True):
        self._check(self._lib.ariel_batch_set_graphs(self._b, 1 if enable else 0))

    def Step(self, nsteps=1):
        self._check(self._lib.ariel_batch_step(self._b, nsteps))

    def step_sequential(self, nsteps=1):
        self._check(self._lib.ariel_batch_step_sequential(self._b, nsteps))


__all__ = ["FlipBatch"]

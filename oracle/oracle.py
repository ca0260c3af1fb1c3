"""ctypes wrapper over oracle/_build/liboracle.so (see flip_oracle.h).

Test infrastructure: the oracle is a deterministic serial C++ restatement of the
reference's src/sim path; this wrapper mirrors the product binding's accessors so
that tests can drive both with the same calls.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ARIEL_ORACLE_LIB selects another build of the same source (the planted-mistake
# test loads deliberately broken copies through it).
LIB = os.environ.get("ARIEL_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

FAITHFUL = 1
EXTENDED = 0


class OracleStats(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("solve_status", ctypes.c_int32), ("fluid_cells", ctypes.c_int32),
                ("particles", ctypes.c_int64), ("rhs_norm", ctypes.c_double),
                ("residual_norm", ctypes.c_double), ("true_residual_norm", ctypes.c_double),
                ("max_divergence", ctypes.c_double), ("max_density", ctypes.c_double),
                ("stage_ms", ctypes.c_double * 10), ("step_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("precond_fallback", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("true_residual_norm64", ctypes.c_double)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    V, I, I64, F, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double
    sig = {
        "oracle_create": (V, [V, I]), "oracle_destroy": (None, [V]), "oracle_init": (I, [V]),
        "oracle_step": (I, [V, I]), "oracle_stage": (I, [V, I, I]),
        "oracle_num_particles": (I64, [V]), "oracle_get_particles": (None, [V, V, V, V]),
        "oracle_set_particles": (None, [V, V, V, V, I64]), "oracle_field_size": (I64, [V, I]),
        "oracle_get_field": (None, [V, I, V]), "oracle_set_field": (None, [V, I, V]),
        "oracle_get_flags": (None, [V, V]), "oracle_set_flags": (None, [V, V]),
        "oracle_get_sort": (None, [V, V, V]), "oracle_get_accept": (None, [V, V]),
        "oracle_get_max_density": (F, [V]), "oracle_set_max_density": (None, [V, F]),
        "oracle_get_stats": (None, [V, V]), "oracle_get_frame": (I, [V]),
        "oracle_set_frame": (None, [V, I]),
        "oracle_apply_A": (None, [V, V, V]), "oracle_apply_A_f64": (None, [V, V, V]), "oracle_apply_precond": (None, [V, V, V]),
        "oracle_get_precond": (None, [V, V]),
        "oracle_solve_fp64": (I, [V, D, I, V, V]),
        "oracle_smooth": (F, [F, F]), "oracle_sharpen": (F, [F, F]),
        "oracle_sqrlength": (F, [V, V]),
        "oracle_hash4": (ctypes.c_uint32, [ctypes.c_uint32] * 4),
        "oracle_interpolate_velocity": (None, [V, I, V, V]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleSim:
    """Oracle FlipSim. `desc` is any ctypes struct laid out as ariel_scene_desc."""

    def __init__(self, desc, threads=1):
        self.lib = load()
        raw = ctypes.create_string_buffer(bytes(desc), len(bytes(desc)))
        self.h = self.lib.oracle_create(ctypes.cast(raw, ctypes.c_void_p), threads)
        if not self.h:
            raise ValueError("oracle_create rejected the descriptor")
        self.n = desc.dims[0]

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self):
        self.lib.oracle_init(self.h)

    def step(self, mode=EXTENDED):
        self.lib.oracle_step(self.h, mode)

    def stage(self, stage, mode=EXTENDED):
        if self.lib.oracle_stage(self.h, stage, mode) != 0:
            raise ValueError(stage)

    @property
    def num_particles(self):
        return int(self.lib.oracle_num_particles(self.h))

    def get_particles(self):
        n = self.num_particles
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        self.lib.oracle_get_particles(self.h, _ptr(pos), _ptr(vel), _ptr(rho))
        return pos, vel, rho

    def set_particles(self, pos, vel, density=None):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        d = None if density is None else np.ascontiguousarray(density, np.float32)
        self.lib.oracle_set_particles(self.h, _ptr(pos), _ptr(vel),
                                      None if d is None else _ptr(d), len(pos))

    def get_field(self, field):
        out = np.zeros(self.lib.oracle_field_size(self.h, field), np.float32)
        self.lib.oracle_get_field(self.h, field, _ptr(out))
        return out

    def set_field(self, field, arr):
        arr = np.ascontiguousarray(arr, np.float32)
        assert arr.size == self.lib.oracle_field_size(self.h, field)
        self.lib.oracle_set_field(self.h, field, _ptr(arr))

    def get_flags(self):
        out = np.zeros(self.n ** 3, np.uint8)
        self.lib.oracle_get_flags(self.h, _ptr(out))
        return out

    def set_flags(self, flags):
        flags = np.ascontiguousarray(flags, np.uint8)
        self.lib.oracle_set_flags(self.h, _ptr(flags))

    def get_sort(self):
        ids = np.zeros(self.num_particles, np.int32)
        offsets = np.zeros(self.n ** 3 + 1, np.int32)
        self.lib.oracle_get_sort(self.h, _ptr(ids), _ptr(offsets))
        return ids, offsets

    def get_accept(self):
        out = np.zeros(self.num_particles, np.uint8)
        self.lib.oracle_get_accept(self.h, _ptr(out))
        return out

    @property
    def max_density(self):
        return float(self.lib.oracle_get_max_density(self.h))

    @max_density.setter
    def max_density(self, v):
        self.lib.oracle_set_max_density(self.h, v)

    @property
    def frame(self):
        return int(self.lib.oracle_get_frame(self.h))

    @frame.setter
    def frame(self, v):
        self.lib.oracle_set_frame(self.h, int(v))

    def stats(self):
        st = OracleStats()
        self.lib.oracle_get_stats(self.h, ctypes.byref(st))
        return st

    def apply_A(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        self.lib.oracle_apply_A(self.h, _ptr(x), _ptr(y))
        return y

    def apply_A_f64(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape, np.float64)
        self.lib.oracle_apply_A_f64(self.h, _ptr(x), _ptr(y))
        return y

    def apply_precond(self, r):
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        self.lib.oracle_apply_precond(self.h, _ptr(r), _ptr(z))
        return z

    def get_precond(self):
        out = np.zeros(self.n ** 3, np.float32)
        self.lib.oracle_get_precond(self.h, _ptr(out))
        return out

    def solve_fp64(self, tol=1e-12, max_iters=5000):
        p = np.zeros(self.n ** 3, np.float64)
        rel = ctypes.c_double(0.0)
        it = self.lib.oracle_solve_fp64(self.h, tol, max_iters, _ptr(p), ctypes.byref(rel))
        return it, p, rel.value

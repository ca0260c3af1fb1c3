"""Python mirror of fluidCore::FlipSim (src/sim/flip.hpp:20-34) over the C ABI.

``FlipSim(desc)`` creates the device state (FlipSim::FlipSim, flip.cpp:15-30);
``Init()``/``Step()`` follow FlipSim::Init/Step; ``GetParticles()``,
``GetDimensions()`` and ``m_frame`` are the reference accessors. The MAC-grid,
flag and sort accessors have no reference counterpart (m_mgrid is private).
Errors raise :class:`ArielError` carrying the ABI status and message.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import SceneDesc, StepStats


class ArielError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ariel error {code}: {msg}")
        self.code = code


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class FlipSim:
    def __init__(self, desc, _handle=None):
        self._lib = _abi.load()
        self.desc = desc
        self.n = desc.dims[0]
        self._owned = _handle is None
        if _handle is not None:
            # a view of a handle owned elsewhere (a batch tank): never destroyed here
            self._h = ctypes.c_void_p(_handle)
            return
        self._h = ctypes.c_void_p()
        rc = self._lib.ariel_create(ctypes.byref(desc), ctypes.byref(self._h))
        if rc != 0:
            raise ArielError(rc, self._lib.ariel_last_error(None).decode())
        if desc.precond == _abi.PRECOND_LEARNED and self.n % 16 == 0:
            # the shipped weights, when present; set_learned_weights replaces them
            from .learned import DEFAULT_WEIGHTS
            if os.path.exists(DEFAULT_WEIGHTS):
                self.set_learned_weights(DEFAULT_WEIGHTS)

    @classmethod
    def from_json(cls, text_or_path, base_dir=None, solver=None):
        """FlipSim with the reference scene's full geometry (ariel_create_from_json):
        settings, forces, mesh files (relative to base_dir, default the JSON file's
        folder), transforms, animated meshes, emitters; solver fields from `solver`."""
        lib = _abi.load()
        text = text_or_path
        if os.path.exists(str(text_or_path)):
            text = open(text_or_path).read()
            if base_dir is None:
                base_dir = os.path.dirname(os.path.abspath(str(text_or_path)))
        h = ctypes.c_void_p()
        rc = lib.ariel_create_from_json(text.encode(), (base_dir or "").encode(),
                                        ctypes.byref(solver) if solver is not None else None,
                                        ctypes.byref(h))
        if rc != 0:
            raise ArielError(rc, lib.ariel_last_error(None).decode())
        d = SceneDesc()
        lib.ariel_get_desc(h, ctypes.byref(d))
        sim = cls(d, _handle=h.value)
        sim._owned = True
        return sim

    # ------------------------------------------------------------ lifecycle
    def _check(self, rc):
        if rc != 0:
            raise ArielError(rc, self._lib.ariel_last_error(self._h).decode())

    def close(self):
        if self._h and self._owned:
            self._lib.ariel_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------- reference interface
    def Init(self):
        self._check(self._lib.ariel_init(self._h))

    def Step(self, nsteps=1):
        self._check(self._lib.ariel_step(self._h, nsteps))

    def StepAutotune(self, candidates=None):
        """One Step whose pressure solve is first timed with each candidate
        preconditioner (ariel_step_autotune); the fastest converged one becomes
        this handle's. Returns (trials, chosen): trials as dicts of precond,
        iterations, status, solve_ms."""
        if candidates is None:
            candidates = [_abi.PRECOND_MIC0_TILED, _abi.PRECOND_MG, _abi.PRECOND_MIC0_BLOCK,
                          _abi.PRECOND_JACOBI]
        c = (ctypes.c_int32 * len(candidates))(*candidates)
        tr = (_abi.SolverTrial * len(candidates))()
        chosen = ctypes.c_int32()
        self._check(self._lib.ariel_step_autotune(self._h, c, len(candidates), tr, ctypes.byref(chosen)))
        return ([{"precond": t.precond, "iterations": t.iterations, "status": t.status,
                  "solve_ms": t.solve_ms} for t in tr], chosen.value)

    def GetParticles(self, types=False):
        """(pos[n,3], vel[n,3], density[n]) in particle-vector (id) order; with
        types=True also Particle::m_type (macgrid.inl:36-42) as uint8[n]."""
        n = self.num_particles
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        rho = np.empty(n, np.float32)
        typ = np.empty(n, np.uint8) if types else None
        cnt = ctypes.c_int64()
        self._check(self._lib.ariel_get_particles(self._h, _p(pos), _p(vel), _p(rho),
                                                  _p(typ) if types else None, n,
                                                  ctypes.byref(cnt)))
        return (pos, vel, rho, typ) if types else (pos, vel, rho)

    def GetScene(self):
        """FlipSim::GetScene (flip.hpp:31): the scene description the handle runs
        with (ariel_get_desc; the descriptor replaces sceneCore::Scene): the one it
        was created from, with the preconditioner an autotuned step picked."""
        d = SceneDesc()
        self._check(self._lib.ariel_get_desc(self._h, ctypes.byref(d)))
        return d

    def GetDimensions(self):
        d = (ctypes.c_int32 * 3)()
        self._check(self._lib.ariel_get_dimensions(self._h, d))
        return tuple(d)

    @property
    def m_frame(self):
        f = ctypes.c_int32()
        self._check(self._lib.ariel_get_frame(self._h, ctypes.byref(f)))
        return f.value

    # ------------------------------------------------------------ extras
    init = Init
    step = Step

    @property
    def num_particles(self):
        n = ctypes.c_int64()
        self._check(self._lib.ariel_particle_count(self._h, ctypes.byref(n)))
        return n.value

    def run_stage(self, stage):
        self._check(self._lib.ariel_run_stage(self._h, stage))

    def sync(self):
        self._check(self._lib.ariel_sync(self._h))

    def set_particles(self, pos, vel=None, density=None):
        pos = np.ascontiguousarray(pos, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        d = None if density is None else np.ascontiguousarray(density, np.float32)
        self._check(self._lib.ariel_set_particles(self._h, _p(pos), None if v is None else _p(v),
                                                  None if d is None else _p(d), len(pos)))

    def field_size(self, which):
        n = self.n
        return (n + 1) * n * n if which <= _abi.FIELD_DUZ else n ** 3

    def get_field(self, which):
        out = np.empty(self.field_size(which), np.float32)
        self._check(self._lib.ariel_get_macgrid(self._h, which, _p(out), out.size))
        return out

    def set_field(self, which, arr):
        arr = np.ascontiguousarray(arr, np.float32).ravel()
        self._check(self._lib.ariel_set_macgrid(self._h, which, _p(arr), arr.size))

    def get_flags(self):
        out = np.empty(self.n ** 3, np.uint8)
        self._check(self._lib.ariel_get_flags(self._h, _p(out), out.size))
        return out

    def set_flags(self, flags):
        flags = np.ascontiguousarray(flags, np.uint8).ravel()
        self._check(self._lib.ariel_set_flags(self._h, _p(flags), flags.size))

    def get_sort(self):
        ids = np.empty(self.num_particles, np.int32)
        offs = np.empty(self.n ** 3 + 1, np.int32)
        self._check(self._lib.ariel_get_sort(self._h, _p(ids), ids.size, _p(offs), offs.size))
        return ids, offs

    def get_accept(self):
        out = np.empty(self.num_particles, np.uint8)
        self._check(self._lib.ariel_get_resample_accept(self._h, _p(out), out.size))
        return out

    @property
    def max_density(self):
        v = ctypes.c_float()
        self._check(self._lib.ariel_get_max_density(self._h, ctypes.byref(v)))
        return v.value

    @max_density.setter
    def max_density(self, v):
        self._check(self._lib.ariel_set_max_density(self._h, float(v)))

    def stats(self):
        st = StepStats()
        self._check(self._lib.ariel_get_stats(self._h, ctypes.byref(st)))
        return st

    def apply_laplacian(self, x):
        x = np.ascontiguousarray(x, np.float32).ravel()
        y = np.empty_like(x)
        self._check(self._lib.ariel_apply_laplacian(self._h, _p(x), _p(y)))
        return y

    def apply_preconditioner(self, r):
        r = np.ascontiguousarray(r, np.float32).ravel()
        z = np.empty_like(r)
        self._check(self._lib.ariel_apply_preconditioner(self._h, _p(r), _p(z)))
        return z

    def set_profiling(self, enable=True):
        self._check(self._lib.ariel_set_profiling(self._h, 1 if enable else 0))

    def reset_kernel_stats(self):
        self._check(self._lib.ariel_reset_kernel_stats(self._h))

    def kernel_stats(self, kernel):
        ks = _abi.KernelStats()
        self._check(self._lib.ariel_get_kernel_stats(self._h, kernel, ctypes.byref(ks)))
        return ks

    def save_state(self, path):
        self._check(self._lib.ariel_save_state(self._h, str(path).encode()))

    def load_state(self, path):
        self._check(self._lib.ariel_load_state(self._h, str(path).encode()))

    def set_learned_weights(self, path_or_bytes=None):
        """Load the learned preconditioner's packed weights (default: the file
        shipped in ariel_amd/learned/)."""
        if path_or_bytes is None:
            from .learned import DEFAULT_WEIGHTS
            path_or_bytes = DEFAULT_WEIGHTS
        blob = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        self._check(self._lib.ariel_set_learned_weights(self._h, buf, len(blob)))

    def solve_rhs(self, rhs):
        """x with A x = rhs (current classes and level set), from x = 0; returns (x, stats)."""
        rhs = np.ascontiguousarray(rhs, np.float32).reshape(-1)
        x = np.empty_like(rhs)
        st = StepStats()
        self._check(self._lib.ariel_solve_rhs(
            self._h, rhs.ctypes.data_as(_abi._PF), x.ctypes.data_as(_abi._PF), ctypes.byref(st)))
        return x, st

    def set_levelset_mode(self, mode):
        """_abi.LEVELSET_SPHERES (default) or _abi.LEVELSET_TRAILS (the reference's
        rasterizeTrails) for later level-set and mesh exports."""
        self._check(self._lib.ariel_set_levelset_mode(self._h, mode))

    def export_levelset(self, path):
        """The particle level set; an OpenVDB file if path ends in ".vdb", else the
        raw volume that read_levelset reads."""
        self._check(self._lib.ariel_export_levelset(self._h, str(path).encode()))

    def export_mesh(self, path):
        """OBJ of the particle level set's zero surface (Scene::ExportParticles' OBJ path)."""
        self._check(self._lib.ariel_export_mesh(self._h, str(path).encode()))

    def export_particles(self, path):
        """Partio particle file (Scene::ExportParticles' Partio path): ".geo" or
        ".bgeo", optionally gzipped (".gz")."""
        self._check(self._lib.ariel_export_particles(self._h, str(path).encode()))


def read_levelset(path):
    """The volume ariel_export_levelset writes: (phi as an (M, M, M) float32 array
    indexed [i, j, k], background value)."""
    with open(path, "rb") as f:
        raw = f.read()
    lines, pos = [], 0
    while len(lines) < 5:
        end = raw.index(b"\n", pos)
        lines.append(raw[pos:end].decode())
        pos = end + 1
    if lines[0] != "ARIEL_LEVELSET 1" or lines[4] != "float32 le":
        raise ValueError("not an ARIEL_LEVELSET 1 file")
    dims = [int(v) for v in lines[1].split()[1:4]]
    background = float(lines[3].split()[1])
    phi = np.frombuffer(raw, "<f4", dims[0] * dims[1] * dims[2], pos).reshape(dims)
    return phi, background


def mesh_sdf(phi, path):
    """The library's mesher on a caller's level set (M^3 node values, indexed
    [i, j, k]), host only: writes an OBJ and returns (vertices, quads)."""
    lib = _abi.load()
    phi = np.ascontiguousarray(phi, np.float32)
    if phi.ndim != 3 or len(set(phi.shape)) != 1:
        raise ValueError("phi must be an M x M x M array")
    nv, nq = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.ariel_mesh_sdf(phi.ctypes.data_as(_abi._PF), phi.shape[0], str(path).encode(),
                            ctypes.byref(nv), ctypes.byref(nq))
    if rc != 0:
        raise ArielError(rc, lib.ariel_last_error(None).decode())
    return nv.value, nq.value


def write_vdb(phi, path, background=3.0):
    """The library's OpenVDB writer on a caller's M^3 node volume (host only)."""
    lib = _abi.load()
    phi = np.ascontiguousarray(phi, np.float32)
    if phi.ndim != 3 or len(set(phi.shape)) != 1:
        raise ValueError("phi must be an M x M x M array")
    rc = lib.ariel_write_vdb(phi.ctypes.data_as(_abi._PF), phi.shape[0], background, str(path).encode())
    if rc != 0:
        raise ArielError(rc, lib.ariel_last_error(None).decode())


def write_particles(path, pos, vel):
    """The library's Partio writer on caller data (positions in grid units), host only."""
    lib = _abi.load()
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    if len(pos) != len(vel):
        raise ValueError("pos and vel differ in length")
    rc = lib.ariel_write_particles(str(path).encode(), len(pos), pos.ctypes.data_as(_abi._PF),
                                   vel.ctypes.data_as(_abi._PF))
    if rc != 0:
        raise ArielError(rc, lib.ariel_last_error(None).decode())


__all__ = ["FlipSim", "ArielError", "SceneDesc", "read_levelset", "mesh_sdf", "write_vdb",
           "write_particles"]

"""Scene descriptors for the dam-break workloads.

A scene is an ``ariel_scene_desc`` (include/ariel.h): the host-side replacement
of the reference's JSON scene (src/scene/sceneloader.cpp:11-131) reduced to what
the FLIP step reads: grid size, lattice density, step size, global forces and
liquid primitives with initial velocities. Positions are unit-box coordinates.
"""
from . import _abi
from ._abi import SceneDesc


def defaults():
    """Reference defaults (sceneloader.cpp:411-443, flip.cpp:15-30, solver.inl:107)."""
    lib = _abi.load()
    d = SceneDesc()
    rc = lib.ariel_desc_defaults(d)
    if rc != 0:
        raise RuntimeError("ariel_desc_defaults failed")
    return d


def raw_defaults():
    """The same defaults without loading the native library (host-only use)."""
    d = SceneDesc()
    d.dims[:] = [32, 32, 32]
    d.density = 0.5
    d.dt = 0.005
    d.picflip = 0.95
    d.n_forces = 0
    d.n_liquids = 0
    d.wall = 0
    d.seed = 0
    d.jitter = 0
    d.precond = _abi.PRECOND_MIC0_TILED
    d.mic_tau = 0.97
    d.mic_sigma = 0.25
    d.tile = 8
    d.tol = 1e-6
    d.max_iters = 500
    d.refine_every = 0
    d.stop_rule = _abi.STOP_RELATIVE
    d.splat = _abi.SPLAT_SHARPEN
    d.resample = 1
    d.sharpen_re = 1.4
    d.spring = 50.0
    return d


def add_force(d, f):
    if d.n_forces >= _abi.MAX_FORCES:
        raise ValueError("too many forces")
    d.forces[d.n_forces][:] = list(f)
    d.n_forces += 1


def add_box(d, lo, hi, velocity=(0.0, 0.0, 0.0)):
    if d.n_liquids >= _abi.MAX_LIQUIDS:
        raise ValueError("too many liquids")
    lq = d.liquids[d.n_liquids]
    lq.kind = _abi.LIQUID_BOX
    lq.lo[:] = list(lo)
    lq.hi[:] = list(hi)
    lq.velocity[:] = list(velocity)
    d.n_liquids += 1


def add_sphere(d, center, radius, velocity=(0.0, 0.0, 0.0)):
    if d.n_liquids >= _abi.MAX_LIQUIDS:
        raise ValueError("too many liquids")
    lq = d.liquids[d.n_liquids]
    lq.kind = _abi.LIQUID_SPHERE
    lq.center[:] = list(center)
    lq.radius = radius
    lq.velocity[:] = list(velocity)
    d.n_liquids += 1


def dam_break(n=32, seed=0, **overrides):
    """Dam break on an n^3 MAC grid with 1-cell SOLID walls: liquid where x<0.4 and
    y<0.6, 8 particles per cell jittered inside each octant, gravity -9.8 in y,
    dt 1/240, FLIP 0.95, MIC(0) tau 0.97 sigma 0.25, relative 1e-6, <=500 iterations."""
    d = raw_defaults()
    d.dims[:] = [n, n, n]
    d.density = 0.5
    d.dt = 1.0 / 240.0
    d.picflip = 0.95
    d.wall = 1
    d.seed = seed
    d.jitter = 1
    d.tol = 1e-6
    d.max_iters = 500
    add_force(d, (0.0, -9.8, 0.0))
    add_box(d, (0.0, 0.0, 0.0), (0.4, 0.6, 1.0))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def dam_break_sphere(n=128, seed=0, **overrides):
    """Dam break plus a liquid sphere of radius 0.12 at (0.7, 0.8, 0.5) moving at (0,-1,0)."""
    d = dam_break(n, seed=seed)
    add_sphere(d, (0.7, 0.8, 0.5), 0.12, (0.0, -1.0, 0.0))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def pool_and_cube(n=256, seed=1, **overrides):
    """Pool of depth 0.3 plus a falling liquid cube of side 0.25 centred at height 0.7,
    jitter seed 1, true-residual recompute every 25 iterations."""
    d = raw_defaults()
    d.dims[:] = [n, n, n]
    d.density = 0.5
    d.dt = 1.0 / 240.0
    d.picflip = 0.95
    d.wall = 1
    d.seed = seed
    d.jitter = 1
    d.tol = 1e-6
    d.max_iters = 500
    d.refine_every = 25
    add_force(d, (0.0, -9.8, 0.0))
    add_box(d, (0.0, 0.0, 0.0), (1.0, 0.3, 1.0))
    add_box(d, (0.375, 0.575, 0.375), (0.625, 0.825, 0.625))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def tank(n, frac, seed=2, **overrides):
    """One tank of the batched workload: a dam break whose liquid block spans
    frac = (fx, fy, fz) of the unit box per axis."""
    d = raw_defaults()
    d.dims[:] = [n, n, n]
    d.density = 0.5
    d.dt = 1.0 / 240.0
    d.picflip = 0.95
    d.wall = 1
    d.seed = seed
    d.jitter = 1
    add_force(d, (0.0, -9.8, 0.0))
    add_box(d, (0.0, 0.0, 0.0), tuple(frac))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def batched_tanks(ntanks=64, seed=2, sizes=(48, 64, 80, 96), lo=0.2, hi=0.6, **overrides):
    """The batched workload: ntanks dam breaks, edges drawn uniformly from `sizes`
    and liquid-block fractions uniformly from [lo, hi] per axis (numpy generator
    seeded with `seed`), each tank seeded with jitter seed `seed`."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ntanks):
        n = int(rng.choice(sizes))
        frac = tuple(float(x) for x in rng.uniform(lo, hi, size=3))
        out.append(tank(n, frac, seed=seed, **overrides))
    return out


def from_json(text_or_path, **overrides):
    """A descriptor from the reference's JSON scene format (a path or the text),
    through ariel_desc_from_json; keyword overrides set descriptor fields after."""
    import os
    text = text_or_path
    if isinstance(text_or_path, str) and os.path.exists(text_or_path):
        text = open(text_or_path).read()
    lib = _abi.load()
    d = SceneDesc()
    rc = lib.ariel_desc_from_json(text.encode(), d)
    if rc != 0:
        from .flipsim import ArielError
        raise ArielError(rc, lib.ariel_last_error(None).decode())
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def emit(text, frame, base_dir=None, solver=None):
    """The particles a JSON scene's liquid geoms emit on `frame` (ariel_scene_emit,
    host only): (pos, vel), float32 arrays of shape (n, 3)."""
    import ctypes
    import numpy as np
    lib = _abi.load()
    n = ctypes.c_int64()
    args = (text.encode(), (base_dir or "").encode(),
            ctypes.byref(solver) if solver is not None else None, frame)
    rc = lib.ariel_scene_emit(*args, None, None, 0, ctypes.byref(n))
    if rc == 0:
        pos = np.zeros((n.value, 3), np.float32)
        vel = np.zeros((n.value, 3), np.float32)
        PF = ctypes.POINTER(ctypes.c_float)
        rc = lib.ariel_scene_emit(*args, pos.ctypes.data_as(PF), vel.ctypes.data_as(PF), n.value,
                                  ctypes.byref(n))
    if rc != 0:
        from .flipsim import ArielError
        raise ArielError(rc, lib.ariel_last_error(None).decode())
    return pos, vel


def copy(d):
    out = SceneDesc()
    out_bytes = bytes(d)
    import ctypes
    ctypes.memmove(ctypes.addressof(out), out_bytes, len(out_bytes))
    return out

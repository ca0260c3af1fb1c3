"""Scenes with the reference's full geometry on the GPU (ariel_create_from_json):
Init seeds the liquid geoms' frame-0 particles, and a persistent emitter adds its
particles at the start of every step (Scene::GenerateParticles at the step's
frame, src/sim/flip.cpp:98, scene.cpp:257-300). A step with emission matches the
oracle stepped from the same particles (the GPU's plus the emitted ones)."""
import json

import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from oracle import OracleSim
from tests.helpers import FIELDS, assert_bitexact

pytestmark = pytest.mark.gpu
N = 32


def _scene():
    return json.dumps({
        "settings": [{"density": 0.5, "step_size": 1.0 / 240.0, "dim": {"x": N, "y": N, "z": N}}],
        "globalforces": [{"x": 0, "y": -9.8, "z": 0}],
        "meshfiles": [{"id": "pool", "mesh_gen": "box", "point0": {"x": 1, "y": 1, "z": 1},
                       "point1": {"x": 31, "y": 9, "z": 31}},
                      {"id": "jet", "mesh_gen": "sphere", "center": {"x": 16, "y": 24, "z": 16}, "radius": 3}],
        "transforms": [{"id": "t"}, {"id": "r", "rotation": {"x": 0, "y": 30, "z": 0}}],
        "geoms": [{"id": "g0", "type": "mesh", "geom_frames": ["pool"], "transform_frames": ["t"],
                   "pre_persist": False, "post_persist": False, "frame_interval": 1, "frame_offset": 0},
                  {"id": "g1", "type": "mesh", "geom_frames": ["jet"], "transform_frames": ["r"],
                   "pre_persist": False, "post_persist": True, "frame_interval": 1, "frame_offset": 0}],
        "sim": [{"geom": "g0", "type": "liquid"},
                {"geom": "g1", "type": "liquid", "velocity": {"x": 0, "y": -2, "z": 0}}]})


def test_emitter_scene_steps_like_the_oracle():
    text = _scene()
    solver = scenes.dam_break(N)
    sim = FlipSim.from_json(text, solver=solver)
    sim.Init()
    e0, _ = scenes.emit(text, 0, solver=solver)
    assert sim.num_particles == len(e0)
    pos0, vel0, rho0 = sim.GetParticles()
    assert np.array_equal(pos0, e0)
    # the oracle from the GPU's state after Init plus the frame-1 emission
    e1, v1 = scenes.emit(text, 1, solver=solver)
    assert len(e1) > 0
    o = OracleSim(sim.GetScene(), threads=16)
    o.set_particles(np.concatenate([pos0, e1]), np.concatenate([vel0, v1]),
                    np.concatenate([rho0, np.full(len(e1), 10.0, np.float32)]))
    for f in FIELDS:
        o.set_field(f, sim.get_field(f))
    o.set_flags(sim.get_flags())
    o.max_density = sim.max_density
    o.frame = 0
    print("emitter scene:", len(pos0), "particles, +", len(e1), "per step", flush=True)
    sim.Step(1)
    o.step()
    assert sim.num_particles == len(pos0) + len(e1)
    assert_bitexact(sim.get_flags(), o.get_flags(), "cell classes after an emitting step")
    pg, vg, _ = sim.GetParticles()
    po, vo, _ = o.get_particles()
    assert pg.shape == po.shape
    assert np.abs(pg - po).max() < 1e-5
    assert np.abs(vg - vo).max() < 1e-3 * max(1.0, float(np.abs(vo).max()))
    # later steps keep emitting
    sim.Step(2)
    more = sum(len(scenes.emit(text, f, solver=solver)[0]) for f in (2, 3))
    assert sim.num_particles == len(pos0) + len(e1) + more
    assert sim.stats().solve_status == 0
    sim.close()
    o.close()

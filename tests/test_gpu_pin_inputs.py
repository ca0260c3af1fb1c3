"""The GPU stages on the inputs the oracle's brute-force and closed-form pins use
(test_oracle_pins.py): unstructured random particles through sort, density,
marks, splat and resample; a linear level set with p = c * L through the
ghost-fluid gradient; random cell classes through the boundary pass. Each is
compared with the oracle bit for bit, which carries those pins over to the GPU."""
import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from oracle import OracleSim
from tests.helpers import assert_bitexact

pytestmark = pytest.mark.gpu


def _pair(n):
    d = scenes.dam_break(n)
    g = FlipSim(d)
    g.Init()
    o = OracleSim(d)
    o.init()
    return d, g, o


def test_random_particles_through_the_particle_stages():
    n, m = 16, 1800
    d, g, o = _pair(n)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.04, 0.62, (m, 3)).astype(np.float32)
    vel = rng.normal(0.0, 1.0, (m, 3)).astype(np.float32)
    for sim in (g, o):
        sim.set_particles(pos, vel, np.zeros(m, np.float32))
    stages = [(_abi.STAGE_SORT, "sort"), (_abi.STAGE_DENSITY, "density"), (_abi.STAGE_MARK, "mark"),
              (_abi.STAGE_P2G, "p2g"), (_abi.STAGE_RESAMPLE, "resample")]
    for st, name in stages:
        print("stage", name, flush=True)
        g.run_stage(st)
        o.stage(st)
        if name == "sort":
            assert_bitexact(g.get_sort()[0], o.get_sort()[0], "sort ids")
        if name in ("density", "resample"):
            for a, b, what in zip(g.GetParticles(), o.get_particles(), ("pos", "vel", "rho")):
                assert_bitexact(a, b, f"{name} {what}")
        if name == "mark":
            assert_bitexact(g.get_flags(), o.get_flags(), "cell classes")
            assert_bitexact(g.get_field(_abi.FIELD_L), o.get_field(_abi.FIELD_L), "cell SDF")
        if name == "p2g":
            for comp in range(3):
                assert_bitexact(g.get_field(_abi.FIELD_UX + comp), o.get_field(_abi.FIELD_UX + comp),
                                f"face velocity {comp}")
        if name == "resample":
            assert_bitexact(g.get_accept(), o.get_accept(), "accept mask")
    g.close()
    o.close()


def _gradient_case(g, o, n, L, P, flags, seed):
    rng = np.random.default_rng(seed)
    for sim in (g, o):
        if flags is not None:
            sim.set_flags(flags)
        sim.set_field(_abi.FIELD_L, L)
        sim.set_field(_abi.FIELD_P, P)
    for comp in range(3):
        base = rng.uniform(-1, 1, o.get_field(_abi.FIELD_UX + comp).size).astype(np.float32)
        g.set_field(_abi.FIELD_UX + comp, base)
        o.set_field(_abi.FIELD_UX + comp, base)
    g.run_stage(_abi.STAGE_GRADIENT)
    o.stage(_abi.STAGE_GRADIENT)
    for comp in range(3):
        assert_bitexact(g.get_field(_abi.FIELD_UX + comp), o.get_field(_abi.FIELD_UX + comp),
                        f"face velocity {comp}")


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [1, -1])
def test_ghost_fluid_gradient_on_a_linear_level_set(axis, side):
    n = 16
    print("ghost gradient axis", axis, "side", side, flush=True)
    _, g, o = _pair(n)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = [i, j, k][axis].astype(np.float32)
    L = (np.float32(side) * np.float32(0.5) * (x - np.float32(7.7))).astype(np.float32)
    P = np.where(L < 0, np.float32(3.0) * L, np.float32(0)).astype(np.float32)
    _gradient_case(g, o, n, L.ravel(), P.ravel(), None, axis)
    g.close()
    o.close()


def test_boundary_pass_on_random_cell_classes():
    n = 16
    print("boundary pass, random classes", flush=True)
    _, g, o = _pair(n)
    rng = np.random.default_rng(13)
    flags = rng.choice([_abi.AIR, _abi.FLUID, _abi.SOLID], size=n ** 3, p=[0.3, 0.4, 0.3]).astype(np.uint8)
    L = rng.uniform(-1, 1, n ** 3).astype(np.float32)
    _gradient_case(g, o, n, L, np.zeros(n ** 3, np.float32), flags, 7)
    g.close()
    o.close()


def test_coincident_particles_through_the_resampler():
    """The resampler's coincident-pair branch (the hash in place of rand()) on the
    isolated coincident groups of the oracle's pin, plus a random cloud around
    them: positions, velocities and the accept mask bit for bit."""
    from tests.test_oracle_pins import coincident_groups
    n = 16
    d = scenes.dam_break(n, seed=9)
    g, o = FlipSim(d), OracleSim(d)
    g.Init()
    o.init()
    pos, _ = coincident_groups(n)
    rng = np.random.default_rng(21)
    cloud = rng.uniform(0.1, 0.9, (600, 3)).astype(np.float32)
    pos = np.concatenate([pos, cloud, cloud[:40]])   # 40 more exact duplicates
    vel = rng.normal(0.0, 1.0, pos.shape).astype(np.float32)
    for sim in (g, o):
        sim.set_particles(pos, vel, np.zeros(len(pos), np.float32))
    print("coincident resample", len(pos), flush=True)
    g.run_stage(_abi.STAGE_RESAMPLE)
    o.stage(_abi.STAGE_RESAMPLE)
    for a, b, what in zip(g.GetParticles(), o.get_particles(), ("pos", "vel", "rho")):
        assert_bitexact(a, b, f"resample {what}")
    assert_bitexact(g.get_accept(), o.get_accept(), "accept mask")
    assert np.abs(g.GetParticles()[0][:60] - pos[:60]).max() > 1e-5   # the groups moved
    g.close()
    o.close()

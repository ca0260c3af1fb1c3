"""Tiled P2G (ARIEL_P2G_TILED): the scatter from 8^3 cell tiles gives the face
velocities of particlegridoperations.inl:146-249 to within 1e-5 of the largest
particle speed, against the oracle (whose sums are the reference's, in the
reference's order), for both splat kernels, on random particles that reach the
domain's faces and on the dam-break and 128^3 scenes; it is deterministic (two
launches give the same bits) and a batch of tanks gets each tank's single-tank
result bit for bit."""
import numpy as np
import pytest

from ariel_amd import FlipBatch, FlipSim, _abi, scenes
from oracle import OracleSim
from tests.helpers import assert_bitexact, copy_state

pytestmark = pytest.mark.gpu

TOL = 1e-5   # max |u_tiled - u_oracle| / max particle speed component


def _faces(sim):
    return [sim.get_field(f) for f in (_abi.FIELD_UX, _abi.FIELD_UY, _abi.FIELD_UZ,
                                       _abi.FIELD_DUX, _abi.FIELD_DUY, _abi.FIELD_DUZ)]


def _check(g, o, vel, what):
    scale = float(np.abs(vel).max())
    for comp, (a, b) in enumerate(zip(_faces(g), _faces(o))):
        err = float(np.abs(a.astype(np.float64) - b).max()) / scale
        print(f"{what} component {comp}: max error {err:.3e} of the largest speed", flush=True)
        assert err <= TOL, f"{what} component {comp}: {err:.3e}"


@pytest.mark.parametrize("splat", [_abi.SPLAT_SHARPEN, _abi.SPLAT_TRILINEAR])
def test_random_particles_against_the_oracle(splat):
    n, m = 24, 40000
    d = scenes.dam_break(n, splat=splat, p2g=_abi.P2G_TILED)
    g = FlipSim(d)
    g.Init()
    o = OracleSim(scenes.dam_break(n, splat=splat))
    o.init()
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 1.0, (m, 3)).astype(np.float32)
    pos[:64] = rng.integers(0, 2, (64, 3)).astype(np.float32)   # on the box's corners and faces
    pos[64:128, 1] = 1.0
    vel = rng.normal(0.0, 2.0, (m, 3)).astype(np.float32)
    for sim in (g, o):
        sim.set_particles(pos, vel, np.zeros(m, np.float32))
    for st in (_abi.STAGE_SORT, _abi.STAGE_DENSITY, _abi.STAGE_MARK):
        g.run_stage(st)
        o.stage(st)
    assert_bitexact(g.get_flags(), o.get_flags(), "cell classes")
    g.run_stage(_abi.STAGE_P2G)
    o.stage(_abi.STAGE_P2G)
    _check(g, o, vel, f"splat {splat}")
    first = _faces(g)
    g.run_stage(_abi.STAGE_P2G)
    for comp, (a, b) in enumerate(zip(_faces(g), first)):
        assert_bitexact(a, b, f"second launch, field {comp}")
    g.close()
    o.close()


@pytest.mark.parametrize("n,steps", [(32, 3), (128, 1)])
def test_scene_state_against_the_oracle(n, steps):
    """Oracle steps the scene (dam break, plus the falling sphere at 128^3); the
    GPU's tiled P2G then runs on the oracle's state."""
    d = scenes.dam_break_sphere(n) if n == 128 else scenes.dam_break(n)
    o = OracleSim(d)
    o.init()
    for _ in range(steps):
        o.step()
    for st in (_abi.STAGE_SORT, _abi.STAGE_DENSITY, _abi.STAGE_MARK):
        o.stage(st)
    dt = scenes.dam_break_sphere(n) if n == 128 else scenes.dam_break(n)
    dt.p2g = _abi.P2G_TILED
    g = FlipSim(dt)
    copy_state(o, g)
    g.run_stage(_abi.STAGE_P2G)
    o.stage(_abi.STAGE_P2G)
    _check(g, o, o.get_particles()[1], f"{n}^3 after {steps} steps")
    g.close()
    o.close()


def test_tiled_batch_matches_single_tanks():
    descs = [scenes.tank(16, (0.3, 0.5, 0.6), p2g=_abi.P2G_TILED),
             scenes.tank(24, (0.5, 0.4, 0.3), p2g=_abi.P2G_TILED),
             scenes.tank(40, (0.25, 0.55, 0.45), p2g=_abi.P2G_TILED)]
    ref = []
    for d in descs:
        s = FlipSim(d)
        s.Init()
        s.Step(2)
        ref.append(s.GetParticles())
        s.close()
    b = FlipBatch(descs)
    b.Init()
    b.Step(2)
    for i, r in enumerate(ref):
        for name, x, y in zip(("pos", "vel", "density"), b.tank(i).GetParticles(), r):
            assert_bitexact(x, y, f"tank {i} {name}")
    b.close()


def test_tiled_mode_needs_a_short_sharpen_radius():
    d = scenes.dam_break(16, p2g=_abi.P2G_TILED, sharpen_re=2.0)
    with pytest.raises(Exception, match="radius"):
        FlipSim(d)


def test_free_running_steps_stay_close_to_the_ordered_mode():
    """Ten full steps of the 32^3 dam break in each mode from the same Init: every
    solve converges, and the tiled sums' last-bit differences grow no further
    than 1e-4 of the largest speed in the particle velocities and 1e-3 cells in
    positions (measured: 3.4e-7 of the largest speed and 1.3e-5 cells, the
    positions passing through the resampler's spring step)."""
    sims = []
    for mode in (_abi.P2G_ORDERED, _abi.P2G_TILED):
        s = FlipSim(scenes.dam_break(32, p2g=mode))
        s.Init()
        for _ in range(10):
            s.Step(1)
            assert s.stats().solve_status == 0
        sims.append(s)
    (p0, v0, _), (p1, v1, _) = (s.GetParticles() for s in sims)
    assert p0.shape == p1.shape
    dv = float(np.abs(v1.astype(np.float64) - v0).max()) / float(np.abs(v0).max())
    dp = float(np.abs(p1.astype(np.float64) - p0).max()) * 32
    print(f"after 10 steps: velocity {dv:.3e} of the largest speed, position {dp:.3e} cells", flush=True)
    assert dv <= 1e-4 and dp <= 1e-3
    for s in sims:
        s.close()

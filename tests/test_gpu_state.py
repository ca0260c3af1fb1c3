"""State dump and restore through the C ABI (ariel_save_state / ariel_load_state)
and the boundary details around a loaded or hand-set state: step k, save, step m;
a fresh handle of the same descriptor loads the file and steps m; particles, cell
classes, the pressure (the solver's warm start, solver.inl:299), the face grids and
the frame (flip.hpp:33) agree bit for bit. Truncated files and files of another
descriptor are refused with ARIEL_ERR_IO."""
import numpy as np
import pytest

from ariel_amd import ArielError, FlipSim, _abi, scenes
from tests.helpers import FIELDS, assert_bitexact

pytestmark = pytest.mark.gpu


def _state(sim):
    pos, vel, rho = sim.GetParticles()
    out = {"pos": pos, "vel": vel, "rho": rho, "flags": sim.get_flags(), "frame": sim.m_frame}
    for f in FIELDS:
        out[f"field{f}"] = sim.get_field(f)
    return out


@pytest.mark.parametrize("desc_fn", [lambda: scenes.dam_break(32),
                                     lambda: scenes.pool_and_cube(32, refine_every=7)])
def test_save_load_round_trip(tmp_path, desc_fn):
    d = desc_fn()
    a = FlipSim(d)
    a.Init()
    a.Step(3)
    path = tmp_path / "state.bin"
    print("save after 3 steps", flush=True)
    a.save_state(path)
    b = FlipSim(d)            # fresh handle, never initialised
    b.load_state(path)
    assert b.m_frame == a.m_frame == 3
    assert b.max_density == a.max_density
    for k, v in _state(a).items():
        if k != "frame":
            assert_bitexact(_state(b)[k], v, f"loaded {k}")
    a.Step(2)
    b.Step(2)
    sa, sb = _state(a), _state(b)
    assert sa["frame"] == sb["frame"] == 5
    for k in sa:
        if k != "frame":
            assert_bitexact(sb[k], sa[k], f"after two more steps: {k}")
    assert a.stats().iterations == b.stats().iterations
    a.close()
    b.close()


def test_load_refuses_truncated_and_foreign_files(tmp_path):
    d = scenes.dam_break(16)
    a = FlipSim(d)
    a.Init()
    a.Step(1)
    path = tmp_path / "state.bin"
    a.save_state(path)
    raw = path.read_bytes()
    cut = tmp_path / "cut.bin"
    cut.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(ArielError) as e:
        a.load_state(cut)
    assert e.value.code == _abi.ERR_IO
    other = FlipSim(scenes.dam_break(16, seed=3))
    with pytest.raises(ArielError) as e:
        other.load_state(path)
    assert e.value.code == _abi.ERR_IO
    with pytest.raises(ArielError) as e:
        a.load_state(tmp_path / "missing.bin")
    assert e.value.code == _abi.ERR_IO
    # the refused loads left the handle usable
    a.Step(1)
    assert a.stats().solve_status == 0
    a.close()
    other.close()


def test_set_particles_on_a_fresh_handle_calibrates_max_density():
    """Without ariel_init, ariel_set_particles runs FlipSim::Init's max-density
    calibration (flip.cpp:42-71) first, so densities are not divided by zero."""
    d = scenes.dam_break(16)
    ref = FlipSim(d)
    ref.Init()
    g = FlipSim(d)
    pos, vel, rho = ref.GetParticles()
    g.set_particles(pos, vel, rho)
    assert g.max_density == ref.max_density > 0
    for sim in (ref, g):
        sim.run_stage(_abi.STAGE_SORT)
        sim.run_stage(_abi.STAGE_DENSITY)
    assert_bitexact(g.GetParticles()[2], ref.GetParticles()[2], "densities")
    assert np.isfinite(g.GetParticles()[2]).all()
    ref.close()
    g.close()


def test_particle_types_and_scene():
    d = scenes.dam_break(16)
    g = FlipSim(d)
    g.Init()
    pos, vel, rho, typ = g.GetParticles(types=True)
    assert typ.shape == (len(pos),) and np.all(typ == _abi.FLUID)
    assert bytes(g.GetScene()) == bytes(d)
    g.close()

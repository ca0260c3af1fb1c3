"""The solver autotuner (ariel_step_autotune; no reference counterpart): one step
whose pressure solve is timed with each candidate preconditioner from the same
warm start; the fastest converged one is kept, and the step must equal a plain
step with it, bit for bit."""
import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from tests.helpers import assert_bitexact

pytestmark = pytest.mark.gpu
N = 64


def _state(sim):
    pos, vel, rho = sim.GetParticles()
    return pos, vel, rho, sim.get_field(_abi.FIELD_P), sim.get_flags()


def _same(a, b, what):
    for x, y, name in zip(a, b, ["positions", "velocities", "densities", "pressure", "classes"]):
        assert_bitexact(x, y, f"{what}: {name}")


def test_autotune_picks_the_fastest_and_steps_like_it():
    cands = [_abi.PRECOND_MIC0_TILED, _abi.PRECOND_MG, _abi.PRECOND_JACOBI]
    sim = FlipSim(scenes.pool_and_cube(N))
    sim.Init()
    sim.Step(2)
    trials, chosen = sim.StepAutotune(cands)
    for t in trials:
        print(f"autotune {N}^3: precond {t['precond']}: {t['iterations']} it, {t['solve_ms']:.2f} ms, "
              f"status {t['status']}", flush=True)
    assert [t["precond"] for t in trials] == cands
    assert all(t["status"] == 0 for t in trials)
    it = {t["precond"]: t["iterations"] for t in trials}
    assert it[_abi.PRECOND_MG] < it[_abi.PRECOND_MIC0_TILED] < it[_abi.PRECOND_JACOBI]
    conv = [i for i, t in enumerate(trials) if t["status"] == 0]
    assert chosen == min(conv, key=lambda i: trials[i]["solve_ms"])
    assert sim.GetScene().precond == cands[chosen]
    sim.close()


def test_single_candidate_autotune_is_a_plain_step():
    a = FlipSim(scenes.pool_and_cube(N))
    b = FlipSim(scenes.pool_and_cube(N))
    for s in (a, b):
        s.Init()
        s.Step(1)
    trials, chosen = a.StepAutotune([_abi.PRECOND_MIC0_TILED])
    b.Step(1)
    assert chosen == 0 and trials[0]["status"] == 0
    assert a.stats().iterations == b.stats().iterations
    _same(_state(a), _state(b), "autotuned step vs plain step")
    # and the next plain steps stay identical
    a.Step(1)
    b.Step(1)
    _same(_state(a), _state(b), "the step after")
    a.close()
    b.close()


def test_autotune_to_multigrid_then_steps_like_a_multigrid_handle():
    # a two-candidate tune whose pick is forced by giving only multigrid and
    # Jacobi (multigrid is the faster at 64^3): the handle then steps exactly
    # like one created with multigrid, from the same state
    a = FlipSim(scenes.pool_and_cube(N))
    a.Init()
    trials, chosen = a.StepAutotune([_abi.PRECOND_JACOBI, _abi.PRECOND_MG])
    assert trials[chosen]["precond"] == _abi.PRECOND_MG
    b = FlipSim(scenes.pool_and_cube(N, precond=_abi.PRECOND_MG))
    b.Init()
    b.Step(1)
    _same(_state(a), _state(b), "tuned to multigrid vs a multigrid handle")
    a.Step(2)
    b.Step(2)
    _same(_state(a), _state(b), "two more steps")
    a.close()
    b.close()

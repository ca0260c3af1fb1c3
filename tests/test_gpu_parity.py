"""Stage-wise parity of the HIP path against the CPU oracle (identical inputs).

Integer and byte outputs (sort order, per-cell counts, cell classes, resample
accept mask) and every stage whose arithmetic is order-preserving are required to
be bit-exact; the pressure solve (FP64 reductions in a different order) agrees to
1e-5 relative at the converged solution.
"""
import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from oracle import OracleSim
from tests.helpers import FIELDS, assert_bitexact, copy_state, rel_err

pytestmark = pytest.mark.gpu


def _pair(desc, warm_steps=0):
    o = OracleSim(desc, threads=16)
    g = FlipSim(desc)
    o.init()
    g.Init()
    for _ in range(warm_steps):
        o.step()
        g.Step(1)
    return o, g


@pytest.mark.parametrize("n,precond", [(32, _abi.PRECOND_MIC0_TILED)])
def test_init_matches(n, precond):
    d = scenes.dam_break(n, precond=precond)
    o, g = _pair(d)
    assert g.max_density == o.max_density
    po, vo, ro = o.get_particles()
    pg, vg, rg = g.GetParticles()
    assert_bitexact(pg, po, "seeded positions")
    assert_bitexact(vg, vo, "seeded velocities")
    assert_bitexact(g.get_flags(), o.get_flags(), "init flags")
    ig, og = g.get_sort()
    io, oo = o.get_sort()
    assert_bitexact(og, oo, "init offsets")
    assert_bitexact(ig, io, "init sort ids")


@pytest.mark.parametrize("desc_fn", [
    lambda: scenes.dam_break(32),
    lambda: scenes.dam_break(32, splat=_abi.SPLAT_TRILINEAR, precond=_abi.PRECOND_JACOBI),
    lambda: scenes.dam_break_sphere(48, precond=_abi.PRECOND_MIC0_BLOCK),
    lambda: scenes.pool_and_cube(32, refine_every=7),
    lambda: scenes.pool_and_cube(64),
    lambda: scenes.dam_break_sphere(128),
])
def test_stagewise_parity(desc_fn):
    d = desc_fn()
    o, g = _pair(d, warm_steps=2)
    copy_state(o, g)
    # SORT
    o.stage(_abi.STAGE_SORT)
    g.run_stage(_abi.STAGE_SORT)
    ig, og = g.get_sort()
    io, oo = o.get_sort()
    assert_bitexact(np.diff(og), np.diff(oo), "per-cell counts")
    assert_bitexact(ig, io, "sort ids")
    # DENSITY (+ forces)
    o.stage(_abi.STAGE_DENSITY)
    g.run_stage(_abi.STAGE_DENSITY)
    po, vo, ro = o.get_particles()
    pg, vg, rg = g.GetParticles()
    assert_bitexact(rg, ro, "density")
    assert_bitexact(vg, vo, "velocity after forces")
    # MARK
    o.stage(_abi.STAGE_MARK)
    g.run_stage(_abi.STAGE_MARK)
    assert_bitexact(g.get_flags(), o.get_flags(), "cell classes")
    assert_bitexact(g.get_field(_abi.FIELD_L), o.get_field(_abi.FIELD_L), "level set L")
    # P2G
    o.stage(_abi.STAGE_P2G)
    g.run_stage(_abi.STAGE_P2G)
    for f in (_abi.FIELD_UX, _abi.FIELD_UY, _abi.FIELD_UZ, _abi.FIELD_DUX, _abi.FIELD_DUY,
              _abi.FIELD_DUZ):
        assert_bitexact(g.get_field(f), o.get_field(f), f"p2g field {f}")
    # DIVERGENCE
    o.stage(_abi.STAGE_DIVERGENCE)
    g.run_stage(_abi.STAGE_DIVERGENCE)
    assert_bitexact(g.get_field(_abi.FIELD_D), o.get_field(_abi.FIELD_D), "divergence")
    # preconditioner itself on the current classes (bit-exact factor and apply)
    r = o.get_field(_abi.FIELD_D) * (o.get_flags() == 1)
    assert_bitexact(g.apply_preconditioner(r), o.apply_precond(r), "preconditioner apply")
    assert_bitexact(g.apply_laplacian(r), o.apply_A(r), "laplacian")
    # SOLVE
    o.stage(_abi.STAGE_SOLVE)
    g.run_stage(_abi.STAGE_SOLVE)
    so, sg = o.stats(), g.stats()
    assert sg.solve_status == 0 and so.solve_status == 0
    assert abs(sg.iterations - so.iterations) <= 1, (sg.iterations, so.iterations)
    assert rel_err(g.get_field(_abi.FIELD_P), o.get_field(_abi.FIELD_P)) < 1e-5
    assert sg.residual_norm <= d.tol * sg.rhs_norm * 1.0000001
    # continue from the oracle's pressure so later stages see identical inputs
    g.set_field(_abi.FIELD_P, o.get_field(_abi.FIELD_P))
    g.set_field(_abi.FIELD_D, o.get_field(_abi.FIELD_D))
    # GRADIENT
    o.stage(_abi.STAGE_GRADIENT)
    g.run_stage(_abi.STAGE_GRADIENT)
    for f in (_abi.FIELD_UX, _abi.FIELD_UY, _abi.FIELD_UZ):
        assert_bitexact(g.get_field(f), o.get_field(f), f"gradient field {f}")
    assert g.stats().max_divergence == pytest.approx(o.stats().max_divergence, rel=1e-6)
    # EXTRAPOLATE
    o.stage(_abi.STAGE_EXTRAPOLATE)
    g.run_stage(_abi.STAGE_EXTRAPOLATE)
    for f in FIELDS[:6]:
        assert_bitexact(g.get_field(f), o.get_field(f), f"extrapolate field {f}")
    # ADVECT
    o.stage(_abi.STAGE_ADVECT)
    g.run_stage(_abi.STAGE_ADVECT)
    po, vo, ro = o.get_particles()
    pg, vg, rg = g.GetParticles()
    assert_bitexact(pg, po, "advected positions")
    assert_bitexact(vg, vo, "PIC/FLIP velocities")
    # RESAMPLE
    o.stage(_abi.STAGE_RESAMPLE)
    g.run_stage(_abi.STAGE_RESAMPLE)
    ig, og = g.get_sort()
    io, oo = o.get_sort()
    assert_bitexact(og, oo, "resample offsets")
    assert_bitexact(ig, io, "resample sort ids")
    assert_bitexact(g.get_accept(), o.get_accept(), "resample accept mask")
    po, vo, ro = o.get_particles()
    pg, vg, rg = g.GetParticles()
    assert_bitexact(pg, po, "resampled positions")
    assert_bitexact(vg, vo, "resampled velocities")


def test_free_running_steps_track_oracle():
    """Five full steps from the same seed: flags and counts stay identical while the
    trajectories agree to FP32 noise (the pressure differs by reduction order)."""
    d = scenes.dam_break(32)
    o, g = _pair(d)
    for step in range(5):
        o.step()
        g.Step(1)
        so, sg = o.stats(), g.stats()
        assert sg.solve_status == 0
        mism = int((g.get_flags() != o.get_flags()).sum())
        assert mism <= 2, f"step {step + 1}: {mism} cell classes differ"
    po, vo, _ = o.get_particles()
    pg, vg, _ = g.GetParticles()
    assert np.abs(pg - po).max() < 1e-4
    assert rel_err(vg, vo) < 1e-3


def test_dam_break_sphere_128_first_steps():
    """The 128^3 dam break + sphere, free running for steps 1-5: cell classes and
    per-cell particle counts identical to the oracle's on every step, and the
    particle totals equal. Every stage up to the solve is bit-exact; the pressure
    agrees to its last bits only (FP64 reduction order), so a particle that sat
    within an ulp of a cell face after advection could land in another cell. For
    this seeded scene none does on steps 1-5 (the measured differences are zero),
    and the test holds it to that."""
    d = scenes.dam_break_sphere(128)
    o, g = _pair(d)
    for step in range(5):
        o.step()
        g.Step(1)
        assert g.stats().solve_status == 0 and o.stats().solve_status == 0
        assert g.num_particles == o.num_particles
        flag_mism = int((g.get_flags() != o.get_flags()).sum())
        count_mism = int((np.diff(g.get_sort()[1]) != np.diff(o.get_sort()[1])).sum())
        print(f"step {step + 1}: {flag_mism} cell classes, {count_mism} cell counts differ", flush=True)
        assert flag_mism == 0, f"step {step + 1}: {flag_mism} cell classes differ"
        assert count_mism == 0, f"step {step + 1}: {count_mism} per-cell counts differ"

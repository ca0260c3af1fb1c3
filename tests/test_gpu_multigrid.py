"""The multigrid preconditioner (ARIEL_PRECOND_MG; no reference counterpart):
one V(1,1) cycle with damped-Jacobi smoothing, trilinear transfers (restriction
the transpose of prolongation over 8) and symmetric red-black Gauss-Seidel on
the coarsest level, inside plain PCG. The cycle must be a symmetric positive
definite operator, its solves must reach the FP64 oracle's pressure to the
solve tolerance, and it must need far fewer iterations than MIC(0)."""
import numpy as np
import pytest

from ariel_amd import ArielError, FlipBatch, FlipSim, _abi, scenes
from oracle import OracleSim
from tests import precision

pytestmark = pytest.mark.gpu


def _mg(d):
    d.precond = _abi.PRECOND_MG
    return d


# 48: levels 48, 24, 12, 6; 72: 72, 36, 18, 9, 5 (an odd edge, coarsened by
# ceil); 16: one coarse level
@pytest.mark.parametrize("n", [48, 72, 16])
def test_vcycle_is_symmetric_and_positive(n):
    sim = FlipSim(_mg(scenes.dam_break_sphere(n)))
    sim.Init()
    sim.Step(2)
    fl = sim.get_flags().ravel() == _abi.FLUID
    rng = np.random.default_rng(n)
    rs = []
    for q in range(4):
        r = np.where(fl, rng.standard_normal(fl.size), 0.0)
        if q == 3:   # a smooth right-hand side
            g = np.indices((n, n, n)).reshape(3, -1) / n
            r = np.where(fl, np.sin(3.0 * g[0]) * np.cos(2.0 * g[1]) + g[2], 0.0)
        rs.append(r.astype(np.float32))
    zs = [sim.apply_preconditioner(r).astype(np.float64) for r in rs]
    print(f"V-cycle n={n}: {fl.sum()} FLUID cells", flush=True)
    for a in range(len(rs)):
        assert np.all(zs[a][~fl] == 0.0)
        assert np.all(np.isfinite(zs[a]))
        assert float(zs[a] @ rs[a]) > 0.0
        for b in range(a + 1, len(rs)):
            x, y = float(zs[a] @ rs[b]), float(rs[a] @ zs[b])
            scale = np.linalg.norm(zs[a]) * np.linalg.norm(rs[b])
            # symmetric in exact arithmetic (the float64 restatement: 1e-18); FP32
            # rounding at free-surface cells, whose diagonals reach ~10^3, leaves
            # up to ~2e-5 (measured)
            assert abs(x - y) <= 1e-4 * scale, (a, b, x, y)
    sim.close()


def test_mgpcg_pressure_against_fp64_oracle_and_iterations():
    n, steps = 64, 3
    g = FlipSim(_mg(scenes.pool_and_cube(n)))
    ref = FlipSim(scenes.pool_and_cube(n))
    g.Init()
    ref.Init()
    chk = OracleSim(scenes.pool_and_cube(n), threads=16)
    for step in range(steps):
        g.Step(1)
        ref.Step(1)
        st, sr = g.stats(), ref.stats()
        assert st.solve_status == 0
        e = precision.errors(chk, g.get_flags(), g.get_field(_abi.FIELD_L),
                             g.get_field(_abi.FIELD_D), g.get_field(_abi.FIELD_P))
        print(f"multigrid n={n} step {step + 1}: {st.iterations} it (MIC(0) {sr.iterations}); "
              f"pressure rel error {e['pressure_rel_error']:.2e}, velocity {e['velocity_rel_error']:.2e}",
              flush=True)
        assert e["pressure_rel_error"] < 1e-5
        assert e["velocity_rel_error"] < 1e-4
        assert st.iterations * 5 <= sr.iterations * 3   # measured: 23 vs 46, 17 vs 33
    g.close()
    ref.close()
    chk.close()


def test_batches_refuse_multigrid():
    with pytest.raises(ArielError):
        FlipBatch([_mg(scenes.dam_break(32)), _mg(scenes.dam_break(32))])


@pytest.mark.parametrize("n", [48, 64])
def test_vcycle_matches_the_numpy_restatement(n):
    from tests import mg_model
    sim = FlipSim(_mg(scenes.pool_and_cube(n)))
    sim.Init()
    sim.Step(1)
    flags = sim.get_flags().reshape(n, n, n)
    L = sim.get_field(_abi.FIELD_L).reshape(n, n, n).astype(np.float64)
    lv = mg_model.levels(flags, L)
    fl = flags == _abi.FLUID
    rng = np.random.default_rng(7)
    r = np.where(fl, rng.standard_normal(fl.shape), 0.0).astype(np.float32)
    z = sim.apply_preconditioner(r).reshape(fl.shape).astype(np.float64)
    zm = mg_model.vcycle(lv, r.astype(np.float64))
    err = np.linalg.norm(z - zm) / np.linalg.norm(zm)
    print(f"V-cycle vs restatement n={n}: levels {[x[0].shape[0] for x in lv]}, rel error {err:.2e}", flush=True)
    assert err < 1e-4
    sim.close()

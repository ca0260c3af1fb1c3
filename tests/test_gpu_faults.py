"""Fault injection through the C ABI on the GPU: the same cases as the oracle's
(empty fluid, the singular all-Neumann pocket, a NaN right-hand side, the
iteration cap) end with the same solve status; NaN or zero network weights
make the learned preconditioner fall back to MIC(0) and still converge."""
import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, learned, scenes
from ariel_amd.learned import net as lnet
from tests.test_oracle_solver import FLUID, _fault_solve, fault_cases

pytestmark = pytest.mark.gpu


def _gpu_solve(flags, b, n=16, max_iters=500, precond=None, blob=None):
    d = scenes.dam_break(n, max_iters=max_iters)
    if precond is not None:
        d.precond = precond
    s = FlipSim(d)
    if blob is not None:
        s.set_learned_weights(blob)
    s.Init()
    s.set_flags(flags)
    s.set_field(_abi.FIELD_L, np.where(flags == FLUID, -0.5, 0.5).astype(np.float32))
    x, st = s.solve_rhs(np.asarray(b, np.float32))
    s.close()
    return st, x


@pytest.mark.parametrize("case", range(5))
def test_fault_statuses_match_oracle(case):
    name, flags, b, max_iters, want = fault_cases()[case]
    print(f"fault case: {name}", flush=True)
    st, x = _gpu_solve(flags, b, max_iters=max_iters)
    so, _ = _fault_solve(flags, b, max_iters=max_iters)
    assert st.solve_status in want, (name, st.solve_status)
    if len(want) == 1:
        assert st.solve_status == so.solve_status
    if name == "empty fluid":
        assert st.iterations == 0 and np.all(x == 0.0)
    if name == "iteration cap 1":
        assert st.iterations == 1
    if st.solve_status != 3:
        assert np.all(np.isfinite(x))


@pytest.mark.parametrize("kind", ["nan", "zero"])
def test_learned_fallback_on_broken_weights(kind):
    n = 32
    flags = np.zeros((n, n, n), np.uint8)
    flags[1:-1, 1:14, 1:-1] = FLUID
    flags = flags.ravel()
    b = np.random.default_rng(3).standard_normal(n ** 3) * (flags == FLUID)
    ws = learned.unpack(lnet.PressureNet().blob())
    ws = [np.full_like(w, np.nan if kind == "nan" else 0.0) for w in ws]
    print(f"learned fallback, {kind} weights", flush=True)
    stl, xl = _gpu_solve(flags, b, n=n, precond=_abi.PRECOND_LEARNED, blob=learned.pack(ws))
    stm, xm = _gpu_solve(flags, b, n=n)
    assert stl.precond_fallback == 2 and stl.fallback_iter == 0
    assert stl.solve_status == 0 and np.all(np.isfinite(xl))
    assert np.linalg.norm(xl - xm) <= 1e-3 * np.linalg.norm(xm)

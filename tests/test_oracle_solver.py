"""Oracle pins for the pressure solve: the matrix-free operator against an
explicitly assembled sparse matrix, PCG against scipy's direct solve, and the
MIC(0)/IC(0) factorisations against their defining identities (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ariel_amd import _abi, scenes
from oracle import OracleSim

AIR, FLUID, SOLID = 0, 1, 2


def assemble(flags, L, n):
    """Sparse operator of ComputeAx (solver.inl:196-230) from its definition:
    for a FLUID cell c: N^2 * (d_c x_c - sum_{FLUID q} x_q), with
    d_c = 6 - #(SOLID or outside) - sum_{AIR q} L_q / min(1e-6, L_c)."""
    idx = np.arange(n ** 3).reshape(n, n, n)
    fl = flags.reshape(n, n, n)
    Lg = L.reshape(n, n, n).astype(np.float64)
    rows, cols, vals = [], [], []
    for i, j, k in zip(*np.nonzero(fl == FLUID)):
        c = idx[i, j, k]
        d = 6.0
        for di, dj, dk in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
            a, b, e = i + di, j + dj, k + dk
            if not (0 <= a < n and 0 <= b < n and 0 <= e < n) or fl[a, b, e] == SOLID:
                d -= 1.0
            elif fl[a, b, e] == AIR:
                d -= Lg[a, b, e] / min(1e-6, Lg[i, j, k])
            else:
                rows.append(c); cols.append(idx[a, b, e]); vals.append(-1.0)
        rows.append(c); cols.append(c); vals.append(d)
    A = sp.csr_matrix((np.array(vals) * n * n, (rows, cols)), shape=(n ** 3, n ** 3))
    return A


def stepped_oracle(n=16, steps=2, **kw):
    d = scenes.dam_break(n, **kw)
    o = OracleSim(d, threads=4)
    o.init()
    for _ in range(steps):
        o.step()
    return o, d


def test_operator_matches_assembled_matrix():
    o, _ = stepped_oracle()
    n = o.n
    flags, L = o.get_flags(), o.get_field(_abi.FIELD_L)
    assert (flags == AIR).any() and (flags == FLUID).any() and (flags == SOLID).any()
    A = assemble(flags, L, n)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n ** 3).astype(np.float32)
    y = o.apply_A(x)
    ref = A @ x.astype(np.float64)
    m = flags == FLUID
    assert np.abs(y[m] - ref[m]).max() <= 1e-5 * np.abs(ref[m]).max()
    assert np.all(y[~m] == 0.0)
    # FP64 evaluation of the same operator (true residual) agrees to double rounding
    y64 = o.apply_A_f64(x)
    assert np.abs(y64 - ref).max() <= 1e-11 * np.abs(ref).max()
    # constant vector: zero on fluid cells whose 6 neighbours are all fluid
    fl = flags.reshape(n, n, n)
    interior = np.zeros_like(fl, bool)
    interior[1:-1, 1:-1, 1:-1] = ((fl[1:-1, 1:-1, 1:-1] == FLUID) & (fl[2:, 1:-1, 1:-1] == FLUID) &
                                  (fl[:-2, 1:-1, 1:-1] == FLUID) & (fl[1:-1, 2:, 1:-1] == FLUID) &
                                  (fl[1:-1, :-2, 1:-1] == FLUID) & (fl[1:-1, 1:-1, 2:] == FLUID) &
                                  (fl[1:-1, 1:-1, :-2] == FLUID))
    ones = o.apply_A(np.ones(n ** 3, np.float32))
    assert np.all(ones[interior.ravel()] == 0.0)


@pytest.mark.parametrize("precond", [_abi.PRECOND_MIC0, _abi.PRECOND_MIC0_TILED,
                                     _abi.PRECOND_MIC0_BLOCK, _abi.PRECOND_JACOBI])
def test_pcg_matches_direct_solve(precond):
    o, d = stepped_oracle(precond=precond)
    n = o.n
    flags, L = o.get_flags(), o.get_field(_abi.FIELD_L)
    A = assemble(flags, L, n)
    m = np.flatnonzero(flags == FLUID)
    b = o.get_field(_abi.FIELD_D).astype(np.float64)[m]
    x = spla.spsolve(A[m][:, m].tocsc(), b)
    st = o.stats()
    assert st.solve_status == 0 and st.iterations > 0
    p = o.get_field(_abi.FIELD_P).astype(np.float64)
    assert np.abs(p[m] - x).max() <= 1e-4 * np.abs(x).max()
    assert st.residual_norm <= 1e-6 * st.rhs_norm * (1 + 1e-6)
    assert np.all(p[np.flatnonzero(flags != FLUID)] == 0.0)
    # FP64 oracle solve to 1e-12 matches the direct solve to double accuracy
    it, p64, rel = o.solve_fp64(1e-12)
    assert it > 0 and rel <= 1e-12
    assert np.abs(p64[m] - x).max() <= 1e-9 * np.abs(x).max()


def _mic_matrix(o, flags, pc_vals, order_lower):
    """M = (F + E) E^-1 (F + E)^T with E = pc^-2 and F the strictly lower part of
    the (unscaled) operator under the ordering."""
    n = o.n
    A = assemble(flags, o.get_field(_abi.FIELD_L), n) / (n * n)
    m = np.flatnonzero(flags == FLUID)
    Af = A[m][:, m].tocoo()
    low = np.array([order_lower(m[r], m[c]) for r, c in zip(Af.row, Af.col)])
    F = sp.csr_matrix((Af.data[low], (Af.row[low], Af.col[low])), shape=Af.shape)
    E = 1.0 / pc_vals[m].astype(np.float64) ** 2
    FE = F + sp.diags(E)
    M = FE @ sp.diags(1.0 / E) @ FE.T
    return A[m][:, m], M


def _natural_lower(n):
    return lambda a, b: b < a


def _tiled_lower(n, T=8):
    def tile(c):
        i, j, k = c // (n * n), (c // n) % n, c % n
        return (i // T, j // T, k // T)

    def lower(a, b):   # is b before a ?
        ta, tb = tile(a), tile(b)
        if ta == tb:
            return b < a
        return (sum(tb) & 1) == 0
    return lower


@pytest.mark.parametrize("kind", [_abi.PRECOND_MIC0, _abi.PRECOND_MIC0_TILED])
def test_ic0_identity_on_pattern(kind):
    """tau = 0 gives IC(0): (F+E)E^-1(F+E)^T equals A on A's sparsity pattern
    (incomplete Cholesky, any ordering)."""
    o, _ = stepped_oracle(n=16, precond=kind, mic_tau=0.0, mic_sigma=0.0)
    n = o.n
    flags = o.get_flags()
    o.apply_precond(np.ones(n ** 3, np.float32))   # builds the factor
    pc = o.get_precond()
    lower = _natural_lower(n) if kind == _abi.PRECOND_MIC0 else _tiled_lower(n)
    A, M = _mic_matrix(o, flags, pc, lower)
    Ac = A.tocoo()
    Md = M.tocsr()
    diff = np.abs(np.asarray(Md[Ac.row, Ac.col]).ravel() - Ac.data)
    assert diff.max() <= 1e-5 * np.abs(Ac.data).max()


@pytest.mark.parametrize("kind", [_abi.PRECOND_MIC0, _abi.PRECOND_MIC0_TILED])
def test_mic0_row_sums(kind):
    """tau = 1 (no safety) gives MIC(0): M 1 = A 1 (modified incomplete Cholesky)."""
    o, _ = stepped_oracle(n=16, precond=kind, mic_tau=1.0, mic_sigma=0.0)
    n = o.n
    flags = o.get_flags()
    o.apply_precond(np.ones(n ** 3, np.float32))
    pc = o.get_precond()
    m = np.flatnonzero(flags == FLUID)
    assert np.all(pc[m] > 0)
    lower = _natural_lower(n) if kind == _abi.PRECOND_MIC0 else _tiled_lower(n)
    A, M = _mic_matrix(o, flags, pc, lower)
    one = np.ones(A.shape[0])
    assert np.abs(M @ one - A @ one).max() <= 1e-4 * np.abs(A.diagonal()).max()


def test_mic0_apply_is_inverse_of_factor():
    o, _ = stepped_oracle(n=16, precond=_abi.PRECOND_MIC0_TILED)
    n = o.n
    flags = o.get_flags()
    rng = np.random.default_rng(5)
    r = (rng.standard_normal(n ** 3) * (flags == FLUID)).astype(np.float32)
    z = o.apply_precond(r)
    pc = o.get_precond()
    A, M = _mic_matrix(o, flags, pc, _tiled_lower(n))
    m = np.flatnonzero(flags == FLUID)
    assert np.abs(M @ z[m].astype(np.float64) - r[m]).max() <= 1e-4 * np.abs(r).max()
    assert np.all(z[flags != FLUID] == 0.0)


def test_reference_preconditioner_reduces_to_diagonal():
    """The reference's PRef returns a value only when it equals FLUID (1.0f)
    (solver.inl:72-78), so BuildPreconditioner/ApplyPreconditioner reduce to
    z = (r * pc) * pc with pc = 1/sqrt(diag)."""
    o, _ = stepped_oracle(n=16, precond=_abi.PRECOND_REFERENCE)
    n = o.n
    flags = o.get_flags()
    rng = np.random.default_rng(2)
    r = (rng.standard_normal(n ** 3) * (flags == FLUID)).astype(np.float32)
    z = o.apply_precond(r)
    pc = o.get_precond()
    A = assemble(flags, o.get_field(_abi.FIELD_L), n) / (n * n)
    dg = np.zeros(n ** 3)
    dg[flags == FLUID] = A.diagonal()[flags == FLUID]
    m = flags == FLUID
    assert np.allclose(pc[m], 1 / np.sqrt(dg[m]), rtol=1e-6)
    assert np.array_equal(z[m], (r[m] * pc[m]) * pc[m])


def test_faithful_reference_solver_terminates():
    d = scenes.dam_break(16, precond=_abi.PRECOND_REFERENCE)
    o = OracleSim(d, threads=4)
    o.init()
    o.step(mode=1)
    st = o.stats()
    assert st.solve_status == 0 and 0 < st.iterations < 16 ** 3


def test_faithful_reference_solver_matches_direct_solve():
    """SolveConjugateGradient verbatim (solver.inl:289-354): it stops at the first
    iteration whose FP32 r.r falls to 1e-2 N^3 (absolute). With a right-hand side
    large enough that this is a tight relative bound, its pressure must match
    scipy's direct solve of the assembled operator, and the true residual of the
    result must meet the same absolute bound up to FP32 drift."""
    o, d = stepped_oracle(n=16, precond=_abi.PRECOND_REFERENCE)
    n = o.n
    flags, L = o.get_flags(), o.get_field(_abi.FIELD_L)
    m = np.flatnonzero(flags == FLUID)
    rng = np.random.default_rng(8)
    b = np.zeros(n ** 3, np.float32)
    b[m] = (rng.standard_normal(m.size) * 1e4).astype(np.float32)
    o.set_field(_abi.FIELD_D, b)
    o.set_field(_abi.FIELD_P, np.zeros(n ** 3, np.float32))
    o.stage(_abi.STAGE_SOLVE, mode=1)
    st = o.stats()
    eps = 1e-2 * n ** 3
    assert st.solve_status == 0 and 1 < st.iterations < n ** 3
    assert st.residual_norm ** 2 <= eps
    A = assemble(flags, L, n)[m][:, m].tocsc()
    x = spla.spsolve(A, b[m].astype(np.float64))
    p = o.get_field(_abi.FIELD_P).astype(np.float64)
    r_true = b[m] - A @ p[m]
    assert (r_true ** 2).sum() <= 1.1 * eps
    bound = np.sqrt(1.1 * eps) / float(np.linalg.norm(b[m]))   # relative residual reached
    assert bound < 1e-4
    # error <= cond * residual; measured 1.4e-4 against a residual bound of 2.5e-5
    assert np.linalg.norm(p[m] - x) <= 20.0 * bound * np.linalg.norm(x)
    assert np.all(p[np.flatnonzero(flags != FLUID)] == 0.0)
    o.close()


def test_precision_measure_on_an_oracle_step():
    """The FP64 error measurement pinned on the oracle's own FP32 step: the FP32
    pressure is within the solver tolerance's reach of the FP64 solution, the
    error of the exact FP64 solution (rounded to FP32) is at FP32 round-off,
    and a planted 1% pressure error is reported as 1%."""
    from tests import precision
    d = scenes.pool_and_cube(32)
    o = OracleSim(d, threads=4)
    o.init()
    o.step()
    flags, L, D, P = (o.get_flags(), o.get_field(_abi.FIELD_L), o.get_field(_abi.FIELD_D),
                      o.get_field(_abi.FIELD_P))
    chk = OracleSim(d, threads=4)
    e = precision.errors(chk, flags, L, D, P)
    assert e["fp64_rel_residual"] <= 1e-10
    assert 0.0 < e["pressure_rel_error"] < 1e-4 and 0.0 < e["velocity_rel_error"] < 1e-3
    it, p64, _ = chk.solve_fp64(tol=1e-10)
    e64 = precision.errors(chk, flags, L, D, p64.astype(np.float32))
    assert e64["pressure_rel_error"] < 1e-6
    fl = flags == FLUID
    e1 = precision.errors(chk, flags, L, D, np.where(fl, p64 * 1.01, 0.0).astype(np.float32))
    assert e1["pressure_rel_error"] == pytest.approx(0.01, rel=1e-3)
    assert e1["velocity_rel_error"] == pytest.approx(0.01, rel=1e-2)


def _fault_solve(flags, D, n=16, max_iters=500):
    d = scenes.dam_break(n, max_iters=max_iters)
    o = OracleSim(d, threads=4)
    o.set_flags(flags)
    o.set_field(_abi.FIELD_L, np.where(flags == FLUID, -0.5, 0.5).astype(np.float32))
    o.set_field(_abi.FIELD_D, np.asarray(D, np.float32))
    o.set_field(_abi.FIELD_P, np.zeros(n ** 3, np.float32))
    o.stage(_abi.STAGE_SOLVE)
    return o.stats(), o.get_field(_abi.FIELD_P)


def fault_cases(n=16, seed=0):
    """(name, flags, rhs, max_iters, expected status or set of statuses)."""
    rng = np.random.default_rng(seed)
    pocket = np.full((n, n, n), SOLID, np.uint8)
    pocket[4:12, 4:12, 4:12] = FLUID
    pocket = pocket.ravel()
    b = rng.standard_normal(n ** 3) * (pocket == FLUID)
    consistent = b.copy()
    consistent[pocket == FLUID] -= consistent[pocket == FLUID].mean()
    dam = np.zeros((n, n, n), np.uint8)
    dam[1:-1, 1:8, 1:-1] = FLUID
    dam = dam.ravel()
    bn = rng.standard_normal(n ** 3) * (dam == FLUID)
    bn[np.flatnonzero(dam == FLUID)[5]] = np.nan
    return [
        ("empty fluid", np.zeros(n ** 3, np.uint8), rng.standard_normal(n ** 3), 500, {0}),
        ("closed pocket, zero-mean rhs", pocket, consistent, 500, {0}),
        ("closed pocket, inconsistent rhs", pocket, b, 500, {1, 2}),
        ("NaN in the rhs", dam, bn, 500, {3}),
        ("iteration cap 1", dam, rng.standard_normal(n ** 3) * (dam == FLUID), 1, {1}),
    ]


@pytest.mark.parametrize("case", range(5))
def test_fault_injection(case):
    """Empty fluid, the singular all-Neumann pocket (consistent and not), a NaN
    right-hand side and the iteration cap: every solve terminates with the
    matching status (converged / max-iters / breakdown / non-finite)."""
    name, flags, b, max_iters, want = fault_cases()[case]
    st, P = _fault_solve(flags, b, max_iters=max_iters)
    assert st.solve_status in want, (name, st.solve_status)
    if name == "empty fluid":
        assert st.iterations == 0 and np.all(P == 0.0)
    if name == "iteration cap 1":
        assert st.iterations == 1
    if st.solve_status != 3:
        assert np.all(np.isfinite(P))

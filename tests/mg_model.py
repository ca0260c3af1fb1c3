"""A numpy restatement of the multigrid V(1,1) cycle of ARIEL_PRECOND_MG (test
infrastructure): dense grids at every level, the same cell classes, operators,
piecewise-constant transfers, smoothing and coarsest-level sweeps as the
kernels, in float64."""
import numpy as np

AIR, FLUID, SOLID = 0, 1, 2
OMEGA = 6.0 / 7.0
COARSEST_SWEEPS = 6
_NBRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def _shift(a, d, fill):
    """out[c] = a[c + d], `fill` outside the grid."""
    out = np.full_like(a, fill)
    src, dst = [slice(None)] * 3, [slice(None)] * 3
    for ax, s in enumerate(d):
        if s == 1:
            src[ax], dst[ax] = slice(1, None), slice(0, -1)
        elif s == -1:
            src[ax], dst[ax] = slice(0, -1), slice(1, None)
    out[tuple(dst)] = a[tuple(src)]
    return out


def fine_diagonal(flags, L):
    """The solve's diagonal: 6, less 1 per SOLID or out-of-range neighbour, plus
    L_q / -min(1e-6, L_c) per AIR neighbour (the ghost-fluid term)."""
    dg = np.full(flags.shape, 6.0)
    for d in _NBRS:
        tq, lq = _shift(flags, d, SOLID), _shift(L, d, 0.0)
        dg -= np.where(tq == SOLID, 1.0, 0.0)
        dg -= np.where(tq == AIR, lq / np.minimum(1e-6, L), 0.0)
    return np.where(flags == FLUID, dg, 6.0)


def apply_A(t, dg, x, h2):
    y = dg * x
    for d in _NBRS:
        y = y - np.where(_shift(t, d, SOLID) == FLUID, _shift(x, d, 0.0), 0.0)
    return np.where(t == FLUID, y / h2, 0.0)


def coarse_classes(t):
    M = t.shape[0]
    Mc = (M + 1) // 2
    tp = np.full((2 * Mc,) * 3, -1, np.int16)
    tp[:M, :M, :M] = t
    ch = tp.reshape(Mc, 2, Mc, 2, Mc, 2).transpose(0, 2, 4, 1, 3, 5).reshape(Mc, Mc, Mc, 8)
    return np.where((ch == FLUID).any(-1), FLUID, np.where((ch == AIR).any(-1), AIR, SOLID)).astype(np.int16)


def coarse_diagonal(t):
    d = np.full(t.shape, 6.0)
    for dd in _NBRS:
        d -= _shift(t, dd, SOLID) == SOLID
    return np.where(t == FLUID, d, 1.0)


def _p1(Mf, Mc):
    """1D prolongation: fine x takes its parent's value x >> 1."""
    P = np.zeros((Mf, Mc))
    for x in range(Mf):
        P[x, x >> 1] = 1.0
    return P


def prolong(e, Mf):
    P = _p1(Mf, e.shape[0])
    x = np.tensordot(P, e, axes=(1, 0))
    x = np.tensordot(P, x, axes=(1, 1)).transpose(1, 0, 2)
    return np.tensordot(x, P, axes=(2, 1))


def restrict(r, Mc):
    P = _p1(r.shape[0], Mc)
    x = np.tensordot(P, r, axes=(0, 0))
    x = np.tensordot(P, x, axes=(0, 1)).transpose(1, 0, 2)
    return np.tensordot(x, P, axes=(2, 0)) / 8.0


def levels(flags, L):
    N = flags.shape[0]
    t = flags.astype(np.int16)
    out = [(t, fine_diagonal(flags, L), 1.0 / (N * N))]
    lv = 0
    while t.shape[0] > 8:
        t = coarse_classes(t)
        lv += 1
        out.append((t, coarse_diagonal(t), 4.0 ** lv / (N * N)))
    return out


def omega(dg):
    """Jacobi weight per cell: 1 where the ghost-fluid terms raise the diagonal
    above 6 (there a full local solve keeps the coarse correction out of the
    stiff cell), OMEGA elsewhere."""
    return np.where(dg > 6.0, 1.0, OMEGA)


def vcycle(lv, r, l=0):
    t, dg, h2 = lv[l]
    fl = t == FLUID
    if l == len(lv) - 1:
        z = np.zeros_like(r)
        col = np.indices(t.shape).sum(0) & 1

        def half(c):
            s = np.zeros_like(z)
            for d in _NBRS:
                s += np.where(_shift(t, d, SOLID) == FLUID, _shift(z, d, 0.0), 0.0)
            return np.where(fl & (col == c), (h2 * r + s) / dg, z)
        z = half(0)
        for _ in range(COARSEST_SWEEPS):
            z = half(1)
            z = half(0)
        return np.where(fl, z, 0.0)
    om = omega(dg)
    z1 = np.where(fl, om * h2 * r / dg, 0.0)
    res = np.where(fl, r - apply_A(t, dg, z1, h2), 0.0)
    tc = lv[l + 1][0]
    rc = np.where(tc == FLUID, restrict(res, tc.shape[0]), 0.0)
    zt = np.where(fl, z1 + prolong(vcycle(lv, rc, l + 1), t.shape[0]), 0.0)
    zo = zt + om * h2 / dg * (r - apply_A(t, dg, zt, h2))
    return np.where(fl, zo, 0.0)

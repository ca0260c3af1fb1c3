"""Particle level set of the frame export — TEST INFRASTRUCTURE ONLY (the checker
of ariel_export_levelset; the product never imports it).

Restates LevelSet(particles, maxdimension) (src/grid/levelset.cpp:101-111) with
the sphere rasterisation (rasterizeSpheres, levelset.cpp:108) instead of the
velocity trails of line 109, whose OpenVDB implementation (ParticlesToLevelSet)
is not in this image: every particle is a sphere of radius 0.5
(ParticleList::getPosRad, levelset.hpp:44-53) centred at p * maxd
(levelset.hpp:45-47) on a grid of voxel size 1; the level set is the minimum
over spheres of the distance to the sphere, limited to the narrow-band value 3
(OpenVDB's default half width of 3 voxels times the voxel size).

Arithmetic in float32 in the library's order: x = p * maxd, dx = i - x,
d = sqrt((dx*dx + dy*dy) + dz*dz), phi = min(min_p d - 0.5, 3) with the
minimum started at 3.5.
"""
import numpy as np

BAND = np.float32(3.0)
RADIUS = np.float32(0.5)


def particle_levelset(pos, maxd, N):
    """pos: (n, 3) positions in [0, 1]; returns phi on the (N+1)^3 nodes, [i, j, k]."""
    M = N + 1
    g = np.asarray(pos, np.float32) * np.float32(maxd)
    best = np.full((M, M, M), BAND + RADIUS, np.float32)
    # nodes within 3.5 of x on an axis: floor(x) - 3 .. floor(x) + 4
    offs = np.arange(-3, 5)
    base = np.floor(g).astype(np.int64)
    for a in offs:
        for b in offs:
            ii = base[:, 0] + a
            jj = base[:, 1] + b
            for c in offs:
                kk = base[:, 2] + c
                ok = (ii >= 0) & (ii < M) & (jj >= 0) & (jj < M) & (kk >= 0) & (kk < M)
                if not ok.any():
                    continue
                i, j, k = ii[ok], jj[ok], kk[ok]
                dx = i.astype(np.float32) - g[ok, 0]
                dy = j.astype(np.float32) - g[ok, 1]
                dz = k.astype(np.float32) - g[ok, 2]
                d = np.sqrt((dx * dx + dy * dy) + dz * dz)
                np.minimum.at(best, (i, j, k), d)
    return np.minimum(best - RADIUS, BAND)


def brute_force_levelset(pos, maxd, N):
    """All nodes against all particles (tiny inputs only): the same value."""
    M = N + 1
    g = np.asarray(pos, np.float32) * np.float32(maxd)
    idx = np.arange(M, dtype=np.float32)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    best = np.full((M, M, M), BAND + RADIUS, np.float32)
    for p in g:
        dx, dy, dz = I - p[0], J - p[1], K - p[2]
        best = np.minimum(best, np.sqrt((dx * dx + dy * dy) + dz * dz))
    return np.minimum(best - RADIUS, BAND)


# ---------------------------------------------------------------------------
# Velocity trails (rasterizeTrails, levelset.cpp:109), restating OpenVDB's
# ParticlesToLevelSet (a third-party library, unvendored in the reference and
# absent here; version unpinned: Nuparu bundles it, README.md:6), whose
# published rule is: per particle, radius R0 (0.5, getPosRadVel,
# levelset.hpp:55-67), position P0 = p * maxd (float32 product, widened),
# velocity v as given; spheres at P with radius R starting at (P0, R0); after
# each, P += 0.5 R n with n = -v / |v|, d = |P - P0|, R = R0 - (R0 - Rmin) d / |v|,
# while d <= |v|; Rmin = 0.01f (raster.setRmin, levelset.cpp:104). A sphere
# writes, at every node c with |c - P|^2 < (R + 3)^2, float(|c - P| - R) when
# smaller, and -3 where |c - P|^2 <= max(0, R - 3)^2. Arithmetic in float64
# (OpenVDB's Real), in this order.
TRAIL_RMIN = np.float64(np.float32(0.01))


def trail_levelset(pos, vel, maxd, N, max_spheres=1 << 20):
    """pos, vel: (n, 3) float32; returns phi on the (N+1)^3 nodes, [i, j, k]."""
    M = N + 1
    band = np.float64(BAND)
    r0 = np.float64(RADIUS)
    phi = np.full((M, M, M), BAND, np.float32)
    p0 = (np.asarray(pos, np.float32) * np.float32(maxd)).astype(np.float64)
    v = np.asarray(vel, np.float32).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        speed = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
        inv = 1.0 / speed
        nrm = -v * inv[:, None]
    P = p0.copy()
    R = np.full(len(p0), r0)
    d = np.zeros(len(p0))
    offs = np.arange(-4, 5)
    live = np.ones(len(p0), bool) & (r0 >= TRAIL_RMIN)
    steps = 0
    while live.any() and steps < max_spheres:
        idx = np.nonzero(live)[0]
        Pl, Rl = P[idx], R[idx]
        mx = Rl + band
        max2 = mx * mx
        inner = np.maximum(Rl - band, 0.0)
        min2 = inner * inner
        base = np.floor(Pl).astype(np.int64)
        for a in offs:
            x = base[:, 0] + a
            x2 = (x - Pl[:, 0]) * (x - Pl[:, 0])
            for b in offs:
                y = base[:, 1] + b
                x2y2 = x2 + (y - Pl[:, 1]) * (y - Pl[:, 1])
                for c in offs:
                    z = base[:, 2] + c
                    d2 = x2y2 + (z - Pl[:, 2]) * (z - Pl[:, 2])
                    ok = (d2 < max2) & (x >= 0) & (x < M) & (y >= 0) & (y < M) & (z >= 0) & (z < M)
                    if not ok.any():
                        continue
                    val = np.where(d2[ok] <= min2[ok], -band, np.sqrt(d2[ok]) - Rl[ok]).astype(np.float32)
                    np.minimum.at(phi, (x[ok], y[ok], z[ok]), val)
        with np.errstate(invalid="ignore"):
            P[idx] = Pl + (0.5 * Rl)[:, None] * nrm[idx]
            e = P[idx] - p0[idx]
            d[idx] = np.sqrt(e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1] + e[:, 2] * e[:, 2])
            R[idx] = r0 - (r0 - TRAIL_RMIN) * d[idx] * inv[idx]
            live[idx] = d[idx] <= speed[idx]
        steps += 1
    return phi

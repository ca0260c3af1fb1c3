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

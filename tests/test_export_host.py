"""Frame export on the CPU: the level-set oracle against closed forms and brute
force, and the library's mesher (ariel_mesh_sdf, host only) against what the
geometry fixes: closed oriented surfaces with the Euler characteristic of the
shape, vertices on the surface, normals out of the negative side."""
import os
import tempfile

import numpy as np
import pytest

from ariel_amd import ArielError, mesh_sdf
from oracle.levelset import BAND, RADIUS, brute_force_levelset, particle_levelset


def read_obj(path):
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                verts.append([float(x) for x in t[1:4]])
            elif t[0] == "f":
                faces.append([int(x) - 1 for x in t[1:]])
    return np.array(verts, np.float64).reshape(-1, 3), faces


def mesh(phi):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "m.obj")
        nv, nq = mesh_sdf(phi, path)
        v, f = read_obj(path)
    assert len(v) == nv and len(f) == nq
    return v, f


def edges_of(faces):
    directed = {}
    for f in faces:
        for a, b in zip(f, f[1:] + f[:1]):
            directed[(a, b)] = directed.get((a, b), 0) + 1
    return directed


def check_closed_oriented(faces):
    """Every edge is used once in each direction: a closed, consistently oriented surface."""
    d = edges_of(faces)
    assert all(c == 1 for c in d.values()), "an edge is used twice in one direction"
    assert all((b, a) in d for (a, b) in d), "an edge has no opposite"
    return len(d) // 2


def euler(v, faces):
    return len(v) - check_closed_oriented(faces) + len(faces)


def grid(M):
    i = np.arange(M, dtype=np.float64)
    return np.meshgrid(i, i, i, indexing="ij")


def sphere_phi(M, c, r):
    I, J, K = grid(M)
    return (np.sqrt((I - c[0]) ** 2 + (J - c[1]) ** 2 + (K - c[2]) ** 2) - r).astype(np.float32)


def test_sphere_mesh_is_a_closed_outward_sphere():
    c, r = (16.3, 15.8, 16.1), 9.3
    v, f = mesh(sphere_phi(33, c, r))
    assert euler(v, f) == 2
    dist = np.linalg.norm(v - np.array(c), axis=1)
    assert np.abs(dist - r).max() < 0.1           # on the surface to a tenth of a voxel
    for q in f[::7]:
        p = v[q]
        n = np.cross(p[1] - p[0], p[3] - p[0]) + np.cross(p[3] - p[2], p[1] - p[2])
        assert np.dot(n, p.mean(axis=0) - np.array(c)) > 0   # faces point out of phi < 0


def test_torus_has_euler_characteristic_zero_and_two_spheres_four():
    I, J, K = grid(41)
    R, rr = 11.0, 4.2
    q = np.sqrt((I - 20.2) ** 2 + (J - 20.1) ** 2) - R
    torus = (np.sqrt(q ** 2 + (K - 20.3) ** 2) - rr).astype(np.float32)
    v, f = mesh(torus)
    assert euler(v, f) == 0
    two = np.minimum(sphere_phi(41, (10.2, 20.1, 20.3), 6.1), sphere_phi(41, (30.4, 20.2, 19.7), 5.3))
    v, f = mesh(two)
    assert euler(v, f) == 4


def test_vertex_is_the_mean_of_the_edge_crossings():
    # one negative node: each of the 8 voxels around it holds a vertex, and its
    # 6 edges each give a quad (a closed cube-like surface: V 8, E 12, F 6)
    phi = np.ones((4, 4, 4), np.float32)
    phi[1, 1, 1] = -0.5
    v, f = mesh(phi)
    assert (len(v), len(f), euler(v, f)) == (8, 6, 2)
    # in the voxel whose lowest corner is that node, the 3 edges leaving it are
    # crossed at t = 0.5 / (0.5 + 1) = 1/3; their mean is 1/9 along each axis
    assert np.any(np.all(np.abs(v - (1.0 + 1.0 / 9.0)) < 1e-6, axis=1))


def test_empty_and_full_level_sets_give_no_faces():
    v, f = mesh(np.ones((5, 5, 5), np.float32))
    assert len(v) == 0 and f == []
    v, f = mesh(-np.ones((5, 5, 5), np.float32))
    assert len(v) == 0 and f == []


def test_mesher_rejects_bad_arguments():
    with pytest.raises(ArielError) as e:
        mesh_sdf(np.zeros((1, 1, 1), np.float32), "/tmp/never.obj")
    assert e.value.code == -1
    with pytest.raises(ArielError) as e:
        mesh_sdf(np.ones((3, 3, 3), np.float32), "/nonexistent-dir/m.obj")
    assert e.value.code == -7


def test_levelset_oracle_single_particle_is_the_sphere_distance():
    p = np.array([[0.4, 0.55, 0.31]], np.float32)
    N = 16
    phi = particle_levelset(p, N, N)
    x = p[0] * np.float32(N)
    I, J, K = grid(N + 1)
    exact = np.minimum(np.sqrt((I - x[0]) ** 2 + (J - x[1]) ** 2 + (K - x[2]) ** 2) - 0.5, 3.0)
    assert np.abs(phi - exact).max() < 1e-5
    assert phi.max() == BAND and phi.min() > -RADIUS
    assert np.all(phi[exact >= 3.0 + 1e-5] == BAND)   # the band value, exactly, far away


def test_levelset_oracle_matches_brute_force_bit_for_bit():
    rng = np.random.default_rng(3)
    N = 12
    p = rng.uniform(0.05, 0.95, size=(60, 3)).astype(np.float32)
    p[:5] = np.float32(0.5)                       # coincident particles
    p[5] = [0.0, 0.0, 0.0]
    p[6] = [1.0, 1.0, 1.0]                        # on the domain's corners
    a = particle_levelset(p, N, N)
    b = brute_force_levelset(p, N, N)
    assert np.array_equal(a, b)


# ------------------------------------------------------------------ files
# The OpenVDB and Partio writers on caller data (host only): written, then read
# back by the package's own readers (ariel_amd.fileio), which parse the format's
# layout independently of the writer.
def _band_volume(M, seed):
    rng = np.random.default_rng(seed)
    I, J, K = grid(M)
    c = rng.uniform(0.3 * M, 0.7 * M, 3)
    phi = np.sqrt((I - c[0]) ** 2 + (J - c[1]) ** 2 + (K - c[2]) ** 2) - 0.3 * M
    phi = np.minimum(phi + rng.normal(0, 0.05, phi.shape), BAND).astype(np.float32)
    return phi


@pytest.mark.parametrize("M,seed", [(9, 0), (33, 1), (47, 2)])
def test_vdb_round_trip(tmp_path, M, seed):
    from ariel_amd import read_vdb, write_vdb
    phi = _band_volume(M, seed)
    phi[phi < -BAND] = -BAND          # inactive nodes off the background are kept too
    path = tmp_path / "a.vdb"
    write_vdb(phi, path, float(BAND))
    back, bg, active, meta = read_vdb(path, M)
    assert bg == BAND
    assert np.array_equal(back, phi)
    assert np.array_equal(active, np.abs(phi) < BAND)
    assert meta["class"] == "level set" and meta["name"] == "surface" and meta["version"] == 222
    assert meta["file_voxel_count"] == int(active.sum())
    lo, hi = np.argwhere(active).min(axis=0), np.argwhere(active).max(axis=0)
    assert tuple(lo) == meta["file_bbox_min"] and tuple(hi) == meta["file_bbox_max"]
    # leaves: the 8^3 blocks holding a node off the background
    blocks = set(map(tuple, (np.argwhere(phi != BAND) // 8).tolist()))
    assert meta["leaves"] == len(blocks)
    raw = open(path, "rb").read()
    assert raw[:8] == (0x56444220).to_bytes(8, "little")


def test_vdb_of_an_empty_band(tmp_path):
    from ariel_amd import read_vdb, write_vdb
    phi = np.full((17, 17, 17), BAND, np.float32)
    write_vdb(phi, tmp_path / "e.vdb", float(BAND))
    back, bg, active, meta = read_vdb(tmp_path / "e.vdb", 17)
    assert meta["leaves"] == 0 and not active.any() and np.all(back == BAND)


@pytest.mark.parametrize("ext", [".geo", ".bgeo", ".geo.gz", ".bgeo.gz"])
def test_partio_round_trip(tmp_path, ext):
    from ariel_amd import read_partio, write_particles
    rng = np.random.default_rng(3)
    pos = (rng.random((777, 3)) * 64).astype(np.float32)
    vel = rng.standard_normal((777, 3)).astype(np.float32)
    path = tmp_path / ("p" + ext)
    write_particles(path, pos, vel)
    p, v, i = read_partio(path)
    assert np.array_equal(p, pos) and np.array_equal(v, vel) and np.array_equal(i, np.arange(777))
    if ext.endswith(".gz"):
        assert open(path, "rb").read(2) == b"\x1f\x8b"


def test_partio_refuses_other_formats(tmp_path):
    from ariel_amd import write_particles
    with pytest.raises(ArielError) as e:
        write_particles(tmp_path / "p.pdb", np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32))
    assert e.value.code == -6


# ------------------------------------------------------------------ trails
# Pins of the oracle's velocity-trail rasterisation (OpenVDB's rule restated;
# the library is not in this image): what the geometry of a trail fixes.
def test_trails_at_rest_are_the_spheres():
    from oracle.levelset import trail_levelset
    rng = np.random.default_rng(4)
    pos = (rng.random((60, 3)) * 0.6 + 0.2).astype(np.float32)
    a = trail_levelset(pos, np.zeros_like(pos), 16.0, 16)
    b = particle_levelset(pos, 16.0, 16)
    # float64 against float32 distances: within one float ulp of |phi| <= 3.5
    assert np.abs(a - b).max() <= 2.5e-7


def test_trail_of_one_particle_runs_back_along_its_velocity():
    from oracle.levelset import trail_levelset
    p = np.array([[0.5, 0.5, 0.5]], np.float32)   # node (12, 12, 12) of a 24^3 grid
    v = np.array([[3.0, 0.0, 0.0]], np.float32)   # moving +x at 3 voxels per unit time
    phi = trail_levelset(p, v, 24.0, 24)
    c = 12
    # the head sits exactly on a node: OpenVDB's inner rule (|c - P|^2 <=
    # max(0, R - 3)^2) marks such a node -3, deep inside
    assert phi[c, c, c] == -BAND
    # behind the head the radius falls linearly to Rmin over |v| = 3 voxels
    for t in (1, 2):
        r = 0.5 - (0.5 - 0.01) * t / 3.0
        assert -r - 0.02 <= phi[c - t, c, c] <= -r + 0.13, (t, phi[c - t, c, c])
    assert phi[c - 3, c, c] <= 0.02                  # the tail reaches 3 voxels back
    assert phi[c - 4, c, c] > 0.9                    # and no further
    assert phi[c + 1, c, c] == np.float32(0.5)       # ahead: the head sphere only
    # symmetric about the trail's axis
    assert phi[c - 1, c + 1, c] == phi[c - 1, c - 1, c] == phi[c - 1, c, c + 1] == phi[c - 1, c, c - 1]


def test_trails_are_order_free_and_contain_the_heads():
    from oracle.levelset import trail_levelset
    rng = np.random.default_rng(5)
    pos = (rng.random((80, 3)) * 0.6 + 0.2).astype(np.float32)
    vel = (rng.standard_normal((80, 3)) * 1.5).astype(np.float32)
    a = trail_levelset(pos, vel, 20.0, 20)
    perm = rng.permutation(80)
    assert np.array_equal(a, trail_levelset(pos[perm], vel[perm], 20.0, 20))
    assert np.all(a <= particle_levelset(pos, 20.0, 20) + 2.5e-7)
    assert (a < particle_levelset(pos, 20.0, 20) - 0.1).any()

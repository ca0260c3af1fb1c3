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
